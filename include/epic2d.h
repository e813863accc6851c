/* epic2d.h — C-ABI of the MI355X particle loop for ePIC2D_mcc.
 *
 * The reference (koushik9988/ePIC2D_mcc) has no FFI: its per-timestep loop is
 * a set of C++ member functions called from main.cpp:366-455.  Each entry
 * point below replaces one of them (cited per function); the host C++ classes
 * in epic2d_mcc_amd/host/ keep the reference's class and method names and
 * forward here, so main's loop is unchanged.
 *
 * Conventions
 *   - Every function returns an int status: EPIC_OK (0) or a negative code.
 *     No C++ exception crosses this boundary; epic_last_error() gives text.
 *   - Node arrays cross the boundary as contiguous doubles indexed
 *     (i, j) -> i*ny + j, the layout of Matrix<double>(nx, ny) and flatten()
 *     (linalg.h:211-219, function.h:154-169).
 *   - Particles cross as structure-of-arrays doubles; the library stores them
 *     on the device in FP64 or FP32 (epic_domain_desc.precision).
 *   - One context = one HIP stream, used from one host thread.  Separate
 *     contexts may run concurrently.
 */
#ifndef EPIC2D_H_
#define EPIC2D_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPIC_OK 0
#define EPIC_ERR_ARG (-1)      /* bad argument or unsupported option            */
#define EPIC_ERR_HIP (-2)      /* a HIP runtime call failed                     */
#define EPIC_ERR_CAPACITY (-3) /* particle capacity or a fixed limit exceeded  */
#define EPIC_ERR_STATE (-4)    /* call not valid in the current state          */
#define EPIC_ERR_NOCONV (-5)   /* solver hit max_iteration (result returned)   */
#define EPIC_ERR_DEVICE (-6)   /* device error counters non-zero (see below)   */

/* boundary conditions: Domain::bc strings "pbc" | "open" | "reflective" */
#define EPIC_BC_PBC 0
#define EPIC_BC_OPEN 1
#define EPIC_BC_REFLECTIVE 2
/* shape functions: Domain::shapefunction "CIC" | "NGP" */
#define EPIC_SHAPE_CIC 0
#define EPIC_SHAPE_NGP 1
/* potential solvers (Domain::SolverType)                                    */
#define EPIC_SOLVER_PCG 0       /* "cg"|"iter"|"iterative": Jacobi PCG with the
                                   reference recurrence (solvers.h:72-100)     */
#define EPIC_SOLVER_GS 1        /* "gs": SOR w=1.4 (field.cpp:70-139), run as
                                   red-black SOR                               */
#define EPIC_SOLVER_PCG_MIXED 2 /* FP32 Jacobi-PCG inner solves, FP64 true
                                   residual + iterative refinement             */
#define EPIC_SOLVER_SPECTRAL 3  /* exact sine-transform solve of the same
                                   operator in FP64 (the transforms are FP64
                                   matrix products on the matrix cores), plus
                                   FP64 residual refinement                    */
#define EPIC_SOLVER_GS_BOLTZMANN 4 /* "gshybrid": SOR on the Boltzmann-electron
                                   problem lap phi = -(rho - exp(phi)) L^2/LDe^2
                                   (gsfluidsolver, field.cpp:7-68), red-black,
                                   with its boundary rho values               */
#define EPIC_SOLVER_NEWTON_CG 5 /* "hybrid": nrpcgsolver (field.cpp:141-210), ten
                                   Newton steps each solved by plain CG on
                                   A - diag(exp(phi)), as written (its
                                   right-hand side accumulates over the steps);
                                   one workgroup                              */
/* inner solve of EPIC_SOLVER_PCG_MIXED */
#define EPIC_PRECOND_JACOBI 0   /* FP32 Jacobi-preconditioned CG               */
#define EPIC_PRECOND_SPECTRAL 1 /* FP32 sine-transform solve (one application
                                   per refinement step)                        */
/* Poisson operators */
#define EPIC_OP_REF_GHOST 0      /* pcgsolver's A: every node unknown, zero ghost
                                    ring, zero RHS on boundary (field.cpp:221-251) */
#define EPIC_OP_DIRICHLET_WALL 1 /* gssolver's problem: boundary phi = 0, interior
                                    unknowns (field.cpp:89-98)                    */
/* particle storage precision */
#define EPIC_FP64 0
#define EPIC_FP32 1

/* device error counters (epic_error_counters) */
#define EPIC_NERR 8
#define EPIC_ERRC_GATHER_OOB 0  /* gather outside the node grid (reference throws) */
#define EPIC_ERRC_DEPOSIT_OOB 1 /* scatter outside the node grid (reference throws) */
#define EPIC_ERRC_FIXED_OVF 2   /* fixed-point deposit overflow                     */
#define EPIC_ERRC_CAPACITY 3    /* births/injections dropped for lack of capacity   */
#define EPIC_ERRC_NOCONV 4      /* a Poisson solve stopped above its tolerance (PCG
                                   or SOR at max_iteration, the mixed and sine-
                                   transform solves at their correction limit)     */
#define EPIC_ERRC_NAN 5         /* a push met a non-finite new position (a non-finite
                                   velocity or field); the particle is dropped (the
                                   reference's next bounds-checked gather throws)  */
#define EPIC_ERRC_MCC_BOUND 6   /* a collision candidate's probability exceeded the
                                   null-collision bound 1 - exp(-nu_max dt): a speed
                                   beyond what nu_max covers (4x the table's top
                                   speed); it collided with the bound instead      */

typedef struct epic_ctx epic_ctx;

/* The Domain members the hot path reads (domain.h:26-201), already
 * normalised as main.cpp:151-261 leaves them. */
typedef struct epic_domain_desc {
  double x0, y0, dx, dy;
  int32_t nx, ny;            /* node counts: Nx+1, Ny+1 */
  double L, W, LDe, wpe, vel_norm, density, DT;
  int32_t bc, shape;
  int32_t solver, poisson_op;
  double tolerance;          /* PCG: absolute ||r||_2 (reference) unless tol_relative */
  int32_t tol_relative;
  int32_t max_iteration;
  double sor_omega;          /* 1.4 in the reference */
  int32_t precision;         /* EPIC_FP64 | EPIC_FP32 particle storage */
  int32_t sort_interval;     /* counting sort by cell every k steps (0 = never) */
  uint64_t seed;             /* Philox key for native loads, MCC and injection */
  int32_t precond;           /* EPIC_PRECOND_* for EPIC_SOLVER_PCG_MIXED */
  int32_t reserved;
} epic_domain_desc;

/* Species constructor arguments (species.h:65-66). */
typedef struct epic_species_desc {
  char name[32];
  double mass, charge, spwt, temp;
  int64_t numparticle;       /* initial count; normalises KE/PE (species.cpp:264) */
  int64_t capacity;          /* device slots; 0 = numparticle */
  double vsx, vsy, fract_den;
  int32_t pushed;            /* 0 = skip push (ionfixed for "ion", main.cpp:419) */
} epic_species_desc;

/* Emitter (emitter.h:8-20). */
typedef struct epic_emitter_desc {
  double x0, y0, x1, y1, temp;
  int32_t numparticle;
  double vdx, vdy;
  int32_t species_idx1, species_idx2;
} epic_emitter_desc;

/* Internal conductors ("grids", grid.h:13-66; electrode.h:9-20), as
 * main.cpp:287-323 builds them from an ini's [Grids] section.  Coordinates are
 * in the domain's position units (the reference decks use node units, dx = 1).
 *   EPIC_COND_CIRCULAR  CircularGrid: electrode_number disks of radius
 *                       electrode_radius centred on a circle of radius
 *                       grid_radius about (x_center, y_center).
 *   EPIC_COND_RECT      ReactConductorGrid: [min_x, max_x] x [min_y, max_y].
 * voltage is the normalised value main.cpp:310,321 computes. */
#define EPIC_COND_CIRCULAR 0
#define EPIC_COND_RECT 1
typedef struct epic_conductor_desc {
  int32_t kind;
  int32_t x_center, y_center, electrode_number; /* CIRCULAR (main.cpp:305-307) */
  double grid_radius, electrode_radius;         /* CIRCULAR (main.cpp:308-309) */
  int32_t min_x, max_x, min_y, max_y;           /* RECT (main.cpp:317-320) */
  double voltage;
} epic_conductor_desc;

/* Collision processes for epic_mcc_setup. */
#define EPIC_MCC_E_ELASTIC 0
#define EPIC_MCC_E_EXCITATION 1
#define EPIC_MCC_E_IONIZATION 2
#define EPIC_MCC_I_ELASTIC 3
#define EPIC_MCC_I_CX 4
#define EPIC_MCC_NPROC 5

/* Null-collision MCC set-up.  Cross sections are tables on one energy grid,
 * log-spaced from e_min to e_max (eV) with ntab points, linearly interpolated
 * in log E and held at the end values outside; sigma[p*ntab + k] in m^2 for
 * process p.  Electron-neutral energies are 0.5*m_e*v^2 with the neutral at
 * rest (collision.cpp:194-196); ion-neutral energies are 0.5*m_i*g^2 with g
 * relative to a Maxwellian neutral at gas_temp_K. */
typedef struct epic_mcc_desc {
  int32_t electron_sp, ion_sp;   /* species indices; ion_sp receives ionisation ions */
  double gas_density;            /* m^-3 */
  double gas_temp_K;             /* neutral temperature (0 = at rest, reference) */
  double gas_mass;               /* kg; used for the e-neutral mass ratio */
  double e_exc_th, e_ion_th;     /* eV thresholds: 11.5, 15.8 in the reference */
  int32_t ntab;
  double e_min, e_max;
  const double* sigma;           /* EPIC_MCC_NPROC * ntab */
  int32_t enable[EPIC_MCC_NPROC];
  int32_t ref_selection;         /* 1: the reference's two-draw type choice
                                    (collision.cpp:117,125); 0: one draw against
                                    cumulative frequencies */
  int32_t event_cap;             /* ionisation births per step (0: 16384, or one per
                                    64 electron slots when that is more; at most
                                    2^28, larger values are rejected) */
} epic_mcc_desc;

typedef struct epic_mcc_stats {
  int64_t count[EPIC_MCC_NPROC];  /* collisions per process, cumulative */
  double eloss_ev;                /* energy removed by thresholds (eV), cumulative */
  int64_t births;
} epic_mcc_stats;

/* ---- lifetime ---------------------------------------------------------- */
/* Domain(x0,y0,dx,dy,nx,ny) (domain.h:141) + Species(...) x nsp
 * (species.h:65-66) + FieldSolve(domain, grids) (field.h:19). */
int epic_create(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, epic_ctx** out);
int epic_destroy(epic_ctx* ctx);
const char* epic_last_error(const epic_ctx* ctx);
int epic_sync(epic_ctx* ctx);
int epic_device_name(char* buf, int len);

/* ---- particles --------------------------------------------------------- */
/* Species::part_list contents (species.h:63).  Upload replaces the list;
 * ids are the upload order.  Download returns the live particles in storage
 * order (absorbed ones are dropped, order of the rest kept). */
int epic_upload_particles(epic_ctx* ctx, int sp, const double* x, const double* y, const double* vx,
                          const double* vy, const double* vz, int64_t n);
int epic_download_particles(epic_ctx* ctx, int sp, double* x, double* y, double* vx, double* vy,
                            double* vz, uint32_t* id, int64_t cap, int64_t* n_out);
int epic_count(epic_ctx* ctx, int sp, int64_t* n_out);
/* Native load (replaces Init::Init, init.cpp:3-102, in its "random" mode):
 * positions uniform over the domain, velocities Maxwellian 3V with
 * sigma = sqrt(kT/m) plus drift (vsx, vsy)*vel_norm, from Philox keyed by
 * (seed, species, particle). */
int epic_load_species(epic_ctx* ctx, int sp, int64_t n, int vz_thermal);

/* ---- per-step operators ------------------------------------------------ */
int epic_scatter(epic_ctx* ctx, int sp);               /* Species::ScatterSpecies_serial (species.h:71) */
int epic_compute_rho(epic_ctx* ctx);                   /* Domain::ComputeRho (domain.h:169) */
int epic_solve_potential(epic_ctx* ctx, int* iters, double* err); /* FieldSolve::PotentialSolver (field.h:26) */
/* No counterpart in the reference: epic_solve_potential without any device
 * or pinned-host allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The context's solver must be EPIC_SOLVER_PCG_MIXED with
 * EPIC_PRECOND_SPECTRAL; the result is that solve's, within its tolerance.
 * epic_solve_potential allocates the sine-transform matrices and work
 * arrays on a context's first solve; this variant keeps them in a static
 * 5 MiB device array per GPU (held until the process ends), which limits the
 * transform dimensions to 512: node grids up to 514 x 514 with Dirichlet
 * walls, 512 x 512 with pcgsolver's operator (EPIC_ERR_CAPACITY beyond).
 * Every call on a GPU shares that array: calls that overlap on one GPU (on
 * different streams or threads) race. */
int epic_solve_potential_noalloc(epic_ctx* ctx, int* iters, double* err);
int epic_efield(epic_ctx* ctx);                        /* FieldSolve::CalculateEfield (field.h:25) */
int epic_push(epic_ctx* ctx, int sp, int sub_cycle);   /* Species::Push_species_serial (species.h:70) */
int epic_rewind(epic_ctx* ctx, int sp);                /* Species::Rewind_species (species.h:72) */
int epic_inject(epic_ctx* ctx, int step);              /* Emitter::inject x n_emitters (emitter.h:26) */
int epic_add_emitter(epic_ctx* ctx, const epic_emitter_desc* e);
int epic_mcc_setup(epic_ctx* ctx, const epic_mcc_desc* d); /* CollisionHandler(domain) + tables (collision.h:29-31) */
int epic_collide(epic_ctx* ctx, int step);             /* CollisionHandler::handle_collisions (collision.h:33) */
int epic_mcc_get_stats(epic_ctx* ctx, epic_mcc_stats* out);
/* The conductors of FieldSolve(domain, grids) / Push_species_serial(.., grids)
 * (field.h:19, species.h:70); replaces any earlier list, n = 0 clears it.
 *  - A push removes a particle whose new position (after the walls) lies in a
 *    conductor (species.cpp:84-98, 111-124, 152-165).  A rectangle absorbs
 *    from the start; a circular grid's electrodes exist only once an SOR solve
 *    has run (CircularGrid::initialize, grid.cpp:40-50, is called by
 *    gssolver, field.cpp:100-103), so with the PCG solvers it never absorbs.
 *  - Each SOR solve (EPIC_SOLVER_GS) first sets rho := voltage on the
 *    conductors' nodes, in list order: 100 points of each electrode's circle
 *    (electrode.cpp:7-16) and a rectangle's boundary nodes (grid.cpp:68-85).
 *  - EPIC_ERR_ARG if a rectangle's boundary leaves the node grid (the
 *    reference's bounds-checked rho(i, j) throws), or for a batch member.
 * The device loader (epic_load_species) does not avoid conductors; the
 * reference's loader does (init.cpp:71-96), as does the host Init. */
int epic_set_conductors(epic_ctx* ctx, const epic_conductor_desc* c, int n);
int epic_sort(epic_ctx* ctx, int sp);                  /* no reference counterpart: stable counting sort by cell */
int epic_compact(epic_ctx* ctx, int sp);               /* order-preserving removal, as vector::erase (species.cpp:95-121) */

/* One timestep of main.cpp:395-455 (inject, scatter, rho, solve, E, push,
 * collide), with the next step's deposit fused into the push. */
int epic_step(epic_ctx* ctx, int ts, int* iters);
/* The initial field solve and half-step rewind of main.cpp:366-382. */
int epic_initial_field(epic_ctx* ctx);

/* ---- diagnostics ------------------------------------------------------- */
int epic_compute_ke(epic_ctx* ctx, int sp, int norm_sp, double out3[3]);       /* Species::Compute_KE (species.h:73) */
int epic_compute_momentum(epic_ctx* ctx, int sp, int norm_sp, double out3[3]); /* Species::Compute_Momentum (species.h:74) */
int epic_compute_pe(epic_ctx* ctx, int norm_sp, double* out);                  /* Domain::Compute_PE (domain.h:196) */

/* Node arrays: which = 0 rho, 1 phi, 2 efx, 3 efy, 10+s den of species s. */
int epic_get_field(epic_ctx* ctx, int which, double* out);
int epic_set_field(epic_ctx* ctx, int which, const double* in);
int epic_error_counters(epic_ctx* ctx, int64_t out[EPIC_NERR]);
int epic_clear_error_counters(epic_ctx* ctx);

/* ---- timing ------------------------------------------------------------ */
/* Stage timers: HIP event pairs recorded on the context's stream around
 * each stage without synchronising; epic_timing_read resolves them.  Stages:
 * 0 deposit, 1 rho, 2 solve, 3 efield, 4 push(+deposit, one pair per species
 * launch), 5 collide, 6 sort, 7 inject. */
#define EPIC_NSTAGE 8
/* on: 0 off, 1 every stage, or EPIC_TIMING_ONLY | (1 << stage) | ... to record
 * only those stages (each event pair costs the stream a few microseconds).
 * Enabling clears the accumulated times. */
#define EPIC_TIMING_ONLY 0x10000
int epic_timing_enable(epic_ctx* ctx, int on);
int epic_timing_read(epic_ctx* ctx, double ms[EPIC_NSTAGE], int64_t calls[EPIC_NSTAGE]);

/* ---- memory ------------------------------------------------------------- */
/* Device bytes epic_create allocates for this description (MCC tables and
 * the sine-transform solver's matrices come on first use), and what a live
 * context holds.  No reference counterpart: the reference allocates per
 * object on the host. */
int epic_estimate_bytes(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, int64_t* bytes);
int epic_device_bytes(epic_ctx* ctx, int64_t* bytes);
/* Device bytes epic_mcc_setup adds to a context of these species (tables,
 * ionisation event list, real-collision lists of the colliding species). */
int epic_mcc_estimate_bytes(const epic_mcc_desc* d, const epic_species_desc* sp, int nsp, int64_t* bytes);

/* ---- batched ensemble --------------------------------------------------- */
/* No reference counterpart: the reference runs one simulation per process
 * (main.cpp:24-36).  A batch owns member simulations (FP32, CIC, mixed
 * sine-transform Poisson solve) on one stream and advances all of them with
 * one launch per stage; the step can be captured in a HIP graph and
 * replayed.  epic_batch_member returns a member's context for the
 * per-simulation calls (load/upload, epic_initial_field, epic_mcc_setup,
 * epic_add_emitter, diagnostics, download); members are stepped only through
 * epic_batch_step and destroyed only through epic_batch_remove or
 * epic_batch_destroy.  Each member keeps its own step counter (the collision
 * random-stream key) and its own sort schedule. */
typedef struct epic_batch epic_batch;

#define EPIC_BATCH_NPUSH 5 /* push shapes (particles/thread, minimum waves/SIMD): 0 (4,8), 1 (8,1), 2 (16,1), 3 (4,1), 4 (8,5) */
#define EPIC_BATCH_NNODE 4 /* node-stage workgroup sizes: 64, 128, 256, 512 */

typedef struct {
  int refine;       /* FP32 sine-transform corrections per Poisson solve, each followed by an FP64 residual (0: 2) */
  int use_graph;    /* 1: capture the step in a HIP graph once per layout and replay it */
  int push_variant; /* -1: 0 until tuned; else a fixed push shape */
  int node_block;   /* 0: 256 until tuned; else 64, 128, 256 or 512 */
} epic_batch_opts;

typedef struct {
  int push_variant;                  /* chosen push shape */
  int node_block;                    /* chosen node-stage workgroup size */
  double push_ms[EPIC_BATCH_NPUSH];  /* mean push-stage time per step, per shape */
  double node_ms[EPIC_BATCH_NNODE];  /* mean node-stage (rho, solve, E) time per step, per size */
  int steps_used;                    /* real steps the measurement advanced every member by */
} epic_batch_tuning;

typedef struct {
  int members, groups, tiles, segments, chunks, mcc_passes;
  int push_variant, node_block;
  int64_t steps, graph_replays, graph_captures;
  int64_t device_bytes;  /* members' contexts + transform workspaces */
  double relres_max;     /* largest final ||b - A phi|| / ||b|| of the last step */
  double ms_total;       /* step time accumulated while timing is on */
  double stage_ms[4];    /* eager steps: rho+solve, efield, push, collide */
} epic_batch_stats;

int epic_batch_create(const epic_batch_opts* opts, epic_batch** out);
int epic_batch_destroy(epic_batch* b);
const char* epic_batch_last_error(const epic_batch* b);
int epic_batch_add(epic_batch* b, const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, int* id);
int epic_batch_member(epic_batch* b, int id, epic_ctx** ctx);
int epic_batch_remove(epic_batch* b, int id);
/* Marks a freshly loaded member to take its initial field inside the next
 * batched step (main.cpp:366-382: deposit, rho, solve, E, then
 * Rewind_species for every species) instead of a push; that step does not
 * count as one of the member's steps.  The alternative is epic_initial_field
 * on the member's context before its first batched step. */
int epic_batch_init(epic_batch* b, int id);
int epic_batch_step(epic_batch* b, int nsteps);
int epic_batch_sync(epic_batch* b);
/* Times every push shape and node workgroup size on real steps (results do
 * not depend on them) and keeps the fastest. */
int epic_batch_tune(epic_batch* b, int steps_per_candidate, epic_batch_tuning* out);
int epic_batch_relres(epic_batch* b, int id, double* relres);
int epic_batch_member_steps(epic_batch* b, int id, int64_t* steps);
int epic_batch_get_stats(epic_batch* b, epic_batch_stats* out);
/* Changes the options of a live batch (push_variant -1 / node_block 0 keep
 * the current shapes); the next step re-captures its graph. */
int epic_batch_configure(epic_batch* b, const epic_batch_opts* opts);
int epic_batch_timing_enable(epic_batch* b, int on);

#ifdef __cplusplus
}
#endif
#endif /* EPIC2D_H_ */
