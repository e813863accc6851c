// Internal declarations shared by the HIP sources of libepic2d.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/epic2d.h"

namespace epic {

// Physical constants, the values of include/domain.h:204-215.
constexpr double EPS_0 = 8.85418782e-12;
constexpr double QE = 1.602176565e-19;
constexpr double ME = 9.10938215e-31;
constexpr double K_b = 1.380648e-23;
constexpr double PI = 3.141592653;  // the reference's truncated pi
constexpr double eV = 1.602176565e-19;

// Fixed-point deposit scale (bits after the binary point of one particle's
// unit weight).  FP64 storage: 2^44 (per-contribution rounding < 3e-14 of a
// particle, 2^19 particles per node before overflow).  FP32 storage: 2^28,
// products of 14-bit cell offsets (2^-14 cell resolution, 2^35 particles per
// node before overflow).
constexpr int FIX_BITS_FP64 = 44;
constexpr int FIX_BITS_FP32 = 28;

// LDS deposit window: 8192 int64 node accumulators (64 KiB) per workgroup.
constexpr int DEP_WINDOW = 8192;
// MCC selection pass: a wavefront selects MCC_SUB slots at a time within its
// range of slots (a multiple of 64: one range per resident wavefront for a
// single simulation, MCC_RANGE_BATCH slots in batched runs)
constexpr int MCC_SUB = 512;
constexpr int MCC_RANGE_BATCH = 2048;
constexpr int MCC_EVENT_MAX = 1 << 28;  // ionisation event list entries at most
constexpr int PART_BLOCK = 256;  // threads per workgroup in particle kernels

// FP32 particle storage is tiled: the x, y, vx, vy of 256 consecutive slots
// form one 4 KiB tile (x[256] y[256] vx[256] vy[256]), so that a wavefront's
// chunk of the push is one contiguous block of memory instead of four
// streams.  A species' y, vx, vy pointers are its x pointer + 256, + 512,
// + 768, and slot i of any of the four lives at ptile(i); vz and id are plain
// arrays.  FP64 storage is one array per coordinate (pslot<double> = i).
constexpr int64_t PTILE = 256;
__host__ __device__ __forceinline__ int64_t ptile(int64_t i) { return ((i >> 8) << 10) + (i & 255); }
template <typename P>
__host__ __device__ __forceinline__ int64_t pslot(int64_t i) {
  return sizeof(P) == 4 ? ptile(i) : i;
}

struct PushF32;

struct SpeciesDev {
  epic_species_desc desc;
  int64_t cap = 0;
  int sort_mult = 1;        // sorted every dom.sort_interval * sort_mult steps
  int64_t n_host = 0;       // host copy of the live+dead slot count (authoritative unless births)
  // absorbed (NaN x) slots may exist that neither the walls, the conductors
  // nor the device's non-finite count account for: an uploaded NaN position,
  // or a non-finite count cleared before this species' next sort/compaction
  bool maybe_dead = false;
  void* x = nullptr;        // P*  (P = double | float)
  void* y = nullptr;
  void* vx = nullptr;
  void* vy = nullptr;
  void* vz = nullptr;
  uint32_t* id = nullptr;
  // sort/compaction double buffers
  void* x2 = nullptr;
  void* y2 = nullptr;
  void* vx2 = nullptr;
  void* vy2 = nullptr;
  void* vz2 = nullptr;
  uint32_t* id2 = nullptr;
  unsigned long long* Q = nullptr;  // fixed-point deposit accumulator, nnodes
  double* den = nullptr;            // normalised density, nnodes
  int64_t* d_n = nullptr;           // device [slot count, next id] (births append here)
};

struct Emitter {
  epic_emitter_desc d;
};

struct MccState {
  bool on = false;
  epic_mcc_desc d{};
  double* sigma = nullptr;               // device NPROC * ntab
  double nu_max_e = 0, nu_max_i = 0;     // s^-1
  unsigned long long* counts = nullptr;  // device NPROC + 2 (births, dropped)
  double* events = nullptr;              // device event_cap * 8
  int* event_count = nullptr;            // device
  int event_cap = 0;
  // real-collision lists of the two-pass null-collision step (pass 1 appends
  // (slot, process | (event slot + 1) << 3), pass 2 reads): electron list
  // (list_len_e entries) then ion list; their two counts; the electrons'
  // stored ionisation events per selection range (ranges of 64 slots at least)
  uint2* list = nullptr;
  int* list_n = nullptr;
  int* chunk_ev = nullptr;
  int64_t list_len_e = 0, list_len_i = 0;
};

// One absorbing shape on the device: a disk (one electrode, a, b centre and
// c = r*r) or a rectangle [a, b] x [c, d].  Rectangles come first in the
// table, then the electrodes of the circular grids.
enum { COND_DISK = 0, COND_RECT = 1 };
struct CondDev {
  int kind;
  double a, b, c, d;
};

// Whether (x, y) lies in any of the first n shapes: Electrode::Iselectrode
// (electrode.cpp:19-26), ReactConductorGrid::IsGrid (grid.cpp:60-63).
__host__ __device__ inline bool in_conductor(const CondDev* t, int n, double x, double y) {
  for (int k = 0; k < n; k++) {
    const CondDev q = t[k];
    if (q.kind == COND_DISK) {
      const double dx = x - q.a, dy = y - q.b;
      const double d2 = dx * dx + dy * dy;
      if (d2 <= q.c) return true;
    } else if (x >= q.a && x <= q.b && y >= q.c && y <= q.d) {
      return true;
    }
  }
  return false;
}

struct Conductors {
  std::vector<epic_conductor_desc> desc;
  CondDev* table = nullptr;  // device: rectangles, then electrodes
  int nrect = 0, nelec = 0;
  bool elec_live = false;    // the electrodes exist once an SOR solve has run
  uint32_t* cmask = nullptr; // device: one bit per cell, set where a shape's
                             // bounding box (widened by one cell) reaches
  int* node = nullptr;       // device: nodes an SOR solve sets to a voltage
  double* value = nullptr;
  int nnode = 0;
  bool any() const { return nrect + nelec > 0; }
  int active() const { return nrect + (elec_live ? nelec : 0); }
};

// Host-visible result of a device reduction: the values are stored first,
// seq last with a system-scope release, so a host that reads seq (acquire)
// sees the values of that launch.
struct HostScal {
  unsigned long long bb, rr;  // bit patterns of two doubles
  unsigned long long seq;
};

struct Ctx {
  epic_domain_desc dom{};
  int nsp = 0;
  std::vector<SpeciesDev> sp;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false for members of a batch, which share its stream
  size_t bytes = 0;        // device memory allocated at creation
  int device = 0;
  int num_cu = 0;
  int fix_bits = FIX_BITS_FP64;
  size_t nnodes = 0;

  double* rho = nullptr;
  double* phi = nullptr;
  double* efx = nullptr;
  double* efy = nullptr;
  void* E2 = nullptr;  // interleaved (ex, ey) per node in particle precision
  void* Ec = nullptr;  // FP32: per-cell bilinear coefficients of (ex, ey), 8 floats

  // Poisson workspace
  double* b64 = nullptr;
  double* r64 = nullptr;
  void* ws_v = nullptr;   // 2 * nnodes of T (double-buffered search direction)
  void* ws_w = nullptr;   // nnodes of T
  void* ws_x = nullptr;   // nnodes of T (inner solution)
  double* ws_part = nullptr;  // per-block partial sums (3 * max_blocks)
  double* ws_scalar = nullptr;  // solver scalars written back to host
  // pinned host words of the mixed solver's convergence check (written by
  // the device, polled by the host) and the last sequence number used
  HostScal* hscal = nullptr;
  unsigned long long hscal_seq = 0;
  // fused step: the sine-transform solve launches the E-field right after
  // its convergence check (so it runs while the host waits on the check);
  // efield_done: the launched E-field is that of the final phi
  bool efield_in_solve = false;
  bool efield_done = false;
  // fused step: the charge density is finalised by the sine-transform
  // solve's first kernel (rho_finalize's work in the residual's launch)
  bool rho_in_solve = false;
  int* ws_iter = nullptr;
  int max_solver_blocks = 0;
  // co-resident workgroups per CU of each particle kernel (queried once)
  std::unordered_map<const void*, int> occupancy;
  // spectral solver workspace
  double* sp_S = nullptr;       // sine-transform matrices
  double* sp_T = nullptr;       // temporaries
  double* sp_lam = nullptr;
  double* nr_ws = nullptr;      // Newton-CG solver: 6 * nnodes, allocated on first use
  int sp_m = 0, sp_n = 0;
  void* blas = nullptr;

  // sort workspace
  int* sort_count = nullptr;   // ncell + 2
  int* sort_start = nullptr;
  int* sort_fill = nullptr;
  uint32_t* sort_perm = nullptr;
  int64_t sort_cap = 0;
  int* blk_count = nullptr;    // compaction per-block counts
  int64_t blk_cap = 0;
  long long* scan_tiles = nullptr;  // multi-workgroup scan tile sums
  int scan_tiles_cap = 0;

  // diagnostics
  double* red_part = nullptr;  // per-block partials
  double* red_out = nullptr;   // final sums
  int red_blocks = 0;

  unsigned long long* errc = nullptr;  // EPIC_NERR device counters
  // counts the host observes (a solve whose convergence test runs on the
  // host and fails), added to the device counters when they are read
  long long errc_host[EPIC_NERR] = {};

  std::vector<Emitter> emitters;
  MccState mcc;
  Conductors cond;

  int push_variant = 0;        // FP32 push kernel shape (see particles.hip)
  int push_grid_per_cu = 0;    // FP32 / FP64 push workgroups per CU (0: co-resident maximum)
  int push_map = 1;            // push chunk walk: 0 per-XCD ranges, 1 round-robin (FP32 push)
  int part_map = 0;            // same, generic particle kernel
  int part_ppt = 2;            // FP64 particle kernels: particles per thread (2 or 4; generic also 8)
  int push64 = 1;              // FP64 CIC windowed push: 1 dedicated kernel, 0 generic
  int mcc_occ = 2;             // MCC collision pass: register budget for 1 or 2 workgroups per CU
  int mcc_sel_grid = 1;        // MCC selection pass: 1 resident workgroups, 0 one chunk per wavefront
  int sort_gather_wg = 1;      // sort gather: workgroups per CU
  int push_multi = 1;          // one launch for all pushed FP32 species (EPIC_PUSH_MULTI=0: one per species)
  int push_multi_waves = 8;    // register budget of that launch: 8 waves per SIMD (64 VGPRs), 1: the compiler's choice
  int cond_cell_mask = 1;      // 0: every cell flagged (the exact test everywhere)
  int pcg_block = 512;         // single-workgroup PCG size above 4096 nodes: 1024, 512 or 256 threads (EPIC_PCG_BLOCK)

  // species s is due for its counting sort after this step
  bool sort_due(int s) const {
    const long long k = (long long)dom.sort_interval * sp[s].sort_mult;
    return k > 0 && step_count % k == 0;
  }
  bool deposit_valid = false;  // Q holds the deposit of the current positions
  int step_count = 0;

  // timing: event pairs recorded on the stream without synchronising,
  // resolved when read
  bool timing = false;
  unsigned timing_mask = ~0u;  // stages that record events while timing is on
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_stage;  // stage of each recorded pair
  int ev_used = 0;
  int ev_open = -1;           // index of the pair whose start was recorded
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double stage_ms[EPIC_NSTAGE] = {};
  int64_t stage_calls[EPIC_NSTAGE] = {};

  std::string err;
};

// ---- launchers (one .hip file each) ---------------------------------------
// deposit.hip
hipError_t launch_deposit(Ctx* c, int s);
hipError_t launch_rho_finalize(Ctx* c);
hipError_t launch_den_only(Ctx* c, int s);
hipError_t launch_zero_Q(Ctx* c);
// particles.hip
hipError_t launch_push(Ctx* c, int s, int sub_cycle, bool deposit, double rewind_factor);
bool push_all_ok(const Ctx* c);
hipError_t launch_push_all(Ctx* c);
hipError_t launch_load(Ctx* c, int s, int64_t n, int vz_thermal);
hipError_t launch_compact(Ctx* c, int s);
hipError_t launch_sort(Ctx* c, int s);
hipError_t launch_inject(Ctx* c, int step);
int64_t sort_key_space(int nx, int ny);  // size of the tile-major cell key space
// field.hip
hipError_t launch_solve(Ctx* c, int* iters, double* err);
// The FP32 sine-transform solve with its workspace in a static device array
// (no allocation); *status = an EPIC_ERR_* code and *why its text when the
// context cannot use it.
hipError_t launch_spectral_solve_noalloc(Ctx* c, int* solves, double* relres, int* status, const char** why);
hipError_t launch_conductor_rho(Ctx* c);
hipError_t launch_efield(Ctx* c);
struct NormsArgs;
struct FinArgs;
// The finalisation's arguments (all species: mask of every bit, do_rho 1) and
// the accumulator clear that follows it under periodic walls.
FinArgs fin_args(Ctx* c, unsigned mask, int do_rho);
hipError_t zero_q_after_fin(Ctx* c, const FinArgs& a);
// E-field (and the FP32 gather table) with a mixed solve's convergence norms
// in the same launch where the precision allows (FP32), else after them.
hipError_t launch_efield_norms(Ctx* c, const NormsArgs& na);
hipError_t launch_pack_E2(Ctx* c);
// diag.hip
hipError_t launch_particle_moments(Ctx* c, int s, double out7[7]);
hipError_t launch_field_energy(Ctx* c, double* sum_e2);
// mcc.hip
hipError_t launch_mcc(Ctx* c, int step);
double mcc_nu_max(const epic_mcc_desc& d, const std::vector<double>& sig, double m, bool electron);

// batch support
// out_of_place: the push writes species s's second buffer (x2), which the
// caller then swaps in with swap_xyv; otherwise it updates x in place.
PushF32 make_push_f32(Ctx* c, int s, double kv, double kx, bool out_of_place);
void swap_xyv(SpeciesDev& S);
void push_f32_consts(Ctx* c, int s, double* kv, double* kx);
size_t mcc_args_bytes();
size_t append_args_bytes();
// appends the FP32 collision passes and the birth append of context c
// (step counter index slot) as raw argument records
void mcc_batch_collect(Ctx* c, int slot, std::vector<unsigned char>& args, std::vector<unsigned char>& app);
hipError_t launch_mcc_batch(hipStream_t st, int num_cu, const void* d_args, int nargs, const void* d_app, int napp,
                            const int* d_steps);

// helpers (capi.hip)
int64_t live_count(Ctx* c, int s, hipError_t* e);
hipError_t sync_counts(Ctx* c);
void stage_begin(Ctx* c, int stage);
void stage_end(Ctx* c, int stage);
// context creation on an existing stream (ext non-null: not owned)
int ctx_create(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, hipStream_t ext, epic_ctx** out);
Ctx* ctx_of(epic_ctx* h);
size_t ctx_bytes_estimate(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, int num_cu);

}  // namespace epic

#define EPIC_HIP_TRY(expr)                 \
  do {                                     \
    hipError_t _e = (expr);                \
    if (_e != hipSuccess) return _e;       \
  } while (0)
