// ORACLE — test infrastructure only.
//
// A plain FP64 CPU restatement of ePIC2D_mcc's per-timestep particle loop
// (koushik9988/ePIC2D_mcc).  It is the checker the HIP path is compared
// against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and never as the thing measured or shipped.
//
// Every function names the reference file:line it restates.  Storage follows
// the reference (array-of-structs particles kept in insertion order, removal
// by order-preserving erase, node arrays indexed (i, j) -> i*ny + j as
// Matrix<double>(nx, ny) and flatten() do), and so does the floating-point
// operation order, so that this file reproduces the reference bit for bit
// wherever the reference is deterministic.
//
// Pinning (what checks each part, tests/test_oracle_*.py):
//   setup/normalisation   closed-form plasma parameters (LDe, wpe, spwt) at
//                         the 1e13 m^-3 / 1 eV workload, independent of code
//   mt19937 + uniform     the C++ standard's 10000th-output value 4123659995
//   SampleVel / Init      Irwin-Hall moments of the three-uniform sum (mean,
//                         variance kT/m, excess kurtosis -0.4, the 3 sigma
//                         bound and its tail mass), uniform positions (KS)
//   Scatter / ComputeRho  brute-force CIC/NGP weights on hand-placed
//                         particles, charge conservation, edge/fold rules
//   pcg()                 the reference's own linearalgebra/ templates,
//                         compiled into oracle/_ref, bit for bit; plus
//                         manufactured sine eigen-solutions
//   gssolver              manufactured solutions of the Dirichlet problem
//   CalculateEfield       exact central differences of quadratic phi
//   Push / Rewind         uniform-field closed form, harmonic-oscillator
//                         frequency 2 asin(w dt/2)/dt, wrap/absorb rules
//   KE / PE / momentum    closed forms on hand-built states
//   cross sections        thresholds and continuity there, the elastic
//                         fit's E -> 0 limit, the measured argon ionisation
//                         maximum; the other fit coefficients are unpinned
//   direct MCC            per-process counts against 1 - exp(-nu dt) and the
//                         two-draw selection rule (5 sigma); per-process
//                         kinematics: |g| kept (elastic), 11.5 eV removed
//                         (excitation), e_sc + e_ej = E - 15.8 eV with the
//                         ejected-energy law and the coplanar angles
//                         (ionisation), isotropic cos(chi) (KS)
//   Emitter::inject       positions uniform on the emitter box, velocity
//                         moments per target species (positions drawn from a
//                         seeded stream in place of std::random_device)
//   Philox4x32-10         the Random123 known-answer vectors

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <omp.h>

namespace {

// Host threads of the all-cores timing variant (orc_set_threads); 1 runs
// the sequential restatement.  With more, the push and the deposit split the
// particle list into contiguous per-thread ranges (the push keeps the
// survivors' order; the deposit sums per-thread grids, so its FP64 sums are
// not the sequential ones) and the PCG's operator product splits its rows.
// Timing only: every parity check runs on one thread.
int g_threads = 1;

// Physical constants, include/domain.h:204-215.
const double EPS_0 = 8.85418782e-12;
const double QE = 1.602176565e-19;
const double ME = 9.10938215e-31;
const double K_b = 1.380648e-23;
const double PI = 3.141592653;
const double EV_to_K = QE / K_b;
const double eV = 1.602176565e-19;

// include/collision.h:14-15 and include/cross_sections.h:14-15.
const int CS_RANGES = 5000000;
const double DE_CS = 0.005;
const double E_EXC_TH = 11.5;
const double E_ION_TH = 15.8;

enum { BC_PBC = 0, BC_OPEN = 1, BC_REFLECTIVE = 2 };
enum { SHAPE_CIC = 0, SHAPE_NGP = 1 };
enum { SOLVER_PCG = 0, SOLVER_GS = 1, SOLVER_GS_FLUID = 4, SOLVER_NEWTON_CG = 5 };
enum { LOAD_RANDOM = 0, LOAD_UNIFORM = 1, LOAD_SIN = 2, LOAD_COS = 3, LOAD_GAUSSIAN = 4 };

struct Particle {
  double pos[2];
  double vel[3];
};

}  // namespace

extern "C" {

// Raw ini-level inputs that main.cpp:42-149 reads.
struct orc_inputs {
  int Nx, Ny;
  double x0, y0;
  int NUM_TS;
  double DT_coeff;
  double density;
  int norm_scheme;
  double L_scale;
  int T_scale_kind;    // 0 omegape, 1 omegapi, 2 numeric
  double T_scale_value;
  int nsp;
  const double* mass;
  const int* nparticles;
  const double* temps;
  const int* charge_signs;
  const double* frac;
  const double* vsx;
  const double* vsy;
};

// Everything main.cpp:151-261 derives.
struct orc_setup {
  double LDe, wpe, LDi, wpi, CS;
  double L, W, vel_norm;
  double dx, dy, DT;
  double x0, y0, Lx, Ly;
  int nx, ny;
  double normden[16];
  double spwt[16];
  double charge[16];
};

// Restates main.cpp:151-261 plus Domain::Domain (domain.cpp:7-16) and
// Domain::set_normscheme (domain.cpp:123-183).  Domain is built with y0 := x0
// (main.cpp:226).  vel_norm is L*W: set_normscheme reads vel_normscheme,
// which main never copies (main.cpp:83), and its trailing else wins for every
// value except 3 (domain.cpp:129-145).
int orc_compute_setup(const orc_inputs* in, orc_setup* out) {
  if (in->nsp < 2 || in->nsp > 16) return -1;
  double k = 0;
  for (int i = 0; i < in->nsp; i++) k += (-in->charge_signs[i]) * in->frac[i];  // main.cpp:151-155
  double normden[16];
  normden[1] = in->density;       // main.cpp:159
  normden[0] = in->density / k;   // main.cpp:160
  for (int i = 2; i < in->nsp; i++) normden[i] = in->frac[i] * normden[0];  // main.cpp:162-165

  const double* temps = in->temps;
  const double* mass = in->mass;
  // main.cpp:168-172
  double LDe = sqrt((EPS_0 * K_b * temps[0] * EV_to_K) / (normden[0] * QE * QE));
  double wpe = sqrt((normden[0] * QE * QE) / (mass[0] * EPS_0));
  double wpi = sqrt((normden[1] * QE * QE) / (mass[1] * EPS_0));
  double LDi = sqrt((EPS_0 * K_b * temps[1] * EV_to_K) / (normden[1] * QE * QE));
  double CS = sqrt(temps[0] * K_b * EV_to_K / mass[1]);

  double time_scale = 0;  // main.cpp:183-194
  if (in->T_scale_kind == 0) time_scale = wpe;
  else if (in->T_scale_kind == 1) time_scale = wpi;
  else time_scale = in->T_scale_value;
  double lenght_scale = in->L_scale;

  double dx = 0, dy = 0, DT = 0, stepSize;  // main.cpp:196-224
  int ns = in->norm_scheme;
  if (ns == 2 || ns == 4) {
    stepSize = LDi; dx = stepSize / LDi; dy = stepSize / LDi;
    DT = in->DT_coeff * (1.0 / wpi); DT = wpi * DT;
  }
  if (ns == 1 || ns == 3) {
    stepSize = LDe; dx = stepSize / LDe; dy = stepSize / LDe;
    DT = in->DT_coeff * (1.0 / wpe); DT = wpe * DT;
  }
  if (ns == 5) {
    stepSize = lenght_scale; dx = stepSize / lenght_scale; dy = stepSize / lenght_scale;
    DT = in->DT_coeff * (1.0 / time_scale); DT = time_scale * DT;
  }
  if (ns < 1 || ns > 5) return -2;

  out->nx = in->Nx + 1;  // main.cpp:46-47
  out->ny = in->Ny + 1;
  out->x0 = in->x0;
  out->y0 = in->x0;  // main.cpp:226 passes x0 twice
  out->dx = dx; out->dy = dy; out->DT = DT;
  out->Lx = (out->nx - 1) * dx;  // domain.cpp:14-15
  out->Ly = (out->ny - 1) * dy;
  out->LDe = LDe; out->wpe = wpe; out->LDi = LDi; out->wpi = wpi; out->CS = CS;
  if (ns == 1 || ns == 3) { out->L = LDe; out->W = wpe; }           // domain.cpp:125-146
  else if (ns == 2 || ns == 4) { out->L = LDi; out->W = (ns == 2) ? wpi : wpe; }  // :147-168
  else { out->L = lenght_scale; out->W = time_scale; }              // :169-178 (members main never sets; the ini values are used)
  out->vel_norm = out->L * out->W;
  for (int i = 0; i < in->nsp; i++) {
    out->normden[i] = normden[i];
    // main.cpp:256
    out->spwt[i] = normden[i] * out->Lx * out->Ly * out->L * out->L / in->nparticles[i];
    out->charge[i] = in->charge_signs[i] * QE;  // main.cpp:259
  }
  return 0;
}

// The draw sequence of the reference's global Rnd: std::mt19937 seeded 0 and
// uniform_real_distribution<double>(0, 1) (domain.h:218-231, domain.cpp:4).
// The oracle uses the same toolchain library, so the sequence is the same.
void orc_rnd_draws(unsigned seed, int n, double* out) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<double> d(0.0, 1.0);
  for (int i = 0; i < n; i++) out[i] = d(g);
}

unsigned orc_mt19937_nth(unsigned seed, int n) {
  std::mt19937 g(seed);
  unsigned v = 0;
  for (int i = 0; i < n; i++) v = g();
  return v;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Same algorithm as the HIP path's counter-based stream.
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; r++) {
    philox_round(c, k);
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
  }
  for (int i = 0; i < 4; i++) out[i] = c[i];
}

// Philox2x32-10 of the same paper: one multiply-hi/lo by 0xD256D193 per round,
// a 32-bit key advanced by 0x9E3779B9 after each round (the HIP path's
// candidate draw of the collision step).
void orc_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]) {
  uint32_t c0 = ctr[0], c1 = ctr[1];
  for (int r = 0; r < 10; r++) {
    const uint64_t p = (uint64_t)0xD256D193u * c0;
    const uint32_t n0 = (uint32_t)(p >> 32) ^ key ^ c1;
    c1 = (uint32_t)p;
    c0 = n0;
    key += 0x9E3779B9u;
  }
  out[0] = c0;
  out[1] = c1;
}

}  // extern "C"

namespace {

inline double u53(uint32_t a, uint32_t b) {
  // 53 random bits -> [0, 1)
  uint64_t m = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  return (double)m * (1.0 / 9007199254740992.0);
}

struct Sim {
  // Domain (domain.h:26-201): only the members the hot path reads.
  double x0, y0, dx, dy, Lx, Ly;
  int nx, ny;
  double L, W, LDe, wpe, vel_norm, density, DT;
  int bc, shape, solver;
  double tolerance;
  int max_iteration;
  double error = 0;
  int last_iters = 0;
  bool enable_elastic, enable_excitation, enable_ionization;
  double GAS_DENSITY;

  std::vector<double> phi, rho, efx, efy;  // Matrix<double>(nx, ny), i*ny + j

  struct Sp {
    double mass, charge, spwt, temp;
    int numparticle;
    double vsx, vsy;
    int load_type, load_n;
    std::vector<double> den;
    std::vector<Particle> part_list;
  };
  std::vector<Sp> sp;

  std::mt19937 gen{0};                        // domain.h:223
  std::uniform_real_distribution<double> dist{0.0, 1.0};
  double rnd() { return dist(gen); }

  // internal conductors (grid.h:13-66): kind 0 CircularGrid, 1
  // ReactConductorGrid; a circular grid's electrodes (x0, y0, r) exist only
  // after its initialize() (an SOR solve)
  struct Grid {
    int kind;
    int xc, yc, ne;
    double grad, erad;
    int min_x, max_x, min_y, max_y;
    double voltage;
    std::vector<double> ex, ey, er;
  };
  std::vector<Grid> grids;

  // collision tables (collision.cpp:12-42)
  std::vector<double> sigma_tot_e, sigma_ela, sigma_exc, sigma_ionz;
  double delta_g = 0;

  size_t idx(int i, int j) const { return (size_t)i * ny + j; }
};

}  // namespace

extern "C" {

struct orc_domain {
  double x0, y0, dx, dy;
  int nx, ny;
  double L, W, LDe, wpe, vel_norm, density, DT;
  int bc, shape, solver;
  double tolerance;
  int max_iteration;
  int elastic, excitation, ionization;
  double gas_density;
};

struct orc_species {
  double mass, charge, spwt, temp;
  int numparticle;
  double vsx, vsy;
  int load_type, load_n;
};

void* orc_create(const orc_domain* d, const orc_species* s, int nsp) {
  Sim* m = new Sim();
  m->x0 = d->x0; m->y0 = d->y0; m->dx = d->dx; m->dy = d->dy;
  m->nx = d->nx; m->ny = d->ny;
  m->Lx = (d->nx - 1) * d->dx;  // domain.cpp:14-15
  m->Ly = (d->ny - 1) * d->dy;
  m->L = d->L; m->W = d->W; m->LDe = d->LDe; m->wpe = d->wpe;
  m->vel_norm = d->vel_norm; m->density = d->density; m->DT = d->DT;
  m->bc = d->bc; m->shape = d->shape; m->solver = d->solver;
  m->tolerance = d->tolerance; m->max_iteration = d->max_iteration;
  m->enable_elastic = d->elastic; m->enable_excitation = d->excitation;
  m->enable_ionization = d->ionization; m->GAS_DENSITY = d->gas_density;
  size_t n = (size_t)d->nx * d->ny;
  m->phi.assign(n, 0); m->rho.assign(n, 0); m->efx.assign(n, 0); m->efy.assign(n, 0);
  for (int k = 0; k < nsp; k++) {
    Sim::Sp p;
    p.mass = s[k].mass; p.charge = s[k].charge; p.spwt = s[k].spwt; p.temp = s[k].temp;
    p.numparticle = s[k].numparticle; p.vsx = s[k].vsx; p.vsy = s[k].vsy;
    p.load_type = s[k].load_type; p.load_n = s[k].load_n;
    p.den.assign(n, 0);
    m->sp.push_back(p);
  }
  return m;
}

void orc_destroy(void* h) { delete (Sim*)h; }

void orc_reseed(void* h, unsigned seed) { ((Sim*)h)->gen.seed(seed); }

static bool on_any_grid(const Sim* m, double x, double y);

// Init::SampleVel, init.cpp:104-108 (draws sequenced left to right).
static double sample_vel(Sim* m, double mass, double temp) {
  double v_th = sqrt(2 * K_b * temp * EV_to_K / mass);
  double r1 = m->rnd();
  double r2 = m->rnd();
  double r3 = m->rnd();
  return v_th * sqrt(2) * (r1 + r2 + r3 - 1.5);
}

// Init::Init, init.cpp:3-102.  The gaussian load draws from its own
// generator seeded with `gauss_seed` (the reference seeds it from
// std::random_device, init.cpp:11-12).  A position on a conductor is drawn
// again (init.cpp:71-96); a deterministic load that lands on one would loop
// for ever there, and fails here (-1).
int orc_init_species(void* h, int s, unsigned gauss_seed) {
  Sim* m = (Sim*)h;
  Sim::Sp& sp = m->sp[s];
  std::mt19937 gen(gauss_seed);
  double mean_x = (m->x0 + m->Lx) / 2.0;
  double mean_y = (m->y0 + m->Ly) / 2.0;
  std::normal_distribution<> gx(mean_x, m->Lx / 5), gy(mean_y, m->Ly / 5);
  int n = sp.load_n;
  for (int p = 0; p < sp.numparticle; p++) {
    double x = 0, y = 0;
  draw:
    if (sp.load_type == LOAD_RANDOM) {
      x = m->x0 + m->Lx * m->rnd();
      y = m->y0 + m->Ly * m->rnd();
    } else if (sp.load_type == LOAD_UNIFORM) {
      int nxp = (int)sqrt((double)sp.numparticle);
      int px = p % nxp, py = p / nxp;
      x = m->x0 + px * (m->Lx / (nxp - 1));
      y = m->y0 + py * (m->Ly / (nxp - 1));
    } else if (sp.load_type == LOAD_SIN || sp.load_type == LOAD_COS) {
      x = m->x0 + p * (m->Lx / (sp.numparticle - 1));
      y = m->y0 + p * (m->Ly / (sp.numparticle - 1));
      double kx = 2 * PI * n / m->Lx;
      double ky = 2 * PI * n / m->Ly;
      if (sp.load_type == LOAD_SIN) { x += sin(kx * x); y += sin(ky * y); }
      else { x += cos(kx * x); y += cos(ky * y); }
    } else {
      x = gx(gen);
      y = gy(gen);
    }
    if (on_any_grid(m, x, y)) {
      if (sp.load_type != LOAD_RANDOM && sp.load_type != LOAD_GAUSSIAN) return -1;
      goto draw;
    }
    double vx = sample_vel(m, sp.mass, sp.temp) + sp.vsx * m->vel_norm;
    double vy = sample_vel(m, sp.mass, sp.temp) + sp.vsy * m->vel_norm;
    double vz = 0;
    vx /= m->vel_norm; vy /= m->vel_norm; vz /= m->vel_norm;
    Particle q;
    q.pos[0] = x; q.pos[1] = y; q.vel[0] = vx; q.vel[1] = vy; q.vel[2] = vz;
    sp.part_list.push_back(q);
  }
  return 0;
}

void orc_set_particles(void* h, int s, const double* x, const double* y, const double* vx,
                       const double* vy, const double* vz, long n) {
  Sim* m = (Sim*)h;
  auto& pl = m->sp[s].part_list;
  pl.resize(n);
  for (long i = 0; i < n; i++) {
    pl[i].pos[0] = x[i]; pl[i].pos[1] = y[i];
    pl[i].vel[0] = vx[i]; pl[i].vel[1] = vy[i]; pl[i].vel[2] = vz[i];
  }
}

long orc_count(void* h, int s) { return (long)((Sim*)h)->sp[s].part_list.size(); }

void orc_get_particles(void* h, int s, double* x, double* y, double* vx, double* vy, double* vz) {
  Sim* m = (Sim*)h;
  auto& pl = m->sp[s].part_list;
  for (size_t i = 0; i < pl.size(); i++) {
    x[i] = pl[i].pos[0]; y[i] = pl[i].pos[1];
    vx[i] = pl[i].vel[0]; vy[i] = pl[i].vel[1]; vz[i] = pl[i].vel[2];
  }
}

// which: 0 rho, 1 phi, 2 efx, 3 efy, 10+s den of species s
static std::vector<double>* grid_ptr(Sim* m, int which) {
  switch (which) {
    case 0: return &m->rho;
    case 1: return &m->phi;
    case 2: return &m->efx;
    case 3: return &m->efy;
    default: return &m->sp[which - 10].den;
  }
}
void orc_get_grid(void* h, int which, double* out) {
  auto* g = grid_ptr((Sim*)h, which);
  std::copy(g->begin(), g->end(), out);
}
void orc_set_grid(void* h, int which, const double* in) {
  auto* g = grid_ptr((Sim*)h, which);
  std::copy(in, in + g->size(), g->begin());
}

// Domain::XtoL / YtoL, domain.cpp:257-264.
static inline double XtoL(const Sim* m, double p) { return (p - m->x0) / m->dx; }
static inline double YtoL(const Sim* m, double p) { return (p - m->y0) / m->dy; }

// Grid::IsGrid: CustomGrid/CircularGrid test their electrodes
// (grid.cpp:12-35, Electrode::Iselectrode electrode.cpp:19-26);
// ReactConductorGrid::IsGrid (grid.cpp:60-63).
static bool is_grid(const Sim::Grid& g, double x, double y) {
  if (g.kind == 1) return (x >= g.min_x && x <= g.max_x && y >= g.min_y && y <= g.max_y);
  for (size_t e = 0; e < g.ex.size(); e++) {
    double dx = x - g.ex[e];
    double dy = y - g.ey[e];
    double distance_squared = dx * dx + dy * dy;
    double radius_squared = g.er[e] * g.er[e];
    if (distance_squared <= radius_squared) return true;
  }
  return false;
}

static bool on_any_grid(const Sim* m, double x, double y) {
  for (const Sim::Grid& g : m->grids)
    if (is_grid(g, x, y)) return true;
  return false;
}

// Grid::initialize: CircularGrid (grid.cpp:40-50) rebuilds its electrodes,
// each of which writes rho = voltage on 100 points of its circle (Electrode
// constructor, electrode.cpp:4-17); ReactConductorGrid (grid.cpp:68-85)
// writes the boundary of its node rectangle.  Returns -1 where the bounds-
// checked rho(i, j) would throw.
static int grid_initialize(Sim* m, Sim::Grid& g) {
  if (g.kind == 1) {
    int min_i = (int)XtoL(m, g.min_x), max_i = (int)XtoL(m, g.max_x);
    int min_j = (int)YtoL(m, g.min_y), max_j = (int)YtoL(m, g.max_y);
    for (int i = min_i; i <= max_i; ++i)
      for (int j = min_j; j <= max_j; ++j)
        if (i == min_i || i == max_i || j == min_j || j == max_j) {
          if (i < 0 || i >= m->nx || j < 0 || j >= m->ny) return -1;
          m->rho[m->idx(i, j)] = g.voltage;
        }
    return 0;
  }
  g.ex.clear(); g.ey.clear(); g.er.clear();
  for (int k = 0; k < g.ne; ++k) {
    double theta = k * (2.0 * M_PI / g.ne);
    double x0 = g.xc + g.grad * cos(theta);
    double y0 = g.yc + g.grad * sin(theta);
    double r = g.erad;
    const int N = 100;
    for (int n = 0; n < N; ++n) {
      double t = 2.0 * n * M_PI / N;
      int x = static_cast<int>(x0 + r * cos(t) + 0.5);
      int y = static_cast<int>(y0 + r * sin(t) + 0.5);
      if (x >= 0 && x < m->nx && y >= 0 && y < m->ny) m->rho[m->idx(x, y)] = g.voltage;
    }
    g.ex.push_back(x0); g.ey.push_back(y0); g.er.push_back(r);
  }
  return 0;
}

// Domain::Scatter, domain.cpp:205-228.  Returns -1 where the reference's
// bounds-checked Matrix::operator() would throw (linalg.h:473-480).
static int scatter(Sim* m, double lx, double ly, double value, std::vector<double>& f) {
  int i = (int)lx;
  int j = (int)ly;
  double di = lx - i;
  double dj = ly - j;
  int ngpx = (int)round(lx);
  int ngpy = (int)round(ly);
  if (m->shape == SHAPE_NGP) {
    if (ngpx < 0 || ngpx >= m->nx || ngpy < 0 || ngpy >= m->ny) return -1;
    f[m->idx(ngpx, ngpy)] += value;
  } else {
    if (i < 0 || i + 1 >= m->nx || j < 0 || j + 1 >= m->ny) return -1;
    f[m->idx(i, j)] += value * (1 - di) * (1 - dj);
    f[m->idx(i + 1, j)] += value * (di) * (1 - dj);
    f[m->idx(i, j + 1)] += value * (1 - di) * (dj);
    f[m->idx(i + 1, j + 1)] += value * (di) * (dj);
  }
  return 0;
}

// Domain::Interpolate, domain.cpp:231-255.
static int interpolate(const Sim* m, double lx, double ly, const std::vector<double>& f, double* val) {
  int i = (int)lx;
  int j = (int)ly;
  double di = lx - i;
  double dj = ly - j;
  int ngpx = (int)round(lx);
  int ngpy = (int)round(ly);
  if (m->shape == SHAPE_NGP) {
    if (ngpx < 0 || ngpx >= m->nx || ngpy < 0 || ngpy >= m->ny) return -1;
    *val = f[m->idx(ngpx, ngpy)];
  } else {
    if (i < 0 || i + 1 >= m->nx || j < 0 || j + 1 >= m->ny) return -1;
    *val = f[m->idx(i, j)] * (1 - di) * (1 - dj) + f[m->idx(i + 1, j)] * (di) * (1 - dj) +
           f[m->idx(i, j + 1)] * (1 - di) * (dj) + f[m->idx(i + 1, j + 1)] * (di) * (dj);
  }
  return 0;
}

// Species::ScatterSpecies_serial, species.cpp:172-226.
int orc_scatter(void* h, int s) {
  Sim* m = (Sim*)h;
  Sim::Sp& sp = m->sp[s];
  std::fill(sp.den.begin(), sp.den.end(), 0.0);
  int bad = 0;
  if (g_threads <= 1) {
    for (const Particle& p : sp.part_list) {
      double lx = XtoL(m, p.pos[0]);
      double ly = YtoL(m, p.pos[1]);
      if (scatter(m, lx, ly, sp.spwt, sp.den)) bad++;
    }
  } else {
    const size_t np = sp.part_list.size();
    std::vector<std::vector<double>> part(g_threads, std::vector<double>(sp.den.size(), 0.0));
#pragma omp parallel num_threads(g_threads) reduction(+ : bad)
    {
      const int t = omp_get_thread_num();
      const size_t b = np * t / g_threads, e = np * (t + 1) / g_threads;
      for (size_t k = b; k < e; k++) {
        const Particle& p = sp.part_list[k];
        if (scatter(m, XtoL(m, p.pos[0]), YtoL(m, p.pos[1]), sp.spwt, part[t])) bad++;
      }
    }
    for (int t = 0; t < g_threads; t++)
      for (size_t k = 0; k < sp.den.size(); k++) sp.den[k] += part[t][k];
  }
  double cell_area = m->dx * m->dy;
  double L2 = m->L * m->L;
  double norm = (cell_area * L2 * m->density);
  for (double& v : sp.den) v /= norm;  // Matrix::operator/= (linalg.h:339-349)
  int nx = m->nx, ny = m->ny;
  auto den = [&](int i, int j) -> double& { return sp.den[m->idx(i, j)]; };
  if (m->bc == BC_PBC) {
    for (int i = 0; i < nx; i++) { den(i, 0) += den(i, ny - 1); den(i, ny - 1) = den(i, 0); }
    for (int j = 0; j < ny; j++) { den(0, j) += den(nx - 1, j); den(nx - 1, j) = den(0, j); }
  } else {
    den(0, 0) *= 0.25; den(0, ny - 1) *= 0.25; den(nx - 1, 0) *= 0.25; den(nx - 1, ny - 1) *= 0.25;
    for (int i = 1; i < nx - 1; i++) { den(i, 0) *= 0.5; den(i, ny - 1) *= 0.5; }
    for (int j = 1; j < ny - 1; j++) { den(0, j) *= 0.5; den(nx - 1, j) *= 0.5; }
  }
  return bad;
}

// Domain::ComputeRho, domain.cpp:194-202: rho = 0; rho += den*(charge*(1/QE)).
void orc_compute_rho(void* h) {
  Sim* m = (Sim*)h;
  std::fill(m->rho.begin(), m->rho.end(), 0.0);
  for (auto& sp : m->sp) {
    double f = sp.charge * (1 / QE);
    for (size_t k = 0; k < m->rho.size(); k++) m->rho[k] += sp.den[k] * f;
  }
}

// Dense Matrix x vec row sum over the reference's 5-point A (field.cpp:223-232,
// linalg.h:397-408), restated matrix-free: the zero entries of a row add
// nothing, and the non-zero ones are summed in increasing column order.
static void apply_A(int nx, int ny, const std::vector<double>& v, std::vector<double>& u) {
  int n = nx * ny;
  for (int i = 0; i < n; i++) {
    int row = i / nx, col = i % nx;
    double sum = 0;
    if (row > 0) sum += 1.0 * v[i - nx];
    if (col > 0) sum += 1.0 * v[i - 1];
    sum += -4.0 * v[i];
    if (col < nx - 1) sum += 1.0 * v[i + 1];
    if (row < ny - 1) sum += 1.0 * v[i + nx];
    u[i] = sum;
  }
}

static double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double r = 0;  // vec::operator* (linalg.h:632-646)
  if (g_threads <= 1) {
    for (size_t i = 0; i < a.size(); i++) r += a[i] * b[i];
    return r;
  }
  const long n = (long)a.size();
#pragma omp parallel for num_threads(g_threads) reduction(+ : r) schedule(static)
  for (long i = 0; i < n; i++) r += a[i] * b[i];
  return r;
}

static void apply_A_par(int nx, int ny, const std::vector<double>& v, std::vector<double>& u) {
  if (g_threads <= 1) return apply_A(nx, ny, v, u);
  const int n = nx * ny;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int i = 0; i < n; i++) {
    int row = i / nx, col = i % nx;
    double sum = 0;
    if (row > 0) sum += 1.0 * v[i - nx];
    if (col > 0) sum += 1.0 * v[i - 1];
    sum += -4.0 * v[i];
    if (col < nx - 1) sum += 1.0 * v[i + 1];
    if (row < ny - 1) sum += 1.0 * v[i + nx];
    u[i] = sum;
  }
}

// pcg(), solvers.h:72-100, with M_inv = diagonalPreconditioner(A) = -1/4 I
// (function.h:96-106).  Iteration cap min(n*n, max_iteration) taken in 64 bits.
int orc_pcg(int nx, int ny, const double* b_in, double* x_out, int max_iteration, double tolerance) {
  int n = nx * ny;
  std::vector<double> b(b_in, b_in + n), x(n, 0.0), Ax(n), r(n), w(n), v(n), u(n);
  apply_A(nx, ny, x, Ax);
  for (int i = 0; i < n; i++) r[i] = b[i] - Ax[i];
  for (int i = 0; i < n; i++) w[i] = (1.0 / -4.0) * r[i];
  v = w;
  double alpha = dot(w, r);
  long long cap = std::min((long long)n * n, (long long)max_iteration);
  int iters = 0;
  for (long long it = 0; it < cap; it++) {
    if (std::sqrt(dot(r, r)) < tolerance) break;
    apply_A_par(nx, ny, v, u);
    double t = alpha / dot(v, u);
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int i = 0; i < n; i++) {
      x[i] = x[i] + v[i] * t;
      r[i] = r[i] - u[i] * t;
      w[i] = (1.0 / -4.0) * r[i];
    }
    double beta = dot(w, r);
    double s = beta / alpha;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int i = 0; i < n; i++) v[i] = w[i] + v[i] * s;
    alpha = beta;
    iters++;
  }
  std::copy(x.begin(), x.end(), x_out);
  return iters;
}

// FieldSolve::pcgsolver, field.cpp:212-264.
static void pcgsolver(Sim* m) {
  double dx2 = m->dx * m->dx;
  int nx = m->nx, ny = m->ny, n = nx * ny;
  std::vector<double> rhom(n);
  double s = (m->L * m->L) / (m->LDe * m->LDe);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) rhom[m->idx(i, j)] = -m->rho[m->idx(i, j)] * dx2 * s;
  for (int i = 0; i < nx; i++) { rhom[m->idx(i, 0)] = 0; rhom[m->idx(i, ny - 1)] = 0; }
  for (int j = 0; j < ny; j++) { rhom[m->idx(0, j)] = 0; rhom[m->idx(nx - 1, j)] = 0; }
  // flatten(rhom) (function.h:154-169) maps (i, j) -> i*ny + j: our layout.
  std::vector<double> x(n);
  m->last_iters = orc_pcg(nx, ny, rhom.data(), x.data(), m->max_iteration, m->tolerance);
  std::vector<double> Ax(n);
  apply_A(nx, ny, x, Ax);
  std::vector<double> err(n);
  for (int i = 0; i < n; i++) err[i] = Ax[i] - rhom[i];
  m->error = std::sqrt(dot(err, err)) / n;
  m->phi = x;  // unflatten(x, nx, ny) (function.h:171-189)
}

// FieldSolve::gssolver, field.cpp:70-139; the conductors' initialize()
// writes rho first (field.cpp:100-103).
static int gssolver(Sim* m) {
  double dx2 = m->dx * m->dx, dy2 = m->dy * m->dy;
  int nx = m->nx, ny = m->ny;
  std::fill(m->phi.begin(), m->phi.end(), 0.0);
  for (Sim::Grid& g : m->grids)
    if (grid_initialize(m, g)) return -1;
  double L2 = 0;
  double coeff = (2 / dx2) + (2 / dy2);
  auto phi = [&](int i, int j) -> double& { return m->phi[m->idx(i, j)]; };
  auto rho = [&](int i, int j) -> double { return m->rho[m->idx(i, j)]; };
  double s = (m->L * m->L) / (m->LDe * m->LDe);
  int it_done = 0;
  bool converged = false;
  for (int iter = 0; iter < m->max_iteration; iter++) {
    for (int i = 1; i < nx - 1; i++)
      for (int j = 1; j < ny - 1; j++) {
        double phi_new = (1 / coeff) * (rho(i, j) * s + ((phi(i + 1, j) + phi(i - 1, j)) / dx2) +
                                        ((phi(i, j + 1) + phi(i, j - 1)) / dy2));
        phi(i, j) = phi(i, j) + 1.4 * (phi_new - phi(i, j));
      }
    it_done = iter + 1;
    if (iter % 25 == 0) {
      double sum = 0;
      for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++) {
          double R = ((phi(i + 1, j) - 2 * phi(i, j) + phi(i - 1, j)) / dx2) +
                     ((phi(i, j + 1) - 2 * phi(i, j) + phi(i, j - 1)) / dy2) + (rho(i, j)) * s;
          sum += R * R;
        }
      L2 = sqrt(sum / (nx * ny));
      m->error = L2;
      if (L2 < m->tolerance) { converged = true; break; }
    }
  }
  (void)converged;
  m->last_iters = it_done;
  return 0;
}

// FieldSolve::gsfluidsolver, field.cpp:7-68: SOR on the Boltzmann-electron
// problem, with its boundary rho values and a residual check every 50
// iterations; conductors are not initialised here.
static void gsfluidsolver(Sim* m) {
  double dx2 = m->dx * m->dx, dy2 = m->dy * m->dy;
  int nx = m->nx, ny = m->ny;
  std::fill(m->phi.begin(), m->phi.end(), 0.0);
  double L2 = 0;
  double coeff = (2 / dx2) + (2 / dy2);
  auto phi = [&](int i, int j) -> double& { return m->phi[m->idx(i, j)]; };
  auto rho = [&](int i, int j) -> double& { return m->rho[m->idx(i, j)]; };
  for (int i = 0; i < nx; i++) {
    rho(i, 0) = 0;
    rho(i, ny - 1) = 10000 * (eV / (K_b * EV_to_K));
  }
  for (int i = 0; i < ny; i++) {
    rho(0, i) = 0;
    rho(nx - 1, i) = 0;
  }
  double s = (m->L * m->L) / (m->LDe * m->LDe);
  int it_done = 0;
  for (int iter = 0; iter < m->max_iteration; iter++) {
    for (int i = 1; i < nx - 1; i++)
      for (int j = 1; j < ny - 1; j++) {
        double phi_new = (1 / coeff) * ((rho(i, j) - exp(phi(i, j))) * s +
                                        ((phi(i + 1, j) + phi(i - 1, j)) / dx2) +
                                        ((phi(i, j + 1) + phi(i, j - 1)) / dy2));
        phi(i, j) = phi(i, j) + 1.4 * (phi_new - phi(i, j));
      }
    it_done = iter + 1;
    if (iter % 50 == 0) {
      double sum = 0;
      for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++) {
          double R = ((phi(i + 1, j) - 2 * phi(i, j) + phi(i - 1, j)) / dx2) +
                     ((phi(i, j + 1) - 2 * phi(i, j) + phi(i, j - 1)) / dy2) +
                     (rho(i, j) - exp(phi(i, j))) * s;
          sum += R * R;
        }
      L2 = sqrt(sum / (nx * ny));
      m->error = L2;
      if (L2 < m->tolerance) break;
    }
  }
  m->last_iters = it_done;
}

// FieldSolve::nrpcgsolver, field.cpp:141-210: ten Newton steps on the
// Boltzmann-electron problem, each a cg() (solvers.h:32-69: plain CG from
// x = 0, absolute ||r|| < 1e-4, at most min(n, 10000) iterations) on
// J = A - diag(exp(x)).  Kept as written: the right-hand side accumulates
// exp(x) L^2/LDe^2 over the Newton steps (rhovec = rhovec + rhovecx), while
// the Jacobian's diagonal takes exp(x) without the factor.  A is the
// reference's matrix (apply_A; J x adds (-exp(x_i)) x_i to the diagonal term
// in column order).
static void nrpcgsolver(Sim* m) {
  double dx2 = m->dx * m->dx;
  int nx = m->nx, ny = m->ny, n = nx * ny;
  double s = (m->L * m->L) / (m->LDe * m->LDe);
  std::vector<double> rhovec(n), rhovecx(n), x(n, 0.0), Ax(n), F(n), ex(n);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) rhovec[m->idx(i, j)] = -m->rho[m->idx(i, j)] * dx2 * s;
  for (int i = 0; i < nx; i++) { rhovec[m->idx(i, 0)] = 0; rhovec[m->idx(i, ny - 1)] = 0; }
  for (int j = 0; j < ny; j++) { rhovec[m->idx(0, j)] = 0; rhovec[m->idx(nx - 1, j)] = 0; }
  // J v: the row of A with its diagonal -4 replaced by (-4) - exp(x_i)
  auto apply_J = [&](const std::vector<double>& v, std::vector<double>& u) {
    for (int i = 0; i < n; i++) {
      int row = i / nx, col = i % nx;
      double sum = 0;
      if (row > 0) sum += 1.0 * v[i - nx];
      if (col > 0) sum += 1.0 * v[i - 1];
      sum += (-4.0 - ex[i]) * v[i];
      if (col < nx - 1) sum += 1.0 * v[i + 1];
      if (row < ny - 1) sum += 1.0 * v[i + nx];
      u[i] = sum;
    }
  };
  int total = 0;
  for (int iter = 0; iter < 10; iter++) {
    for (int i = 0; i < n; i++) rhovecx[i] = exp(x[i]) * s;
    for (int i = 0; i < n; i++) rhovec[i] = rhovec[i] + rhovecx[i];
    apply_A(nx, ny, x, Ax);
    for (int i = 0; i < n; i++) F[i] = Ax[i] - rhovec[i];
    for (int i = 0; i < n; i++) ex[i] = exp(x[i]);
    // cg(J, x_ini = 0, F, 10000, 1e-4)
    std::vector<double> y(n, 0.0), r0(n), p0(n), Ap(n);
    apply_J(y, Ap);
    for (int i = 0; i < n; i++) r0[i] = F[i] - Ap[i];
    p0 = r0;
    int cap = std::min(n, 10000);
    for (int it = 0; it < cap; it++) {
      if (std::sqrt(dot(r0, r0)) < 1e-4) break;
      apply_J(p0, Ap);
      double r02 = dot(r0, r0);
      double pAp = dot(p0, Ap);
      double alpha = r02 / pAp;
      for (int i = 0; i < n; i++) y[i] = y[i] + p0[i] * alpha;
      for (int i = 0; i < n; i++) r0[i] = r0[i] - Ap[i] * alpha;
      double r12 = dot(r0, r0);
      double beta = r12 / r02;
      for (int i = 0; i < n; i++) p0[i] = r0[i] + p0[i] * beta;
      total++;
    }
    for (int i = 0; i < n; i++) x[i] = x[i] - y[i];
  }
  m->error = 0;
  m->phi = x;
  m->last_iters = total;
}

// FieldSolve::PotentialSolver, field.cpp:266-272.
int orc_solve(void* h, double* err) {
  Sim* m = (Sim*)h;
  if (m->solver == SOLVER_GS) {
    if (gssolver(m)) return -1;
  } else if (m->solver == SOLVER_GS_FLUID) {
    gsfluidsolver(m);
  } else if (m->solver == SOLVER_NEWTON_CG) {
    nrpcgsolver(m);
  } else {
    pcgsolver(m);
  }
  if (err) *err = m->error;
  return m->last_iters;
}

// FieldSolve::CalculateEfield, field.cpp:276-331.
void orc_efield(void* h) {
  Sim* m = (Sim*)h;
  int nx = m->nx, ny = m->ny;
  auto phi = [&](int i, int j) { return m->phi[m->idx(i, j)]; };
  auto efx = [&](int i, int j) -> double& { return m->efx[m->idx(i, j)]; };
  auto efy = [&](int i, int j) -> double& { return m->efy[m->idx(i, j)]; };
  for (int i = 1; i < nx - 1; i++)
    for (int j = 0; j < ny; j++) efx(i, j) = -(phi(i + 1, j) - phi(i - 1, j)) / (2 * m->dx);
  if (m->bc == BC_OPEN || m->bc == BC_REFLECTIVE)
    for (int j = 0; j < ny; j++) {
      efx(0, j) = -(phi(1, j) - phi(0, j)) / m->dx;
      efx(nx - 1, j) = -(phi(nx - 1, j) - phi(nx - 2, j)) / m->dx;
    }
  if (m->bc == BC_PBC)
    for (int j = 0; j < ny; j++) {
      efx(0, j) = -(phi(1, j) - phi(nx - 2, j)) / (2 * m->dx);
      efx(nx - 1, j) = efx(0, j);
    }
  for (int i = 0; i < nx; i++)
    for (int j = 1; j < ny - 1; j++) efy(i, j) = -(phi(i, j + 1) - phi(i, j - 1)) / (2 * m->dy);
  if (m->bc == BC_OPEN || m->bc == BC_REFLECTIVE)
    for (int i = 0; i < nx; i++) {
      efy(i, 0) = -(phi(i, 1) - phi(i, 0)) / m->dy;
      efy(i, ny - 1) = -(phi(i, ny - 1) - phi(i, ny - 2)) / m->dy;
    }
  if (m->bc == BC_PBC)
    for (int i = 0; i < nx; i++) {
      efy(i, 0) = -(phi(i, 1) - phi(i, ny - 2)) / (2 * m->dy);
      efy(i, ny - 1) = efy(i, 0);
    }
}

// Species::Push_species_serial, species.cpp:29-170; after the walls a
// particle on a conductor is erased (:84-98, :111-124, :152-165).
// Returns the number of particles whose gather would have thrown.
int orc_push(void* h, int s, int sub_cycle) {
  Sim* m = (Sim*)h;
  Sim::Sp& sp = m->sp[s];
  int bad = 0;
  auto& pl = sp.part_list;
  auto one = [&](Particle part, std::vector<Particle>& kept, int& bad) {
    double qm = sp.charge / sp.mass;
    double lx = XtoL(m, part.pos[0]);
    double ly = YtoL(m, part.pos[1]);
    double ex = 0, ey = 0;
    if (interpolate(m, lx, ly, m->efx, &ex) || interpolate(m, lx, ly, m->efy, &ey)) bad++;
    double K = (m->density * QE * m->L) / (EPS_0 * m->W * m->vel_norm);
    part.vel[0] += qm * K * ex * (m->DT * sub_cycle);
    part.vel[1] += qm * K * ey * (m->DT * sub_cycle);
    double c = (m->vel_norm) / (m->L * m->W);
    part.pos[0] += c * part.vel[0] * (m->DT * sub_cycle);
    part.pos[1] += c * part.vel[1] * (m->DT * sub_cycle);
    if (m->bc == BC_PBC) {
      if (part.pos[0] < m->x0) part.pos[0] = part.pos[0] + m->Lx;
      else if (part.pos[0] >= m->x0 + m->Lx) part.pos[0] = part.pos[0] - m->Lx;
      if (part.pos[1] < m->x0) part.pos[1] = part.pos[1] + m->Ly;
      else if (part.pos[1] >= m->x0 + m->Ly) part.pos[1] = part.pos[1] - m->Ly;
    } else if (m->bc == BC_OPEN) {
      if (part.pos[0] < m->x0 || part.pos[0] >= m->x0 + m->Lx || part.pos[1] < m->x0 ||
          part.pos[1] >= m->x0 + m->Ly)
        return;  // erase (species.cpp:103-107); order of the rest is kept
    } else {
      if (part.pos[0] < m->x0) { part.pos[0] = m->x0 + (m->x0 - part.pos[0]); part.vel[0] = -part.vel[0]; }
      else if (part.pos[0] >= m->x0 + m->Lx) {
        part.pos[0] = m->x0 + m->Lx - (part.pos[0] - (m->x0 + m->Lx)); part.vel[0] = -part.vel[0];
      }
      if (part.pos[1] < m->x0) { part.pos[1] = m->x0 + (m->x0 - part.pos[1]); part.vel[1] = -part.vel[1]; }
      else if (part.pos[1] >= m->x0 + m->Ly) {
        part.pos[1] = m->x0 + m->Ly - (part.pos[1] - (m->x0 + m->Ly)); part.vel[1] = -part.vel[1];
      }
    }
    if (on_any_grid(m, part.pos[0], part.pos[1])) return;
    kept.push_back(part);
  };
  std::vector<Particle> kept;
  if (g_threads <= 1) {
    kept.reserve(pl.size());
    for (size_t k = 0; k < pl.size(); k++) one(pl[k], kept, bad);
  } else {
    const size_t np = pl.size();
    std::vector<std::vector<Particle>> part(g_threads);
#pragma omp parallel num_threads(g_threads) reduction(+ : bad)
    {
      const int t = omp_get_thread_num();
      const size_t b = np * t / g_threads, e = np * (t + 1) / g_threads;
      part[t].reserve(e - b);
      for (size_t k = b; k < e; k++) one(pl[k], part[t], bad);
    }
    std::vector<size_t> off(g_threads + 1, 0);
    for (int t = 0; t < g_threads; t++) off[t + 1] = off[t] + part[t].size();
    kept.resize(off[g_threads]);
#pragma omp parallel num_threads(g_threads)
    {
      const int t = omp_get_thread_num();
      std::copy(part[t].begin(), part[t].end(), kept.begin() + off[t]);
      std::vector<Particle>().swap(part[t]);
    }
  }
  pl.swap(kept);
  return bad;
}

// Species::Rewind_species, species.cpp:229-246.
int orc_rewind(void* h, int s) {
  Sim* m = (Sim*)h;
  Sim::Sp& sp = m->sp[s];
  int bad = 0;
  for (Particle& part : sp.part_list) {
    double qm = sp.charge / sp.mass;
    double lx = XtoL(m, part.pos[0]);
    double ly = YtoL(m, part.pos[1]);
    double ex = 0, ey = 0;
    if (interpolate(m, lx, ly, m->efx, &ex) || interpolate(m, lx, ly, m->efy, &ey)) bad++;
    double K = (m->density * QE * m->L) / (EPS_0 * m->W * m->vel_norm);
    part.vel[0] -= 0.5 * qm * K * ex * m->DT;
    part.vel[1] -= 0.5 * qm * K * ey * m->DT;
  }
  return bad;
}

// Species::Compute_KE, species.cpp:248-271 (this = species s, argument = ns).
void orc_ke(void* h, int s, int ns, double* out3) {
  Sim* m = (Sim*)h;
  Sim::Sp& sp = m->sp[s];
  Sim::Sp& nsp = m->sp[ns];
  double ke[3] = {0, 0, 0};
  for (const Particle& p : sp.part_list)
    for (int c = 0; c < 3; c++) ke[c] += (p.vel[c] * p.vel[c]) * (m->vel_norm) * (m->vel_norm);
  for (int c = 0; c < 3; c++) ke[c] *= 0.5 * (sp.spwt * sp.mass);
  double Th = (nsp.temp * eV) * (nsp.spwt) * nsp.numparticle;
  for (int c = 0; c < 3; c++) out3[c] = ke[c] / Th;
}

// Species::Compute_Momentum, species.cpp:273-297.
void orc_momentum(void* h, int s, int ns, double* out3) {
  Sim* m = (Sim*)h;
  Sim::Sp& sp = m->sp[s];
  Sim::Sp& nsp = m->sp[ns];
  double p3[3] = {0, 0, 0};
  for (const Particle& p : sp.part_list)
    for (int c = 0; c < 3; c++) p3[c] += (p.vel[c]) * m->vel_norm;
  for (int c = 0; c < 3; c++) p3[c] *= (sp.spwt * sp.mass);
  double Thp = sqrt(nsp.temp * eV / ME) * nsp.spwt * nsp.numparticle * ME;
  for (int c = 0; c < 3; c++) out3[c] = p3[c] / Thp;
}

// Domain::Compute_PE, domain.cpp:266-286.
double orc_pe(void* h, int ns) {
  Sim* m = (Sim*)h;
  Sim::Sp& nsp = m->sp[ns];
  double pe = 0;
  double unnorm = (m->density * QE * m->L) / EPS_0;
  for (int i = 0; i < m->nx; i++)
    for (int j = 0; j < m->ny; j++) {
      double ef2 = m->efx[m->idx(i, j)] * m->efx[m->idx(i, j)] + m->efy[m->idx(i, j)] * m->efy[m->idx(i, j)];
      pe += ef2 * (m->L * m->L);
    }
  double pe_norm = 0.5 * EPS_0 * pe * unnorm * unnorm;
  double Th = (nsp.temp * eV) * nsp.spwt * nsp.numparticle;
  return pe_norm / Th;
}

// compute_elastic_CS / compute_excitation_CS / compute_ionization_CS,
// cross_section.cpp:4-39 (the enable flags are applied by the caller).
double orc_sigma_elastic(double energy) {
  return (1e-20 * fabs(6.0 / pow(1.0 + (energy / 0.1) + pow(energy / 0.6, 2.0), 3.3) -
                       1.1 * pow(energy, 1.4) / (1.0 + pow(energy / 15.0, 1.2)) /
                           sqrt(1.0 + pow(energy / 5.5, 2.5) + pow(energy / 60.0, 4.1)) +
                       0.05 / pow(1.0 + energy / 10.0, 2.0) +
                       0.01 * pow(energy, 3.0) / (1.0 + pow(energy / 12.0, 6.0))));
}
double orc_sigma_excitation(double energy) {
  if (energy > E_EXC_TH)
    return (1e-20 * (0.034 * pow(energy - E_EXC_TH, 1.1) * (1.0 + pow(energy / 15.0, 2.8)) /
                         (1.0 + pow(energy / 23.0, 5.5)) +
                     0.023 * (energy - E_EXC_TH) / pow(1.0 + energy / 80.0, 1.9)));
  return 0.0;
}
double orc_sigma_ionization(double energy) {
  if (energy > E_ION_TH)
    return (1e-20 * (970.0 * (energy - E_ION_TH) / pow(70.0 + energy, 2.0) +
                     0.06 * pow(energy - E_ION_TH, 2.0) * exp(-energy / 9)));
  return 0.0;
}

// CollisionHandler::set_electron_cross_sections + calc_total_cross_sections,
// collision.cpp:21-42: 5,000,000-point linear tables, energy[0] = DE_CS,
// energy[i] = i*DE_CS.
void orc_mcc_build_tables(void* h) {
  Sim* m = (Sim*)h;
  m->sigma_ela.assign(CS_RANGES, 0); m->sigma_exc.assign(CS_RANGES, 0);
  m->sigma_ionz.assign(CS_RANGES, 0); m->sigma_tot_e.assign(CS_RANGES, 0);
  for (int i = 0; i < CS_RANGES; i++) {
    double e = (i == 0) ? DE_CS : DE_CS * i;
    m->sigma_ela[i] = m->enable_elastic ? orc_sigma_elastic(e) : 0.0;
    m->sigma_exc[i] = m->enable_excitation ? orc_sigma_excitation(e) : 0.0;
    m->sigma_ionz[i] = m->enable_ionization ? orc_sigma_ionization(e) : 0.0;
    m->sigma_tot_e[i] = (m->sigma_ela[i] + m->sigma_exc[i] + m->sigma_ionz[i]) * m->GAS_DENSITY;
  }
}

// CollisionHandler::collision_electron, collision.cpp:60-188.  Counts the
// process chosen in counts[0..2] (elastic, excitation, ionisation).
static void collision_electron(Sim* m, double xe, double ye, double& vxe, double& vye, double& vze,
                               int eindex, int s1, int s2, long* counts) {
  Sim::Sp& sp2 = m->sp[s2];
  const double F1 = ME / (ME + sp2.mass);
  const double F2 = sp2.mass / (ME + sp2.mass);
  double t0, t1, t2, g, g2, gx, gy, gz, wx, wy, wz, theta, phi;
  double chi, eta, chi2, eta2, sc, cc, se, ce, st, ct, sp, cp, energy, e_sc, e_ej;
  gx = vxe; gy = vye; gz = vze;
  g = sqrt(gx * gx + gy * gy + gz * gz);
  double g_before = g;
  wx = F1 * vxe; wy = F1 * vye; wz = F1 * vze;
  if (gx == 0) theta = 0.5 * PI;
  else theta = atan2(sqrt(gy * gy + gz * gz), gx);
  if (gy == 0) phi = (gz > 0) ? 0.5 * PI : -0.5 * PI;
  else phi = atan2(gz, gy);
  st = sin(theta); ct = cos(theta); sp = sin(phi); cp = cos(phi);
  t0 = m->sigma_ela[eindex];
  t1 = t0 + m->sigma_exc[eindex];
  t2 = t1 + m->sigma_ionz[eindex];
  if (m->rnd() < (t0 / t2)) {
    chi = acos(1.0 - 2.0 * m->rnd());
    eta = 2 * PI * m->rnd();
    counts[0]++;
  } else if (m->rnd() < (t1 / t2)) {
    energy = 0.5 * ME * g * g * m->vel_norm * m->vel_norm;
    energy = fabs(energy - E_EXC_TH * eV);
    g = sqrt(2.0 * energy / ME);
    g = g / m->vel_norm;
    chi = acos(1.0 - 2.0 * m->rnd());
    eta = 2 * PI * m->rnd();
    counts[1]++;
  } else {
    energy = 0.5 * ME * g * g * m->vel_norm * m->vel_norm;
    energy = fabs(energy - E_ION_TH * eV);
    e_ej = 10.0 * tan(m->rnd() * atan(energy / eV / 20.0)) * eV;
    e_sc = fabs(energy - e_ej);
    g = sqrt(2.0 * e_sc / ME);
    g2 = sqrt(2.0 * e_ej / ME);
    g = g / m->vel_norm;
    g2 = g2 / m->vel_norm;
    chi = acos(sqrt(e_sc / energy));
    chi2 = acos(sqrt(e_ej / energy));
    eta = 2 * PI * m->rnd();
    eta2 = eta + PI;
    sc = sin(chi2); cc = cos(chi2); se = sin(eta2); ce = cos(eta2);
    gx = g2 * (ct * cc - st * sc * ce);
    gy = g2 * (st * cp * cc + ct * cp * sc * ce - sp * sc * se);
    gz = g2 * (st * sp * cc + ct * sp * sc * ce + cp * sc * se);
    Particle ne;
    ne.pos[0] = xe; ne.pos[1] = ye;
    ne.vel[0] = wx + F2 * gx; ne.vel[1] = wy + F2 * gy; ne.vel[2] = wz + F2 * gz;
    m->sp[s1].part_list.push_back(ne);
    Particle ni;
    ni.pos[0] = xe; ni.pos[1] = ye;
    // collision.cpp:164: SampleVel in m/s, not divided by vel_norm.
    double a = sample_vel(m, sp2.mass, sp2.temp);
    double b = sample_vel(m, sp2.mass, sp2.temp);
    double c = sample_vel(m, sp2.mass, sp2.temp);
    ni.vel[0] = a; ni.vel[1] = b; ni.vel[2] = c;
    m->sp[s2].part_list.push_back(ni);
    counts[2]++;
  }
  sc = sin(chi); cc = cos(chi); se = sin(eta); ce = cos(eta);
  gx = g * (ct * cc - st * sc * ce);
  gy = g * (st * cp * cc + ct * cp * sc * ce - sp * sc * se);
  gz = g * (st * sp * cc + ct * sp * sc * ce + cp * sc * se);
  vxe = wx + F2 * gx; vye = wy + F2 * gy; vze = wz + F2 * gz;
  double g_after = sqrt(vxe * vxe + vye * vye + vze * vze);
  m->delta_g = fabs(g_after - g_before);
}

// CollisionHandler::handle_collisions, collision.cpp:190-207.  The reference
// iterates with a range-for while ionisation appends to the same vector
// (undefined behaviour); this restatement visits the electrons present when
// the call starts, by index.  counts[3] returns the energy (eV) removed by
// excitation and ionisation thresholds.
void orc_collide_direct(void* h, int se, int st, long* counts, double* eloss_ev) {
  Sim* m = (Sim*)h;
  size_t n0 = m->sp[se].part_list.size();
  double lost = 0;
  for (size_t k = 0; k < n0; k++) {
    Particle& part = m->sp[se].part_list[k];
    double v_sqr = (part.vel[0] * part.vel[0] + part.vel[1] * part.vel[1] + part.vel[2] * part.vel[2]) *
                   m->vel_norm * m->vel_norm;
    double velocity = sqrt(v_sqr);
    double energy = (0.5 * ME * v_sqr) / eV;
    int energy_index = std::min((int)(energy / DE_CS + 0.5), CS_RANGES - 1);
    double nu = m->sigma_tot_e[energy_index] * velocity;
    double p_coll = 1 - exp(-nu * (m->DT / m->W));
    if (p_coll > m->rnd()) {
      long before[3] = {counts[0], counts[1], counts[2]};
      double px = part.pos[0], py = part.pos[1];
      double vx = part.vel[0], vy = part.vel[1], vz = part.vel[2];
      collision_electron(m, px, py, vx, vy, vz, energy_index, se, st, counts);
      Particle& again = m->sp[se].part_list[k];  // the vector may have grown
      again.vel[0] = vx; again.vel[1] = vy; again.vel[2] = vz;
      if (counts[1] > before[1]) lost += E_EXC_TH;
      if (counts[2] > before[2]) lost += E_ION_TH;
    }
  }
  if (eloss_ev) *eloss_ev = lost;
}

// The grids main.cpp:287-323 builds (same fields as the device library's
// conductor descriptor).
struct orc_conductor {
  int kind;
  int x_center, y_center, electrode_number;
  double grid_radius, electrode_radius;
  int min_x, max_x, min_y, max_y;
  double voltage;
};

void orc_set_conductors(void* h, const orc_conductor* c, int n) {
  Sim* m = (Sim*)h;
  m->grids.clear();
  for (int k = 0; k < n; k++) {
    Sim::Grid g;
    g.kind = c[k].kind;
    g.xc = c[k].x_center; g.yc = c[k].y_center; g.ne = c[k].electrode_number;
    g.grad = c[k].grid_radius; g.erad = c[k].electrode_radius;
    g.min_x = c[k].min_x; g.max_x = c[k].max_x; g.min_y = c[k].min_y; g.max_y = c[k].max_y;
    g.voltage = c[k].voltage;
    m->grids.push_back(g);
  }
}

int orc_is_conductor(void* h, double x, double y) { return on_any_grid((Sim*)h, x, y) ? 1 : 0; }

// n draws of Init::SampleVel(species, temp) (init.cpp:110-114) from the
// global stream, in m/s.
void orc_sample_vel(void* h, double mass, double temp, int n, double* out) {
  Sim* m = (Sim*)h;
  for (int k = 0; k < n; k++) out[k] = sample_vel(m, mass, temp);
}

// One call of collision_electron (collision.cpp:60-188) on the electron
// velocity v3 (normalised), at table index eindex, position (x, y).  v3 is
// updated in place; births are appended to species se and st as the
// reference appends them.  Returns the process (0 elastic, 1 excitation,
// 2 ionisation).
int orc_collide_one(void* h, int se, int st, int eindex, double x, double y, double* v3) {
  Sim* m = (Sim*)h;
  long counts[3] = {0, 0, 0};
  collision_electron(m, x, y, v3[0], v3[1], v3[2], eindex, se, st, counts);
  return counts[0] ? 0 : (counts[1] ? 1 : 2);
}

// The emitter descriptor (emitter.h:8-20), laid out as the device
// library's epic_emitter_desc.
struct orc_emitter {
  double x0, y0, x1, y1, temp;
  int numparticle;
  double vdx, vdy;
  int species_idx1, species_idx2;
};

// Emitter::inject, emitter.cpp:5-36.  Positions: Domain::unirand
// (domain.cpp:288-299) draws from a generator seeded by std::random_device
// on every call, which no run can reproduce; here they come from one
// mt19937(pos_seed) through the same uniform_real_distribution form
// (x in [x0, x1), then y in [y0, y1)).  Velocities: SampleVel on the global
// stream (vx, then vy, per target species), plus the drift times vel_norm,
// normalised by vel_norm; vz = 0.
void orc_inject(void* h, const orc_emitter* e, unsigned pos_seed) {
  Sim* m = (Sim*)h;
  std::mt19937 pg(pos_seed);
  for (int p = 0; p < e->numparticle; p++) {
    double x = std::uniform_real_distribution<double>(e->x0, e->x1)(pg);
    double y = std::uniform_real_distribution<double>(e->y0, e->y1)(pg);
    int targets[2] = {e->species_idx1, e->species_idx2};
    int nt = e->species_idx1 != e->species_idx2 ? 2 : 1;
    for (int w = 0; w < nt; w++) {
      Sim::Sp& sp = m->sp[targets[w]];
      double vx = sample_vel(m, sp.mass, e->temp) + e->vdx * m->vel_norm;
      double vy = sample_vel(m, sp.mass, e->temp) + e->vdy * m->vel_norm;
      double vz = 0.0;
      vx /= m->vel_norm;
      vy /= m->vel_norm;
      vz /= m->vel_norm;
      Particle q;
      q.pos[0] = x; q.pos[1] = y;
      q.vel[0] = vx; q.vel[1] = vy; q.vel[2] = vz;
      sp.part_list.push_back(q);
    }
  }
}

int orc_last_iters(void* h) { return ((Sim*)h)->last_iters; }

// Host threads of the all-cores timing variant (1: the sequential oracle).
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

void orc_set_tolerance(void* h, double tol) { ((Sim*)h)->tolerance = tol; }

}  // extern "C"
