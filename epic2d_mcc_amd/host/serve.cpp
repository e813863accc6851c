// Request loop over the batched ensemble (include/epic2d_serve.h).
//
// Jobs arrive as a Poisson process; each is admitted first-come first-served
// when the device-memory budget and the member cap allow, loaded on the
// device (seeded Philox particles, sorted by cell), and added to the batch
// with its initial field pending, so that it joins the very next batched
// step.  Every loop iteration advances all admitted jobs by one step; jobs
// that reach their step count are digested (kinetic energies, ||phi||,
// collision count), removed, and their memory returned to the budget.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "epic2d_serve.h"

namespace {

std::string g_err;

// include/domain.h:204-215
constexpr double EPS_0 = 8.85418782e-12;
constexpr double QE = 1.602176565e-19;
constexpr double ME = 9.10938215e-31;
constexpr double K_b = 1.380648e-23;
constexpr double EV_to_K = QE / K_b;
constexpr double M_ION = 1836.0 * ME;
constexpr double M_AR = 6.63352090e-26;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Uniform [0, 1) from 53 bits of a 64-bit Mersenne twister.
double u01(std::mt19937_64& r) { return (double)(r() >> 11) * (1.0 / 9007199254740992.0); }

struct Species {
  const char* name;
  double mass;
  int64_t n;
  double temp;
  int sign;
  double frac, vsx;
};

}  // namespace

extern "C" {

const char* epic_serve_last_error(void) { return g_err.c_str(); }

int epic_serve_make_jobs(const epic_serve_config* c, epic_serve_job* out, int cap, int* n) {
  if (!c || !n || c->rate_hz <= 0 || c->duration_s < 0 || c->n_grids < 1 || c->n_grids > EPIC_SERVE_MAX_GRIDS)
    return fail(EPIC_ERR_ARG, "bad stream configuration");
  if (c->ppc_min < 1 || c->ppc_max < c->ppc_min || c->steps_min < 1 || c->steps_max < c->steps_min)
    return fail(EPIC_ERR_ARG, "bad job ranges");
  std::mt19937_64 r(c->seed);
  double t = 0;
  int k = 0;
  for (;;) {
    t += -std::log(1.0 - u01(r)) / c->rate_hz;
    if (t >= c->duration_s) break;
    if (c->max_jobs > 0 && k >= c->max_jobs) break;
    epic_serve_job j{};
    j.index = k;
    j.arrival_s = t;
    j.grid = c->grid_sizes[std::min((int)(u01(r) * c->n_grids), c->n_grids - 1)];
    j.ppc = c->ppc_min + std::min((int)(u01(r) * (c->ppc_max - c->ppc_min + 1)), c->ppc_max - c->ppc_min);
    j.steps = c->steps_min + std::min((int)(u01(r) * (c->steps_max - c->steps_min + 1)), c->steps_max - c->steps_min);
    j.te = c->te_min + u01(r) * (c->te_max - c->te_min);
    j.ti = c->ti_min + u01(r) * (c->ti_max - c->ti_min);
    // a collisional job has one electron species (the MCC target list of
    // handle_collisions, collision.cpp:190-207): it is never two-stream
    j.collisional = u01(r) < c->collisional_fraction ? 1 : 0;
    const double u_ts = u01(r), u_vd = u01(r);
    j.two_stream = (!j.collisional && u_ts < c->two_stream_fraction) ? 1 : 0;
    j.vd = j.two_stream ? u_vd * c->vd_max : 0.0;
    if (c->fault_job == k + 1) j.ti = std::nan("");
    j.particles = 2ll * j.ppc * j.grid * j.grid;
    if (out && k < cap) out[k] = j;
    k++;
  }
  *n = k;
  return EPIC_OK;
}

// main.cpp:151-261 with norm_scheme 1 (L = LDe, W = omega_pe), x0 = y0 = 0,
// periodic, CIC; species in the reference's order (electrons first, the
// ion second).
int epic_serve_job_setup(const epic_serve_job* j, const epic_serve_config* c, int precision, int solver,
                         epic_domain_desc* dom, epic_species_desc sp[3], int* nsp) {
  if (!j || !c || !dom || !sp || !nsp) return fail(EPIC_ERR_ARG, "null argument");
  const int64_t n = (int64_t)j->ppc * j->grid * j->grid;
  const double mi = j->collisional ? M_AR : M_ION;
  std::vector<Species> s;
  if (j->two_stream) {
    s.push_back({"beam_plus", ME, n / 2, j->te, -1, 1.0, j->vd});
    s.push_back({"ion", mi, n, j->ti, 1, 0.0, 0.0});
    s.push_back({"beam_minus", ME, n - n / 2, j->te, -1, 1.0, -j->vd});
  } else {
    s.push_back({"electron", ME, n, j->te, -1, 1.0, 0.0});
    s.push_back({"ion", mi, n, j->ti, 1, 0.0, 0.0});
  }
  double k = 0;
  for (const Species& q : s) k += (-q.sign) * q.frac;
  std::vector<double> normden(s.size());
  normden[1] = c->density;
  normden[0] = c->density / k;
  for (size_t i = 2; i < s.size(); i++) normden[i] = s[i].frac * normden[0];
  const double LDe = std::sqrt((EPS_0 * K_b * s[0].temp * EV_to_K) / (normden[0] * QE * QE));
  const double wpe = std::sqrt((normden[0] * QE * QE) / (s[0].mass * EPS_0));
  const double dx = LDe / LDe, dy = LDe / LDe;
  double DT = c->dt_coeff * (1.0 / wpe);
  DT = wpe * DT;
  epic_domain_desc d{};
  d.x0 = 0;
  d.y0 = 0;
  d.dx = dx;
  d.dy = dy;
  d.nx = j->grid + 1;
  d.ny = j->grid + 1;
  d.L = LDe;
  d.W = wpe;
  d.LDe = LDe;
  d.wpe = wpe;
  d.vel_norm = d.L * d.W;
  d.density = c->density;
  d.DT = DT;
  d.bc = EPIC_BC_PBC;
  d.shape = EPIC_SHAPE_CIC;
  d.solver = solver;
  d.poisson_op = EPIC_OP_REF_GHOST;
  d.tolerance = 1e-10;
  d.tol_relative = 1;
  d.max_iteration = 20000;
  d.sor_omega = 1.4;
  d.precision = precision;
  d.sort_interval = c->sort_interval;
  d.seed = c->seed * 1000003ull + (uint64_t)j->index;
  d.precond = EPIC_PRECOND_SPECTRAL;
  const double Lx = (d.nx - 1) * dx, Ly = (d.ny - 1) * dy;
  // ionisation births need room in both species
  const double capf = j->collisional ? 1.25 : 1.0;
  for (size_t i = 0; i < s.size(); i++) {
    epic_species_desc e{};
    std::snprintf(e.name, sizeof(e.name), "%s", s[i].name);
    e.mass = s[i].mass;
    e.charge = s[i].sign * QE;
    e.spwt = normden[i] * Lx * Ly * d.L * d.L / (double)s[i].n;
    e.temp = s[i].temp;
    e.numparticle = s[i].n;
    e.capacity = (int64_t)std::ceil((double)s[i].n * capf);
    e.vsx = s[i].vsx;
    e.vsy = 0;
    e.fract_den = s[i].frac;
    e.pushed = 1;
    sp[i] = e;
  }
  *dom = d;
  *nsp = (int)s.size();
  return EPIC_OK;
}

}  // extern "C"

namespace {

struct Active {
  int job;
  int id;
  int64_t bytes;
  epic_ctx* ctx;
};

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double pos = p * (double)(v.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

// Device bytes a job holds while admitted: its context, MCC tables and its
// share of the batch's FP32 transform workspace (X and Y, P*Q each: the
// transform dimensions padded to multiples of 32).
int64_t job_bytes(const epic_domain_desc& d, const epic_species_desc* sp, int nsp, const epic_serve_job& j,
                  const epic_serve_config* c) {
  int64_t b = 0;
  if (epic_estimate_bytes(&d, sp, nsp, &b) != EPIC_OK) return -1;
  if (j.collisional && c->mcc_template) {
    epic_mcc_desc m = *c->mcc_template;
    m.electron_sp = 0;
    m.ion_sp = 1;
    int64_t mb = 0;
    if (epic_mcc_estimate_bytes(&m, sp, nsp, &mb) != EPIC_OK) return -1;
    b += mb;
  }
  auto pad32 = [](int64_t p) { return (p + 31) / 32 * 32; };
  b += 2ll * pad32(d.nx) * pad32(d.ny) * 4 * 3 / 2;  // X, Y (the workspace grows by 1.5x)
  return b;
}

// Loads, sorts and adds a job to the batch; its field starts in the next step.
int admit(epic_batch* B, const epic_serve_config* c, epic_serve_job& j, Active* a) {
  epic_domain_desc d;
  epic_species_desc sp[3];
  int nsp = 0;
  int rc = epic_serve_job_setup(&j, c, EPIC_FP32, EPIC_SOLVER_PCG_MIXED, &d, sp, &nsp);
  if (rc != EPIC_OK) return rc;
  int id = -1;
  if ((rc = epic_batch_add(B, &d, sp, nsp, &id)) != EPIC_OK)
    return fail(rc, std::string("batch add: ") + epic_batch_last_error(B));
  epic_ctx* ctx = nullptr;
  epic_batch_member(B, id, &ctx);
  for (int s = 0; s < nsp; s++) {
    if ((rc = epic_load_species(ctx, s, sp[s].numparticle, 0)) != EPIC_OK) return fail(rc, epic_last_error(ctx));
    if ((rc = epic_sort(ctx, s)) != EPIC_OK) return fail(rc, epic_last_error(ctx));
  }
  if (j.collisional && c->mcc_template) {
    epic_mcc_desc m = *c->mcc_template;
    m.electron_sp = 0;
    m.ion_sp = 1;
    m.gas_density = c->gas_density;
    m.gas_mass = M_AR;
    if ((rc = epic_mcc_setup(ctx, &m)) != EPIC_OK) return fail(rc, epic_last_error(ctx));
  }
  if ((rc = epic_batch_init(B, id)) != EPIC_OK) return fail(rc, epic_batch_last_error(B));
  a->id = id;
  a->ctx = ctx;
  return EPIC_OK;
}

// Digest of a finished job: per-species kinetic energy, ||phi||_2, MCC count.
void digest(epic_ctx* ctx, int nsp, int nn, epic_serve_job& j) {
  for (int s = 0; s < 3; s++) j.ke[s] = 0;
  for (int s = 0; s < nsp && s < 3; s++) {
    double ke[3];
    if (epic_compute_ke(ctx, s, 0, ke) == EPIC_OK) j.ke[s] = ke[0] + ke[1] + ke[2];
  }
  std::vector<double> phi((size_t)nn);
  double s2 = 0;
  if (epic_get_field(ctx, 1, phi.data()) == EPIC_OK)
    for (double v : phi) s2 += v * v;
  j.phi_l2 = std::sqrt(s2);
  epic_mcc_stats st{};
  j.collisions = 0;
  if (j.collisional && epic_mcc_get_stats(ctx, &st) == EPIC_OK)
    for (int p = 0; p < EPIC_MCC_NPROC; p++) j.collisions += st.count[p];
}

using Clock = std::chrono::steady_clock;

// Health of an admitted job from its error counters and last residual: 0,
// or 1 + the EPIC_ERRC_* index that fails it.  A member's NOCONV count
// alone (a step that ended just above the solver tolerance) is recorded
// but does not fail it; its residual has to exceed fail_relres.
int job_health(epic_batch* B, const Active& a, const epic_serve_config* c, epic_serve_job& j) {
  int64_t e[EPIC_NERR] = {};
  if (epic_error_counters(a.ctx, e) != EPIC_OK) return 1 + EPIC_ERRC_NAN;
  for (int k = 0; k < EPIC_NERR; k++) j.errors[k] = e[k];
  double rr = 0;
  if (epic_batch_relres(B, a.id, &rr) == EPIC_OK) j.relres = rr;
  static const int fatal[] = {EPIC_ERRC_NAN, EPIC_ERRC_CAPACITY, EPIC_ERRC_FIXED_OVF, EPIC_ERRC_GATHER_OOB,
                              EPIC_ERRC_DEPOSIT_OOB};
  for (int k : fatal)
    if (e[k]) return 1 + k;
  const double lim = c->fail_relres > 0 ? c->fail_relres : 1e-6;
  if (!(j.relres <= lim)) return 1 + EPIC_ERRC_NOCONV;
  return 0;
}

}  // namespace

extern "C" int epic_serve_run(const epic_serve_config* c, epic_serve_report* rep, epic_serve_job* jobs_out,
                              int cap) {
  if (!c || !rep) return fail(EPIC_ERR_ARG, "null argument");
  int n = 0;
  int rc = epic_serve_make_jobs(c, nullptr, 0, &n);
  if (rc != EPIC_OK) return rc;
  std::vector<epic_serve_job> jobs((size_t)n);
  if ((rc = epic_serve_make_jobs(c, jobs.data(), n, &n)) != EPIC_OK) return rc;
  std::vector<int64_t> bytes((size_t)n);
  std::vector<int> nsp((size_t)n), nn((size_t)n);
  for (int k = 0; k < n; k++) {
    epic_domain_desc d;
    epic_species_desc sp[3];
    if ((rc = epic_serve_job_setup(&jobs[k], c, EPIC_FP32, EPIC_SOLVER_PCG_MIXED, &d, sp, &nsp[k])) != EPIC_OK)
      return rc;
    bytes[k] = job_bytes(d, sp, nsp[k], jobs[k], c);
    nn[k] = d.nx * d.ny;
    jobs[k].bytes = bytes[k];
    jobs[k].particles = 0;
    for (int s = 0; s < nsp[k]; s++) jobs[k].particles += sp[s].numparticle;
    if (bytes[k] < 0 || bytes[k] > c->budget_bytes) return fail(EPIC_ERR_CAPACITY, "a job exceeds the budget");
  }
  epic_batch_opts o{};
  o.refine = c->refine;
  o.use_graph = c->use_graph;
  o.push_variant = -1;
  o.node_block = 0;
  epic_batch* B = nullptr;
  if ((rc = epic_batch_create(&o, &B)) != EPIC_OK) return fail(rc, "batch create");
  auto done_batch = [&](int code) {
    epic_batch_destroy(B);
    return code;
  };
  // kernel shapes from the first jobs, tuned before the clock starts
  if (c->tune_steps > 0 && n > 0) {
    epic_batch* T = nullptr;
    if ((rc = epic_batch_create(&o, &T)) != EPIC_OK) return done_batch(fail(rc, "tuning batch"));
    for (int k = 0; k < std::min(n, 8); k++) {
      Active a{};
      if ((rc = admit(T, c, jobs[k], &a)) != EPIC_OK) {
        epic_batch_destroy(T);
        return done_batch(rc);
      }
    }
    epic_batch_tuning t{};
    rc = epic_batch_tune(T, c->tune_steps, &t);
    epic_batch_destroy(T);
    if (rc != EPIC_OK) return done_batch(fail(rc, "tuning"));
    o.push_variant = t.push_variant;
    o.node_block = t.node_block;
    epic_batch_configure(B, &o);
  }

  std::deque<int> queue;
  std::vector<Active> active;
  int next = 0, done = 0, failed = 0, max_members = 0, since_check = 0;
  const int interval = c->health_interval > 0 ? c->health_interval : 10;
  int64_t used = 0;
  double peak = 0, busy = 0, last_print = 0;
  const auto t0 = Clock::now();
  auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  while (done + failed < n) {
    double t = now();
    while (next < n && jobs[next].arrival_s <= t) queue.push_back(next++);
    while (!queue.empty()) {
      const int k = queue.front();
      if ((int)active.size() >= c->max_members || used + bytes[k] > (int64_t)c->budget_bytes) break;
      Active a{};
      a.job = k;
      a.bytes = bytes[k];
      if ((rc = admit(B, c, jobs[k], &a)) != EPIC_OK) return done_batch(rc);
      jobs[k].admit_s = now();
      used += a.bytes;
      peak = std::max(peak, (double)used);
      active.push_back(a);
      queue.pop_front();
    }
    if (active.empty()) {
      if (next < n) {
        const double wait = jobs[next].arrival_s - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      }
      continue;
    }
    max_members = std::max(max_members, (int)active.size());
    const double ts = now();
    if ((rc = epic_batch_step(B, 1)) != EPIC_OK) return done_batch(fail(rc, epic_batch_last_error(B)));
    // which jobs finish, and (every interval steps, and for the finishing
    // ones) which have failed; read before any member leaves the batch
    const bool check = ++since_check >= interval;
    if (check) since_check = 0;
    std::vector<int> verdict(active.size(), 0);  // 0 stays, 1 completes, 2 evicted
    for (size_t m = 0; m < active.size(); m++) {
      int64_t st = 0;
      epic_batch_member_steps(B, active[m].id, &st);
      epic_serve_job& j = jobs[active[m].job];
      const bool fin = st >= j.steps;
      if (check || fin) {
        j.failed = job_health(B, active[m], c, j);
        if (j.failed) verdict[m] = 2;
        else if (fin) verdict[m] = 1;
      }
    }
    for (size_t m = active.size(); m-- > 0;) {
      if (!verdict[m]) continue;
      epic_serve_job& j = jobs[active[m].job];
      if (verdict[m] == 1) {
        digest(active[m].ctx, nsp[active[m].job], nn[active[m].job], j);  // synchronises
        done++;
      } else {
        failed++;
        std::fprintf(stderr, "[serve] job %d evicted (error counter %d) after %.1f s\n", j.index, j.failed - 1,
                     now() - j.arrival_s);
      }
      j.done_s = now();
      epic_batch_remove(B, active[m].id);
      used -= active[m].bytes;
      active.erase(active.begin() + m);
    }
    busy += now() - ts;
    if (c->progress_s > 0 && now() - last_print >= c->progress_s) {
      last_print = now();
      std::fprintf(stderr, "[serve] t=%.1fs arrived %d admitted %d done %d active %d queued %d used %.2f GB\n",
                   last_print, next, next - (int)queue.size(), done, (int)active.size(), (int)queue.size(),
                   (double)used / 1e9);
      std::fflush(stderr);
    }
  }
  epic_batch_sync(B);
  epic_batch_stats bs{};
  epic_batch_get_stats(B, &bs);
  std::vector<double> lat, wait;
  double ps = 0, first = n ? jobs[0].arrival_s : 0, last = 0;
  for (const epic_serve_job& j : jobs) {
    last = std::max(last, j.done_s);
    if (j.failed) continue;
    lat.push_back(j.done_s - j.arrival_s);
    wait.push_back(j.admit_s - j.arrival_s);
    ps += (double)j.particles * j.steps;
  }
  epic_serve_report r{};
  r.jobs = n;
  r.jobs_done = done;
  r.wall_s = last - first;
  r.busy_s = busy;
  r.latency_p50_s = pct(lat, 0.5);
  r.latency_p99_s = pct(lat, 0.99);
  double sum = 0, mx = 0;
  for (double v : lat) {
    sum += v;
    mx = std::max(mx, v);
  }
  r.latency_mean_s = lat.empty() ? 0 : sum / (double)lat.size();
  r.latency_max_s = mx;
  r.wait_p50_s = pct(wait, 0.5);
  r.wait_p99_s = pct(wait, 0.99);
  r.particle_steps = ps;
  r.throughput = r.wall_s > 0 ? ps / r.wall_s : 0;
  r.max_members = max_members;
  r.peak_bytes = peak;
  r.batch_steps = bs.steps;
  r.graph_captures = bs.graph_captures;
  r.push_variant = bs.push_variant;
  r.node_block = bs.node_block;
  r.jobs_failed = failed;
  *rep = r;
  if (jobs_out)
    for (int k = 0; k < std::min(n, cap); k++) jobs_out[k] = jobs[k];
  return done_batch(EPIC_OK);
}
