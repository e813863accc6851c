// extern "C" boundary of libepic2d: context lifetime, particle upload and
// download, the per-step operators and diagnostics.  Status codes only; no
// exception crosses the boundary.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "epic_internal.hpp"

struct epic_ctx {
  epic::Ctx c;
};

namespace epic {

static int fail(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

static int hip_fail(Ctx* c, hipError_t e, const char* where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  return fail(c, EPIC_ERR_HIP, m);
}

#define CK(expr, where)                                       \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return hip_fail(c, _e, where);      \
  } while (0)

template <typename T>
static hipError_t dalloc(T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemset(*p, 0, count * sizeof(T));
}

template <typename T>
static hipError_t dalloc_c(Ctx* c, T** p, size_t count) {
  if (count == 0) count = 1;
  c->bytes += count * sizeof(T);
  return dalloc(p, count);
}

static size_t psize(const Ctx* c) { return c->dom.precision == EPIC_FP64 ? 8 : 4; }

hipError_t sync_counts(Ctx* c) {
  for (auto& S : c->sp) {
    int64_t n[2];
    EPIC_HIP_TRY(hipMemcpyAsync(n, S.d_n, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
    S.n_host = n[0];
  }
  return hipSuccess;
}

// Resolve every recorded pair into the per-stage totals.
static void resolve_events(Ctx* c) {
  if (c->ev_used == 0) return;
  hipEventSynchronize(c->ev_pool[2 * (c->ev_used - 1) + 1]);
  for (int k = 0; k < c->ev_used; k++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev_pool[2 * k], c->ev_pool[2 * k + 1]) == hipSuccess) {
      c->stage_ms[c->ev_stage[k]] += ms;
      c->stage_calls[c->ev_stage[k]]++;
    }
  }
  c->ev_used = 0;
}

void stage_begin(Ctx* c, int stage) {
  if (!c->timing || !((c->timing_mask >> stage) & 1u)) return;
  if (c->ev_used * 2 + 2 > (int)c->ev_pool.size()) resolve_events(c);
  if (c->ev_used * 2 + 2 > (int)c->ev_pool.size()) return;
  hipEventRecord(c->ev_pool[2 * c->ev_used], c->stream);
  c->ev_open = c->ev_used;
}
void stage_end(Ctx* c, int stage) {
  if (!c->timing || c->ev_open < 0) return;
  hipEventRecord(c->ev_pool[2 * c->ev_open + 1], c->stream);
  c->ev_stage[c->ev_open] = stage;
  c->ev_used = c->ev_open + 1;
  c->ev_open = -1;
}

void spec_destroy(Ctx* c);

static void free_all(Ctx* c) {
  spec_destroy(c);
  const bool tiled = c->dom.precision == EPIC_FP32;
  for (auto& S : c->sp) {
    // tiled storage: y, vx, vy point into x's allocation
    void* ps[] = {S.x, tiled ? nullptr : S.y, tiled ? nullptr : S.vx, tiled ? nullptr : S.vy, S.vz, S.id,
                  S.x2, tiled ? nullptr : S.y2, tiled ? nullptr : S.vx2, tiled ? nullptr : S.vy2, S.vz2, S.id2,
                  S.Q, S.den, S.d_n};
    for (void* p : ps)
      if (p) hipFree(p);
  }
  void* gs[] = {c->rho, c->phi, c->efx, c->efy, c->E2, c->Ec, c->b64, c->r64, c->ws_v, c->ws_w, c->ws_x,
                c->ws_part, c->ws_scalar, c->ws_iter, c->sort_count, c->sort_fill, c->sort_perm, c->blk_count,
                c->red_part, c->red_out, c->errc, c->scan_tiles, c->mcc.sigma, c->mcc.counts, c->mcc.events,
                c->mcc.event_count, c->mcc.list, c->mcc.list_n, c->mcc.chunk_ev, c->sp_S, c->sp_T, c->sp_lam,
                c->cond.table, c->cond.cmask, c->cond.node, c->cond.value, c->nr_ws};
  for (void* p : gs)
    if (p) hipFree(p);
  if (c->hscal) hipHostFree(c->hscal);
  c->hscal = nullptr;
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (hipEvent_t ev : c->ev_pool) hipEventDestroy(ev);
  c->ev_pool.clear();
  if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
}

Ctx* ctx_of(epic_ctx* h) { return h ? &h->c : nullptr; }

// Device bytes ctx_create allocates for this description (the allocations
// below, in the same order); MCC tables and the sine-transform solver's
// matrices are allocated later, on first use.
size_t ctx_bytes_estimate(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, int num_cu) {
  const size_t nn = (size_t)dom->nx * dom->ny;
  const size_t ps = dom->precision == EPIC_FP64 ? 8 : 4;
  auto b = [](size_t count, size_t sz) { return (count == 0 ? 1 : count) * sz; };
  size_t tot = 0;
  int64_t maxcap = 1;
  for (int s = 0; s < nsp; s++) {
    const int64_t cap = sp[s].capacity > 0 ? sp[s].capacity : std::max<int64_t>(sp[s].numparticle, 1);
    maxcap = std::max(maxcap, cap);
    const size_t cap_pad = (size_t)((cap + PTILE - 1) / PTILE * PTILE);
    tot += 10 * b(cap_pad * ps, 1) + 2 * b(cap_pad, 4) + 2 * b(nn, 8) + b(2, 8);
  }
  tot += 4 * b(nn, 8) + b(2 * nn * ps, 1);
  if (dom->precision == EPIC_FP32) tot += b((size_t)(dom->nx - 1) * (dom->ny - 1) * 8 * sizeof(float), 1);
  tot += 2 * b(nn, 8) + b(2 * nn, 8) + 2 * b(nn, 8);
  tot += b((size_t)16 * std::max(num_cu, 1024), 8) + b(8, 8) + b(4, 4);
  const size_t ncell = (size_t)sort_key_space(dom->nx, dom->ny);
  tot += 2 * b(ncell + 2, 4) + b(2 * (size_t)maxcap, 4);
  const int64_t blk_cap = maxcap / 2048 + 4;
  tot += b((size_t)blk_cap, 4);
  const int scan_cap = (int)(std::max<int64_t>((int64_t)ncell + 2, blk_cap) / 1024 + 4);
  tot += b((size_t)scan_cap, 8);
  tot += b((size_t)8 * num_cu * 4, 8) + b(8, 8) + b(EPIC_NERR, 8);
  return tot;
}

}  // namespace epic

using namespace epic;

extern "C" {

int epic_device_name(char* buf, int len) {
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return EPIC_ERR_HIP;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return EPIC_ERR_HIP;
  snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return EPIC_OK;
}

}  // extern "C"

namespace epic {

int ctx_create(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, hipStream_t ext, epic_ctx** out) {
  if (!dom || !sp || !out || nsp < 1 || nsp > 16) return EPIC_ERR_ARG;
  if (dom->nx < 3 || dom->ny < 3) return EPIC_ERR_ARG;
  if (dom->precision != EPIC_FP64 && dom->precision != EPIC_FP32) return EPIC_ERR_ARG;
  if (dom->solver < EPIC_SOLVER_PCG || dom->solver > EPIC_SOLVER_NEWTON_CG) return EPIC_ERR_ARG;
  if (dom->bc < EPIC_BC_PBC || dom->bc > EPIC_BC_REFLECTIVE) return EPIC_ERR_ARG;
  if (dom->shape != EPIC_SHAPE_CIC && dom->shape != EPIC_SHAPE_NGP) return EPIC_ERR_ARG;
  if ((int64_t)dom->nx * dom->ny > (1ll << 30)) return EPIC_ERR_ARG;
  if (sort_key_space(dom->nx, dom->ny) >= (1ll << 25)) return EPIC_ERR_ARG;  // sort packs key * 64 + lane
  epic_ctx* h = new (std::nothrow) epic_ctx();
  if (!h) return EPIC_ERR_CAPACITY;
  Ctx* c = &h->c;
  *out = nullptr;
  c->dom = *dom;
  if (c->dom.sor_omega == 0) c->dom.sor_omega = 1.4;
  c->nsp = nsp;
  c->fix_bits = dom->precision == EPIC_FP64 ? FIX_BITS_FP64 : FIX_BITS_FP32;
  c->nnodes = (size_t)dom->nx * dom->ny;
  hipError_t e;
#define CKC(expr, where)                          \
  do {                                            \
    e = (expr);                                   \
    if (e != hipSuccess) {                        \
      c->err = std::string(where) + ": " + hipGetErrorString(e); \
      free_all(c);                                \
      delete h;                                   \
      return EPIC_ERR_HIP;                        \
    }                                             \
  } while (0)
  CKC(hipGetDevice(&c->device), "hipGetDevice");
  hipDeviceProp_t prop;
  CKC(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
  c->num_cu = prop.multiProcessorCount;
  if (const char* v = getenv("EPIC_PUSH_VARIANT")) c->push_variant = atoi(v);
  if (const char* v = getenv("EPIC_PUSH_GRID")) c->push_grid_per_cu = atoi(v);
  if (const char* v = getenv("EPIC_PUSH_MAP")) c->push_map = atoi(v);
  if (const char* v = getenv("EPIC_PART_MAP")) c->part_map = atoi(v);
  if (const char* v = getenv("EPIC_PART_PPT")) c->part_ppt = atoi(v);
  if (const char* v = getenv("EPIC_PUSH64")) c->push64 = atoi(v);
  if (const char* v = getenv("EPIC_MCC_OCC")) c->mcc_occ = atoi(v);
  if (const char* v = getenv("EPIC_MCC_SEL_GRID")) c->mcc_sel_grid = atoi(v);
  if (const char* v = getenv("EPIC_SORT_GATHER_WG")) c->sort_gather_wg = std::max(1, atoi(v));
  if (const char* v = getenv("EPIC_PUSH_MULTI")) c->push_multi = atoi(v);
  if (const char* v = getenv("EPIC_COND_CELL_MASK")) c->cond_cell_mask = atoi(v);
  if (const char* v = getenv("EPIC_PUSH_MULTI_WAVES")) c->push_multi_waves = atoi(v);
  if (const char* v = getenv("EPIC_PCG_BLOCK")) c->pcg_block = atoi(v);
  // many-workgroup PCG / SOR: one workgroup per two CUs.  Every iteration
  // ends in grid barriers whose cost grows with the number of workgroups
  // arriving on one counter; at 513^2 nodes 128 workgroups (2-4 nodes per
  // thread) run the FP64 PCG iteration in 20 us against 28 us with 256.
  c->max_solver_blocks = std::max(1, c->num_cu / 2);
  if (const char* v = getenv("EPIC_SOLVER_GRID")) {
    const int g = atoi(v);
    if (g > 0 && g <= c->num_cu) c->max_solver_blocks = g;
  }
  if (ext) {
    c->stream = ext;
    c->own_stream = false;
  } else {
    CKC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  }
  CKC(hipEventCreate(&c->ev0), "hipEventCreate");
  CKC(hipEventCreate(&c->ev1), "hipEventCreate");
  const size_t nn = c->nnodes;
  const size_t ps = psize(c);
  int64_t maxcap = 1;
  c->sp.resize(nsp);
  // Sort schedule: the sort keeps a wavefront's particles in few cells, and
  // particles leave their cells at their thermal speed (a drift moves a
  // species' particles together and does not disorder them).  A species
  // whose thermal speed is k times below the fastest one's is sorted k times
  // less often (k up to 64), so cold ions are not re-sorted every interval.
  // Sorting changes no result.  EPIC_SORT_SCALE=0: every species every
  // interval.
  double vmax = 0;
  for (int s = 0; s < nsp; s++) vmax = std::max(vmax, std::sqrt(std::max(sp[s].temp, 1e-30) / sp[s].mass));
  const char* scale_env = getenv("EPIC_SORT_SCALE");
  const bool scale = !(scale_env && atoi(scale_env) == 0);
  for (int s = 0; s < nsp; s++) {
    SpeciesDev& S = c->sp[s];
    S.desc = sp[s];
    const double v = std::sqrt(std::max(sp[s].temp, 1e-30) / sp[s].mass);
    S.sort_mult = scale ? (int)std::min(64.0, std::max(1.0, std::floor(vmax / v))) : 1;
    S.cap = sp[s].capacity > 0 ? sp[s].capacity : std::max<int64_t>(sp[s].numparticle, 1);
    if (S.cap > (int64_t)0x7fffffff) {
      free_all(c);
      delete h;
      return EPIC_ERR_CAPACITY;
    }
    maxcap = std::max(maxcap, S.cap);
    // pad to whole tiles (FP32) or 32-byte runs (FP64) so vector loads of
    // the last run stay inside
    const size_t cap_pad = (size_t)((S.cap + PTILE - 1) / PTILE * PTILE);
    if (dom->precision == EPIC_FP32) {
      // x, y, vx, vy in tiles of PTILE slots: one allocation per buffer
      CKC(dalloc_c(c, (float**)&S.x, 4 * cap_pad), "particle arrays");
      CKC(dalloc_c(c, (float**)&S.x2, 4 * cap_pad), "particle arrays");
      float* t[2] = {(float*)S.x, (float*)S.x2};
      S.y = t[0] + PTILE; S.vx = t[0] + 2 * PTILE; S.vy = t[0] + 3 * PTILE;
      S.y2 = t[1] + PTILE; S.vx2 = t[1] + 2 * PTILE; S.vy2 = t[1] + 3 * PTILE;
      CKC(dalloc_c(c, (char**)&S.vz, cap_pad * ps), "particle arrays");
      CKC(dalloc_c(c, (char**)&S.vz2, cap_pad * ps), "particle arrays");
    } else {
      char** arrs[] = {(char**)&S.x, (char**)&S.y, (char**)&S.vx, (char**)&S.vy, (char**)&S.vz,
                       (char**)&S.x2, (char**)&S.y2, (char**)&S.vx2, (char**)&S.vy2, (char**)&S.vz2};
      for (char** a : arrs) CKC(dalloc_c(c, a, cap_pad * ps), "particle arrays");
    }
    CKC(dalloc_c(c, &S.id, cap_pad), "ids");
    CKC(dalloc_c(c, &S.id2, cap_pad), "ids");
    CKC(dalloc_c(c, &S.Q, nn), "Q");
    CKC(dalloc_c(c, &S.den, nn), "den");
    CKC(dalloc_c(c, &S.d_n, 2), "count");
  }
  CKC(dalloc_c(c, &c->rho, nn), "rho");
  CKC(dalloc_c(c, &c->phi, nn), "phi");
  CKC(dalloc_c(c, &c->efx, nn), "efx");
  CKC(dalloc_c(c, &c->efy, nn), "efy");
  CKC(dalloc_c(c, (char**)&c->E2, 2 * nn * ps), "E2");
  if (dom->precision == EPIC_FP32)
    CKC(dalloc_c(c, (char**)&c->Ec, (size_t)(dom->nx - 1) * (dom->ny - 1) * 8 * sizeof(float)), "Ec");
  CKC(dalloc_c(c, &c->b64, nn), "b");
  CKC(dalloc_c(c, &c->r64, nn), "r");
  CKC(dalloc_c(c, (double**)&c->ws_v, 2 * nn), "v");
  CKC(dalloc_c(c, (double**)&c->ws_w, nn), "w");
  CKC(dalloc_c(c, (double**)&c->ws_x, nn), "x");
  CKC(dalloc_c(c, &c->ws_part, (size_t)16 * std::max(c->max_solver_blocks, 1024)), "partials");
  CKC(dalloc_c(c, &c->ws_scalar, 8), "scalars");
  CKC(dalloc_c(c, &c->ws_iter, 4), "iters");  // [0] iterations, [1] grid barrier
  const size_t ncell = (size_t)sort_key_space(dom->nx, dom->ny);
  CKC(dalloc_c(c, &c->sort_count, ncell + 2), "sort");
  CKC(dalloc_c(c, &c->sort_fill, ncell + 2), "sort");
  c->sort_cap = maxcap;
  CKC(dalloc_c(c, &c->sort_perm, 2 * (size_t)maxcap), "sort");
  c->blk_cap = maxcap / 2048 + 4;
  CKC(dalloc_c(c, &c->blk_count, (size_t)c->blk_cap), "compact");
  c->scan_tiles_cap = (int)(std::max<int64_t>((int64_t)ncell + 2, c->blk_cap) / 1024 + 4);
  CKC(dalloc_c(c, &c->scan_tiles, (size_t)c->scan_tiles_cap), "scan");
  c->red_blocks = c->num_cu * 4;
  CKC(dalloc_c(c, &c->red_part, (size_t)8 * c->red_blocks), "reduce");
  CKC(dalloc_c(c, &c->red_out, 8), "reduce");
  CKC(dalloc_c(c, &c->errc, EPIC_NERR), "errc");
#undef CKC
  *out = h;
  return EPIC_OK;
}

}  // namespace epic

extern "C" {

int epic_create(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, epic_ctx** out) {
  return ctx_create(dom, sp, nsp, nullptr, out);
}

int epic_estimate_bytes(const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, int64_t* bytes) {
  if (!dom || !sp || !bytes || nsp < 1 || nsp > 16) return EPIC_ERR_ARG;
  int dev = 0, ncu = 256;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    ncu = prop.multiProcessorCount;
  *bytes = (int64_t)ctx_bytes_estimate(dom, sp, nsp, ncu);
  return EPIC_OK;
}

// What epic_mcc_setup allocates: table, counters, event list, the
// real-collision lists (8 bytes per slot of each colliding species, capacity
// rounded up to MCC_SUB) with their two counts, and the electrons' event
// counts per selection range (ranges of 64 slots at least).
// Default ionisation event list: 16384 events, or one per 64 electron slots
// for large species (a hot 1e8-electron plasma ionises ~1e5 per step).
static int mcc_event_cap(int event_cap, int64_t len_e) {
  return event_cap > 0 ? event_cap
                       : (int)std::min<int64_t>(MCC_EVENT_MAX, std::max<int64_t>(16384, (len_e + 63) / 64));
}

static size_t mcc_bytes_of(int ntab, int event_cap, int64_t len, int64_t len_e) {
  const int cap = mcc_event_cap(event_cap, len_e);
  return (size_t)EPIC_MCC_NPROC * std::max(ntab, 1) * sizeof(double) + (EPIC_MCC_NPROC + 2) * sizeof(uint64_t) +
         (size_t)cap * 8 * sizeof(double) + sizeof(int) + (size_t)std::max<int64_t>(len, 1) * sizeof(uint2) +
         2 * sizeof(int) + (size_t)std::max<int64_t>(len_e / 64, 1) * sizeof(int);
}

static int64_t mcc_list_len(const epic_species_desc* sp, int s) {
  if (s < 0) return 0;
  const int64_t cap = sp[s].capacity > 0 ? sp[s].capacity : std::max<int64_t>(sp[s].numparticle, 1);
  return (cap + MCC_SUB - 1) / MCC_SUB * MCC_SUB;
}

int epic_mcc_estimate_bytes(const epic_mcc_desc* d, const epic_species_desc* sp, int nsp, int64_t* bytes) {
  if (!d || !sp || !bytes || nsp < 1 || d->electron_sp >= nsp || d->ion_sp >= nsp) return EPIC_ERR_ARG;
  const int64_t le = mcc_list_len(sp, d->electron_sp);
  *bytes = (int64_t)mcc_bytes_of(d->ntab, d->event_cap, le + mcc_list_len(sp, d->ion_sp), le);
  return EPIC_OK;
}

int epic_device_bytes(epic_ctx* h, int64_t* bytes) {
  if (!h || !bytes) return EPIC_ERR_ARG;
  *bytes = (int64_t)h->c.bytes;
  return EPIC_OK;
}

int epic_destroy(epic_ctx* h) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  hipStreamSynchronize(c->stream);
  free_all(c);
  delete h;
  return EPIC_OK;
}

const char* epic_last_error(const epic_ctx* h) { return h ? h->c.err.c_str() : "null context"; }

int epic_sync(epic_ctx* h) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  CK(hipStreamSynchronize(c->stream), "sync");
  return EPIC_OK;
}

// ---- particles ------------------------------------------------------------
int epic_upload_particles(epic_ctx* h, int s, const double* x, const double* y, const double* vx, const double* vy,
                          const double* vz, int64_t n) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp || n < 0) return fail(c, EPIC_ERR_ARG, "bad species or count");
  SpeciesDev& S = c->sp[s];
  if (n > S.cap) return fail(c, EPIC_ERR_CAPACITY, "upload exceeds species capacity");
  const double* src[5] = {x, y, vx, vy, vz};
  void* dst[5] = {S.x, S.y, S.vx, S.vy, S.vz};
  if (c->dom.precision == EPIC_FP64) {
    for (int k = 0; k < 5; k++) {
      if (!src[k]) return fail(c, EPIC_ERR_ARG, "null array");
      CK(hipMemcpy(dst[k], src[k], n * sizeof(double), hipMemcpyHostToDevice), "upload");
    }
  } else {
    // Rounding to FP32 can lift a position just below the upper wall onto
    // it (x0 + Lx), which the reference never holds for a live particle
    // (its loader and push keep x < x0 + Lx): such positions are stored as
    // the largest float below the wall.
    const double hi[2] = {c->dom.x0 + (c->dom.nx - 1) * c->dom.dx, c->dom.y0 + (c->dom.ny - 1) * c->dom.dy};
    // x, y, vx, vy into the tiled layout (whole tiles), vz as it is
    const int64_t ntile = (n + PTILE - 1) / PTILE;
    std::vector<float> tiles((size_t)ntile * 4 * PTILE, 0.f), tmp(n);
    for (int k = 0; k < 5; k++) {
      if (!src[k]) return fail(c, EPIC_ERR_ARG, "null array");
      for (int64_t i = 0; i < n; i++) tmp[i] = (float)src[k][i];
      if (k < 2) {
        const float hf = (float)hi[k], below = std::nextafter(hf, -INFINITY);
        for (int64_t i = 0; i < n; i++)
          if (src[k][i] < hi[k] && tmp[i] >= hf) tmp[i] = below;
      }
      if (k < 4) {
        for (int64_t i = 0; i < n; i++) tiles[ptile(i) + k * PTILE] = tmp[i];
      } else {
        CK(hipMemcpy(dst[k], tmp.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload");
      }
    }
    CK(hipMemcpy(S.x, tiles.data(), tiles.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
  }
  std::vector<uint32_t> ids(n);
  for (int64_t i = 0; i < n; i++) ids[i] = (uint32_t)i;
  CK(hipMemcpy(S.id, ids.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice), "upload ids");
  int64_t cnt[2] = {n, n};
  CK(hipMemcpy(S.d_n, cnt, sizeof(cnt), hipMemcpyHostToDevice), "upload count");
  S.n_host = n;
  S.maybe_dead = false;
  for (int64_t i = 0; i < n && !S.maybe_dead; i++) S.maybe_dead = std::isnan(x[i]);
  c->deposit_valid = false;
  return EPIC_OK;
}

int epic_download_particles(epic_ctx* h, int s, double* x, double* y, double* vx, double* vy, double* vz,
                            uint32_t* id, int64_t cap, int64_t* n_out) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  CK(sync_counts(c), "counts");
  SpeciesDev& S = c->sp[s];
  const int64_t n = S.n_host;
  std::vector<double> a[5];
  void* srcs[5] = {S.x, S.y, S.vx, S.vy, S.vz};
  std::vector<float> tiles;
  if (c->dom.precision == EPIC_FP32) {
    tiles.resize((size_t)((n + PTILE - 1) / PTILE) * 4 * PTILE);
    CK(hipMemcpy(tiles.data(), S.x, tiles.size() * sizeof(float), hipMemcpyDeviceToHost), "download");
  }
  for (int k = 0; k < 5; k++) {
    a[k].resize(n);
    if (c->dom.precision == EPIC_FP64) {
      CK(hipMemcpy(a[k].data(), srcs[k], n * sizeof(double), hipMemcpyDeviceToHost), "download");
    } else if (k < 4) {
      for (int64_t i = 0; i < n; i++) a[k][i] = tiles[ptile(i) + k * PTILE];
    } else {
      std::vector<float> t(n);
      CK(hipMemcpy(t.data(), srcs[k], n * sizeof(float), hipMemcpyDeviceToHost), "download");
      for (int64_t i = 0; i < n; i++) a[k][i] = t[i];
    }
  }
  std::vector<uint32_t> ids(n);
  CK(hipMemcpy(ids.data(), S.id, n * sizeof(uint32_t), hipMemcpyDeviceToHost), "download ids");
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++) {
    if (std::isnan(a[0][i])) continue;
    if (m < cap) {
      if (x) x[m] = a[0][i];
      if (y) y[m] = a[1][i];
      if (vx) vx[m] = a[2][i];
      if (vy) vy[m] = a[3][i];
      if (vz) vz[m] = a[4][i];
      if (id) id[m] = ids[i];
    }
    m++;
  }
  if (n_out) *n_out = m;
  if (m > cap) return fail(c, EPIC_ERR_CAPACITY, "download buffer too small");
  return EPIC_OK;
}

int epic_count(epic_ctx* h, int s, int64_t* n_out) {
  if (!h || !n_out) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  double m[7];
  CK(launch_particle_moments(c, s, m), "count");
  *n_out = (int64_t)llround(m[6]);
  return EPIC_OK;
}

int epic_load_species(epic_ctx* h, int s, int64_t n, int vz_thermal) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp || n < 0) return fail(c, EPIC_ERR_ARG, "bad species or count");
  if (n > c->sp[s].cap) return fail(c, EPIC_ERR_CAPACITY, "load exceeds capacity");
  CK(launch_load(c, s, n, vz_thermal), "load");
  c->deposit_valid = false;
  return EPIC_OK;
}

// ---- per-step operators --------------------------------------------------
int epic_scatter(epic_ctx* h, int s) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  stage_begin(c, 0);
  CK(launch_zero_Q(c), "zero");
  CK(launch_deposit(c, s), "deposit");
  CK(launch_den_only(c, s), "den");
  stage_end(c, 0);
  c->deposit_valid = false;
  return EPIC_OK;
}

int epic_compute_rho(epic_ctx* h) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  stage_begin(c, 1);
  CK(launch_den_only(c, -1), "rho");
  stage_end(c, 1);
  return EPIC_OK;
}

int epic_solve_potential(epic_ctx* h, int* iters, double* err) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  int it = 0;
  double e = 0;
  stage_begin(c, 2);
  CK(launch_solve(c, &it, &e), "solve");
  stage_end(c, 2);
  if (iters) *iters = it;
  if (err) *err = e;
  return EPIC_OK;
}

int epic_solve_potential_noalloc(epic_ctx* h, int* iters, double* err) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  const epic_domain_desc& d = c->dom;
  if (!(d.solver == EPIC_SOLVER_PCG_MIXED && d.precond == EPIC_PRECOND_SPECTRAL))
    return fail(c, EPIC_ERR_ARG, "epic_solve_potential_noalloc: the context's solver must be the FP32 sine-transform "
                                 "mixed solve (EPIC_SOLVER_PCG_MIXED with EPIC_PRECOND_SPECTRAL)");
  int it = 0, status = 0;
  double e = 0;
  const char* why = "";
  CK(launch_spectral_solve_noalloc(c, &it, &e, &status, &why), "solve (no allocation)");
  if (status) return fail(c, status, why);
  if (iters) *iters = it;
  if (err) *err = e;
  return EPIC_OK;
}

int epic_efield(epic_ctx* h) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  stage_begin(c, 3);
  CK(launch_efield(c), "efield");
  stage_end(c, 3);
  return EPIC_OK;
}

int epic_push(epic_ctx* h, int s, int sub_cycle) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  stage_begin(c, 4);
  CK(launch_push(c, s, sub_cycle, false, 0.0), "push");
  stage_end(c, 4);
  c->deposit_valid = false;
  return EPIC_OK;
}

int epic_rewind(epic_ctx* h, int s) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  CK(launch_push(c, s, 1, false, 1.0), "rewind");
  return EPIC_OK;
}

int epic_add_emitter(epic_ctx* h, const epic_emitter_desc* e) {
  if (!h || !e) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (e->species_idx1 < 0 || e->species_idx1 >= c->nsp || e->species_idx2 < 0 || e->species_idx2 >= c->nsp ||
      e->numparticle < 0)
    return fail(c, EPIC_ERR_ARG, "bad emitter");
  c->emitters.push_back(Emitter{*e});
  return EPIC_OK;
}

// ---- internal conductors ----------------------------------------------------
// The absorbing shapes and the SOR's rho := voltage nodes of a conductor list,
// computed on the host in the reference's order and arithmetic:
// CircularGrid::initialize (grid.cpp:40-50) places electrode k at
// (x_center + R cos t, y_center + R sin t), t = k (2 pi / n); each Electrode
// (electrode.cpp:4-17) marks 100 points (int)(x0 + r cos u + 0.5) of its
// circle, u = 2 m pi / 100; ReactConductorGrid::initialize (grid.cpp:68-85)
// marks the boundary of its node rectangle.  Later writes to a node win.
static int conductor_plan(const epic_domain_desc& d, const epic_conductor_desc* cd, int n,
                          std::vector<CondDev>& rects, std::vector<CondDev>& elecs, std::vector<int>& node,
                          std::vector<double>& value, std::string& why) {
  const int nx = d.nx, ny = d.ny;
  std::vector<double> val((size_t)nx * ny, 0.0);
  std::vector<char> set((size_t)nx * ny, 0);
  auto write = [&](int i, int j, double v) {
    val[(size_t)i * ny + j] = v;
    set[(size_t)i * ny + j] = 1;
  };
  for (int k = 0; k < n; k++) {
    const epic_conductor_desc& q = cd[k];
    if (q.kind == EPIC_COND_RECT) {
      rects.push_back(CondDev{COND_RECT, (double)q.min_x, (double)q.max_x, (double)q.min_y, (double)q.max_y});
      const int min_i = (int)((q.min_x - d.x0) / d.dx), max_i = (int)((q.max_x - d.x0) / d.dx);
      const int min_j = (int)((q.min_y - d.y0) / d.dy), max_j = (int)((q.max_y - d.y0) / d.dy);
      for (int i = min_i; i <= max_i; ++i)
        for (int j = min_j; j <= max_j; ++j)
          if (i == min_i || i == max_i || j == min_j || j == max_j) {
            if (i < 0 || i >= nx || j < 0 || j >= ny) {
              why = "conductor " + std::to_string(k) + ": rectangle boundary outside the node grid";
              return EPIC_ERR_ARG;
            }
            write(i, j, q.voltage);
          }
    } else if (q.kind == EPIC_COND_CIRCULAR) {
      if (q.electrode_number < 0) {
        why = "conductor " + std::to_string(k) + ": negative electrode_number";
        return EPIC_ERR_ARG;
      }
      for (int e = 0; e < q.electrode_number; ++e) {
        const double theta = e * (2.0 * M_PI / q.electrode_number);
        const double x0 = q.x_center + q.grid_radius * cos(theta);
        const double y0 = q.y_center + q.grid_radius * sin(theta);
        const double r = q.electrode_radius;
        elecs.push_back(CondDev{COND_DISK, x0, y0, r * r, 0.0});
        const int N = 100;
        for (int m = 0; m < N; ++m) {
          const double u = 2.0 * m * M_PI / N;
          const int x = (int)(x0 + r * cos(u) + 0.5);
          const int y = (int)(y0 + r * sin(u) + 0.5);
          if (x >= 0 && x < nx && y >= 0 && y < ny) write(x, y, q.voltage);
        }
      }
    } else {
      why = "conductor " + std::to_string(k) + ": unknown kind";
      return EPIC_ERR_ARG;
    }
  }
  for (size_t t = 0; t < set.size(); t++)
    if (set[t]) {
      node.push_back((int)t);
      value.push_back(val[t]);
    }
  return EPIC_OK;
}

int epic_set_conductors(epic_ctx* h, const epic_conductor_desc* cd, int n) {
  if (!h || n < 0 || (n > 0 && !cd)) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (!c->own_stream && n > 0) return fail(c, EPIC_ERR_ARG, "conductors are not supported on batch members");
  std::vector<CondDev> rects, elecs;
  std::vector<int> node;
  std::vector<double> value;
  std::string why;
  const int rc = conductor_plan(c->dom, cd, n, rects, elecs, node, value, why);
  if (rc != EPIC_OK) return fail(c, rc, why.c_str());
  Conductors& q = c->cond;
  CK(hipStreamSynchronize(c->stream), "sync");
  if (q.table) hipFree(q.table);
  if (q.cmask) hipFree(q.cmask);
  if (q.node) hipFree(q.node);
  if (q.value) hipFree(q.value);
  q.table = nullptr;
  q.cmask = nullptr;
  q.node = nullptr;
  q.value = nullptr;
  q.desc.assign(cd, cd + n);
  q.nrect = (int)rects.size();
  q.nelec = (int)elecs.size();
  q.nnode = (int)node.size();
  std::vector<CondDev> table(rects);
  table.insert(table.end(), elecs.begin(), elecs.end());
  if (!table.empty()) {
    CK(hipMalloc(&q.table, table.size() * sizeof(CondDev)), "conductors");
    CK(hipMemcpy(q.table, table.data(), table.size() * sizeof(CondDev), hipMemcpyHostToDevice), "conductors");
    // Cells (i * (ny - 1) + j) that a shape's bounding box reaches, widened
    // by one cell on each side: a particle in an unflagged cell is on no
    // conductor (the kernels compute its cell in FP32, whose rounding moves
    // it by at most one cell), so the FP32 pushes skip the exact test there.
    const epic_domain_desc& d = c->dom;
    const int ncx = d.nx - 1, ncy = d.ny - 1;
    std::vector<uint32_t> cm(((size_t)ncx * ncy + 31) / 32, c->cond_cell_mask ? 0u : ~0u);
    auto cell_of = [](double v, double v0, double h, int ncell) {
      const double t = std::floor((v - v0) / h);
      return (int)std::max(-2.0, std::min(t, (double)ncell + 1));
    };
    for (const CondDev& t : table) {
      double xl, xh, yl, yh;
      if (t.kind == COND_DISK) {
        const double r = std::sqrt(std::max(t.c, 0.0));
        xl = t.a - r; xh = t.a + r; yl = t.b - r; yh = t.b + r;
      } else {
        xl = t.a; xh = t.b; yl = t.c; yh = t.d;
      }
      const int i0 = std::max(0, cell_of(xl, d.x0, d.dx, ncx) - 1), i1 = std::min(ncx - 1, cell_of(xh, d.x0, d.dx, ncx) + 1);
      const int j0 = std::max(0, cell_of(yl, d.y0, d.dy, ncy) - 1), j1 = std::min(ncy - 1, cell_of(yh, d.y0, d.dy, ncy) + 1);
      for (int i = i0; i <= i1; i++)
        for (int j = j0; j <= j1; j++) {
          const size_t k = (size_t)i * ncy + j;
          cm[k >> 5] |= 1u << (k & 31);
        }
    }
    CK(hipMalloc(&q.cmask, cm.size() * sizeof(uint32_t)), "conductors");
    CK(hipMemcpy(q.cmask, cm.data(), cm.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "conductors");
  }
  if (q.nnode > 0) {
    CK(hipMalloc(&q.node, node.size() * sizeof(int)), "conductors");
    CK(hipMalloc(&q.value, value.size() * sizeof(double)), "conductors");
    CK(hipMemcpy(q.node, node.data(), node.size() * sizeof(int), hipMemcpyHostToDevice), "conductors");
    CK(hipMemcpy(q.value, value.data(), value.size() * sizeof(double), hipMemcpyHostToDevice), "conductors");
  }
  return EPIC_OK;
}

int epic_inject(epic_ctx* h, int step) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  stage_begin(c, 7);
  CK(launch_inject(c, step), "inject");
  stage_end(c, 7);
  return EPIC_OK;
}

int epic_mcc_setup(epic_ctx* h, const epic_mcc_desc* d) {
  if (!h || !d) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (d->ntab < 2 || !d->sigma || !(d->e_min > 0) || !(d->e_max > d->e_min)) return fail(c, EPIC_ERR_ARG, "bad table");
  if (d->event_cap > MCC_EVENT_MAX) return fail(c, EPIC_ERR_ARG, "event_cap above 2^28");
  if (d->electron_sp >= c->nsp || d->ion_sp >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  MccState& M = c->mcc;
  // a repeated setup replaces the table and the event list
  if (M.sigma) c->bytes -= (size_t)EPIC_MCC_NPROC * M.d.ntab * sizeof(double);
  if (M.events) c->bytes -= (size_t)M.event_cap * 8 * sizeof(double);
  M.d = *d;
  std::vector<double> sig(d->sigma, d->sigma + (size_t)EPIC_MCC_NPROC * d->ntab);
  if (M.sigma) hipFree(M.sigma);
  M.sigma = nullptr;
  CK(dalloc_c(c, &M.sigma, sig.size()), "mcc table");
  CK(hipMemcpy(M.sigma, sig.data(), sig.size() * sizeof(double), hipMemcpyHostToDevice), "mcc table");
  M.d.sigma = nullptr;
  M.event_cap = mcc_event_cap(d->event_cap, d->electron_sp >= 0 ? (c->sp[d->electron_sp].cap + MCC_SUB - 1) /
                                                                      MCC_SUB * MCC_SUB
                                                                : 0);
  if (!M.counts) CK(dalloc_c(c, &M.counts, EPIC_MCC_NPROC + 2), "mcc counts");
  if (M.events) hipFree(M.events);
  M.events = nullptr;
  CK(dalloc_c(c, &M.events, (size_t)M.event_cap * 8), "mcc events");
  if (!M.event_count) CK(dalloc_c(c, &M.event_count, 1), "mcc events");
  // real-collision lists of the colliding species, sized by their
  // capacities, their counts and the electrons' per-chunk event counts
  auto list_bytes = [](int64_t len, int64_t len_e) -> size_t {
    return (size_t)std::max<int64_t>(len, 1) * sizeof(uint2) + 2 * sizeof(int) +
           (size_t)std::max<int64_t>(len_e / 64, 1) * sizeof(int);
  };
  if (M.list) {
    c->bytes -= list_bytes(M.list_len_e + M.list_len_i, M.list_len_e);
    hipFree(M.list);
    hipFree(M.list_n);
    hipFree(M.chunk_ev);
    M.list = nullptr;
    M.list_n = nullptr;
    M.chunk_ev = nullptr;
  }
  auto list_len = [&](int s) -> int64_t {
    return s >= 0 ? (c->sp[s].cap + MCC_SUB - 1) / MCC_SUB * MCC_SUB : 0;
  };
  M.list_len_e = list_len(d->electron_sp);
  M.list_len_i = list_len(d->ion_sp);
  const int64_t len = M.list_len_e + M.list_len_i;
  CK(dalloc_c(c, &M.list, (size_t)len), "mcc collision lists");
  CK(dalloc_c(c, &M.list_n, 2), "mcc collision lists");
  CK(dalloc_c(c, &M.chunk_ev, (size_t)(M.list_len_e / 64)), "mcc collision lists");
  M.nu_max_e = d->electron_sp >= 0 ? mcc_nu_max(M.d, sig, ME, true) : 0.0;
  M.nu_max_i = d->ion_sp >= 0 ? mcc_nu_max(M.d, sig, c->sp[d->ion_sp].desc.mass, false) : 0.0;
  M.on = true;
  return EPIC_OK;
}

int epic_collide(epic_ctx* h, int step) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  stage_begin(c, 5);
  CK(launch_mcc(c, step), "collide");
  stage_end(c, 5);
  return EPIC_OK;
}

int epic_mcc_get_stats(epic_ctx* h, epic_mcc_stats* out) {
  if (!h || !out) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  unsigned long long v[EPIC_MCC_NPROC + 2] = {};
  if (c->mcc.counts) {
    CK(hipMemcpyAsync(v, c->mcc.counts, sizeof(v), hipMemcpyDeviceToHost, c->stream), "stats");
    CK(hipStreamSynchronize(c->stream), "stats");
  }
  for (int p = 0; p < EPIC_MCC_NPROC; p++) out->count[p] = (int64_t)v[p];
  out->births = (int64_t)v[EPIC_MCC_NPROC];
  out->eloss_ev = v[EPIC_MCC_E_EXCITATION] * c->mcc.d.e_exc_th + v[EPIC_MCC_E_IONIZATION] * c->mcc.d.e_ion_th;
  return EPIC_OK;
}

int epic_sort(epic_ctx* h, int s) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  stage_begin(c, 6);
  CK(sync_counts(c), "counts");
  CK(launch_sort(c, s), "sort");
  stage_end(c, 6);
  return EPIC_OK;
}

int epic_compact(epic_ctx* h, int s) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  CK(sync_counts(c), "counts");
  CK(launch_compact(c, s), "compact");
  return EPIC_OK;
}

// One pass of main.cpp:395-423: inject, scatter, ComputeRho, PotentialSolver,
// CalculateEfield, push (next deposit fused), collide; sort on schedule.
int epic_step(epic_ctx* h, int ts, int* iters) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (!c->emitters.empty()) {
    stage_begin(c, 7);
    CK(launch_inject(c, ts), "inject");
    stage_end(c, 7);
  }
  if (!c->deposit_valid) {
    stage_begin(c, 0);
    CK(launch_zero_Q(c), "zero");
    for (int s = 0; s < c->nsp; s++) CK(launch_deposit(c, s), "deposit");
    stage_end(c, 0);
  }
  // the sine-transform solve finalises rho in its first kernel (unless the
  // rho stage is being timed on its own)
  const epic_domain_desc& dd = c->dom;
  const bool spectral = dd.solver == EPIC_SOLVER_SPECTRAL ||
                        (dd.solver == EPIC_SOLVER_PCG_MIXED && dd.precond == EPIC_PRECOND_SPECTRAL);
  c->rho_in_solve = spectral && !(c->timing && ((c->timing_mask >> 1) & 1u));
  if (!c->rho_in_solve) {
    stage_begin(c, 1);
    CK(launch_rho_finalize(c), "rho");
    stage_end(c, 1);
  }
  int it = 0;
  double err = 0;
  stage_begin(c, 2);
  // the sine-transform solve launches the E-field itself after its last
  // convergence check (timed with the solve then)
  c->efield_in_solve = true;
  c->efield_done = false;
  const hipError_t solve_rc = launch_solve(c, &it, &err);
  c->efield_in_solve = false;
  c->rho_in_solve = false;
  CK(solve_rc, "solve");
  stage_end(c, 2);
  if (iters) *iters = it;
  if (!c->efield_done) {
    stage_begin(c, 3);
    CK(launch_efield(c), "efield");
    stage_end(c, 3);
  }
  c->efield_done = false;
  // every pushed species in one launch where the FP32 hot kernel applies
  const bool all = push_all_ok(c);
  if (all) {
    stage_begin(c, 4);
    CK(launch_push_all(c), "push");
    stage_end(c, 4);
  }
  for (int s = 0; s < c->nsp; s++) {
    if (all && c->sp[s].desc.pushed) continue;
    stage_begin(c, 4);
    if (c->sp[s].desc.pushed) CK(launch_push(c, s, 1, true, 0.0), "push");
    else CK(launch_deposit(c, s), "deposit");
    stage_end(c, 4);
  }
  c->deposit_valid = true;
  if (c->mcc.on) {
    stage_begin(c, 5);
    CK(launch_mcc(c, ts), "collide");
    stage_end(c, 5);
  }
  c->step_count++;
  bool any_due = false;
  for (int s = 0; s < c->nsp; s++) any_due = any_due || c->sort_due(s);
  if (any_due) {
    stage_begin(c, 6);
    CK(sync_counts(c), "counts");
    for (int s = 0; s < c->nsp; s++)
      if (c->sort_due(s)) CK(launch_sort(c, s), "sort");
    stage_end(c, 6);
  }
  return EPIC_OK;
}

// main.cpp:366-382: scatter every species, ComputeRho, solve, E, rewind.
int epic_initial_field(epic_ctx* h) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  CK(launch_zero_Q(c), "zero");
  for (int s = 0; s < c->nsp; s++) CK(launch_deposit(c, s), "deposit");
  CK(launch_rho_finalize(c), "rho");
  int it;
  double err;
  CK(launch_solve(c, &it, &err), "solve");
  CK(launch_efield(c), "efield");
  for (int s = 0; s < c->nsp; s++) CK(launch_push(c, s, 1, false, 1.0), "rewind");
  c->deposit_valid = false;
  return EPIC_OK;
}

// ---- diagnostics -----------------------------------------------------------
int epic_compute_ke(epic_ctx* h, int s, int ns, double out3[3]) {
  if (!h || !out3) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp || ns < 0 || ns >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  double m[7];
  CK(launch_particle_moments(c, s, m), "ke");
  const epic_species_desc& a = c->sp[s].desc;
  const epic_species_desc& b = c->sp[ns].desc;
  const double vn = c->dom.vel_norm;
  const double Th = (b.temp * eV) * (b.spwt) * (double)b.numparticle;  // species.cpp:264
  for (int k = 0; k < 3; k++) out3[k] = ((m[k] * vn) * vn) * (0.5 * (a.spwt * a.mass)) / Th;
  return EPIC_OK;
}

int epic_compute_momentum(epic_ctx* h, int s, int ns, double out3[3]) {
  if (!h || !out3) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (s < 0 || s >= c->nsp || ns < 0 || ns >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  double m[7];
  CK(launch_particle_moments(c, s, m), "momentum");
  const epic_species_desc& a = c->sp[s].desc;
  const epic_species_desc& b = c->sp[ns].desc;
  const double vn = c->dom.vel_norm;
  const double Thp = sqrt(b.temp * eV / ME) * b.spwt * (double)b.numparticle * ME;  // species.cpp:290
  for (int k = 0; k < 3; k++) out3[k] = ((m[3 + k] * vn) * (a.spwt * a.mass)) / Thp;
  return EPIC_OK;
}

int epic_compute_pe(epic_ctx* h, int ns, double* out) {
  if (!h || !out) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (ns < 0 || ns >= c->nsp) return fail(c, EPIC_ERR_ARG, "bad species");
  double se2;
  CK(launch_field_energy(c, &se2), "pe");
  const epic_domain_desc& d = c->dom;
  const epic_species_desc& b = c->sp[ns].desc;
  const double pe = se2 * (d.L * d.L);
  const double unnorm = (d.density * QE * d.L) / EPS_0;  // domain.cpp:269
  const double pe_norm = 0.5 * EPS_0 * pe * unnorm * unnorm;
  const double Th = (b.temp * eV) * b.spwt * (double)b.numparticle;
  *out = pe_norm / Th;
  return EPIC_OK;
}

static double* field_ptr(Ctx* c, int which) {
  switch (which) {
    case 0: return c->rho;
    case 1: return c->phi;
    case 2: return c->efx;
    case 3: return c->efy;
    default:
      if (which >= 10 && which < 10 + c->nsp) return c->sp[which - 10].den;
      return nullptr;
  }
}

int epic_get_field(epic_ctx* h, int which, double* out) {
  if (!h || !out) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  double* p = field_ptr(c, which);
  if (!p) return fail(c, EPIC_ERR_ARG, "bad field");
  CK(hipMemcpyAsync(out, p, c->nnodes * sizeof(double), hipMemcpyDeviceToHost, c->stream), "get field");
  CK(hipStreamSynchronize(c->stream), "get field");
  return EPIC_OK;
}

int epic_set_field(epic_ctx* h, int which, const double* in) {
  if (!h || !in) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  double* p = field_ptr(c, which);
  if (!p) return fail(c, EPIC_ERR_ARG, "bad field");
  CK(hipMemcpyAsync(p, in, c->nnodes * sizeof(double), hipMemcpyHostToDevice, c->stream), "set field");
  if (which == 2 || which == 3) CK(launch_pack_E2(c), "pack E");
  CK(hipStreamSynchronize(c->stream), "set field");
  return EPIC_OK;
}

int epic_error_counters(epic_ctx* h, int64_t out[EPIC_NERR]) {
  if (!h || !out) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  unsigned long long v[EPIC_NERR];
  CK(hipMemcpyAsync(v, c->errc, sizeof(v), hipMemcpyDeviceToHost, c->stream), "errc");
  CK(hipStreamSynchronize(c->stream), "errc");
  for (int k = 0; k < EPIC_NERR; k++) out[k] = (int64_t)v[k] + (int64_t)c->errc_host[k];
  return EPIC_OK;
}

int epic_clear_error_counters(epic_ctx* h) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  // particles dropped as non-finite stay as NaN slots until their species'
  // next sort or compaction; the collision pass learns of them from this count
  unsigned long long nan_drops = 0;
  CK(hipMemcpyAsync(&nan_drops, c->errc + EPIC_ERRC_NAN, sizeof(nan_drops), hipMemcpyDeviceToHost, c->stream),
     "errc");
  CK(hipStreamSynchronize(c->stream), "errc");
  if (nan_drops || c->errc_host[EPIC_ERRC_NAN])
    for (int s = 0; s < c->nsp; s++) c->sp[s].maybe_dead = true;
  CK(hipMemsetAsync(c->errc, 0, EPIC_NERR * sizeof(unsigned long long), c->stream), "errc");
  CK(hipStreamSynchronize(c->stream), "errc");
  for (int k = 0; k < EPIC_NERR; k++) c->errc_host[k] = 0;
  return EPIC_OK;
}

int epic_timing_enable(epic_ctx* h, int on) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  if (on && c->ev_pool.empty()) {
    c->ev_pool.resize(2 * 2048);
    c->ev_stage.assign(2048, 0);
    for (auto& ev : c->ev_pool) CK(hipEventCreate(&ev), "hipEventCreate");
  }
  c->ev_used = 0;
  c->ev_open = -1;
  c->timing = on != 0;
  // EPIC_TIMING_ONLY | stage bits: only those stages record events
  c->timing_mask = (on & EPIC_TIMING_ONLY) ? (unsigned)on & ((1u << EPIC_NSTAGE) - 1) : ~0u;
  for (int k = 0; k < EPIC_NSTAGE; k++) {
    c->stage_ms[k] = 0;
    c->stage_calls[k] = 0;
  }
  return EPIC_OK;
}

int epic_timing_read(epic_ctx* h, double ms[EPIC_NSTAGE], int64_t calls[EPIC_NSTAGE]) {
  if (!h) return EPIC_ERR_ARG;
  Ctx* c = &h->c;
  resolve_events(c);
  for (int k = 0; k < EPIC_NSTAGE; k++) {
    if (ms) ms[k] = c->stage_ms[k];
    if (calls) calls[k] = c->stage_calls[k];
  }
  return EPIC_OK;
}

}  // extern "C"
