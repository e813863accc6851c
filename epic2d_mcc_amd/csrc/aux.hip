// Species storage kernels: native load, stable compaction, stable counting
// sort by cell (sort_ops.hpp), emitter injection.
//
// Compaction keeps the order of the live particles, which is what the
// reference's vector::erase removal does (species.cpp:95-121).
#include <algorithm>

#include "sort_ops.hpp"

namespace epic {

constexpr int AUX_BLOCK = 256;
constexpr int CMP_PPT = 8;  // elements per thread in compaction
constexpr int CMP_CH = AUX_BLOCK * CMP_PPT;

// ---- native load ------------------------------------------------------------
template <typename P>
__global__ void load_kernel(P* x, P* y, P* vx, P* vy, P* vz, uint32_t* id, int64_t n, double x0, double y0,
                            double Lx, double Ly, double sig, double ux, double uy, int vz_thermal, uint32_t k0,
                            uint32_t k1, uint32_t sp) {
  const P xhi = (P)(x0 + Lx), yhi = (P)(y0 + Ly);
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t lo = (uint32_t)p, hi = (uint32_t)(p >> 32);
    const u32x4 r0 = philox4x32_10({lo, hi, RNG_LOAD_POS, sp}, k0, k1);
    const u32x4 r1 = philox4x32_10({lo, hi, RNG_LOAD_VEL, sp}, k0, k1);
    const u32x4 r2 = philox4x32_10({lo, hi, RNG_LOAD_VEL2, sp}, k0, k1);
    P px = (P)(x0 + Lx * u01_53(r0.a, r0.b));
    P py = (P)(y0 + Ly * u01_53(r0.c, r0.d));
    if (px >= xhi) px = nextafter(xhi, (P)x0);
    if (py >= yhi) py = nextafter(yhi, (P)y0);
    // Box-Muller: (1 - u) in (0, 1]
    const double u1 = 1.0 - u01_53(r1.a, r1.b), u2 = u01_53(r1.c, r1.d);
    const double u3 = 1.0 - u01_53(r2.a, r2.b), u4 = u01_53(r2.c, r2.d);
    const double m1 = sqrt(-2.0 * log(u1)), m2 = sqrt(-2.0 * log(u3));
    const double g1 = m1 * cos(2.0 * M_PI * u2), g2 = m1 * sin(2.0 * M_PI * u2);
    const double g3 = m2 * cos(2.0 * M_PI * u4);
    const int64_t q = pslot<P>(p);
    x[q] = px;
    y[q] = py;
    vx[q] = (P)(sig * g1 + ux);
    vy[q] = (P)(sig * g2 + uy);
    vz[p] = (P)(vz_thermal ? sig * g3 : 0.0);
    id[p] = (uint32_t)p;
  }
}

__global__ void set_count_kernel(int64_t* d_n, int64_t v) {
  d_n[0] = v;
  d_n[1] = v;
}

hipError_t launch_load(Ctx* c, int s, int64_t n, int vz_thermal) {
  SpeciesDev& S = c->sp[s];
  const epic_domain_desc& d = c->dom;
  const double Lx = (d.nx - 1) * d.dx, Ly = (d.ny - 1) * d.dy;
  // sigma = sqrt(kT/m): the spread of SampleVel's Irwin-Hall sum
  // v_th*sqrt(2)*(R1+R2+R3-1.5) with v_th = sqrt(2 kT/m) (init.cpp:104-108)
  const double kT = S.desc.temp * QE;
  const double sig = sqrt(kT / S.desc.mass) / d.vel_norm;
  const double ux = S.desc.vsx, uy = S.desc.vsy;
  const uint32_t k0 = (uint32_t)d.seed, k1 = (uint32_t)(d.seed >> 32);
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  if (d.precision == EPIC_FP64)
    hipLaunchKernelGGL(load_kernel<double>, dim3(std::max(grid, 1)), dim3(256), 0, c->stream, (double*)S.x,
                       (double*)S.y, (double*)S.vx, (double*)S.vy, (double*)S.vz, S.id, n, d.x0, d.y0, Lx, Ly, sig,
                       ux, uy, vz_thermal, k0, k1, (uint32_t)s);
  else
    hipLaunchKernelGGL(load_kernel<float>, dim3(std::max(grid, 1)), dim3(256), 0, c->stream, (float*)S.x,
                       (float*)S.y, (float*)S.vx, (float*)S.vy, (float*)S.vz, S.id, n, d.x0, d.y0, Lx, Ly, sig, ux,
                       uy, vz_thermal, k0, k1, (uint32_t)s);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(set_count_kernel, dim3(1), dim3(1), 0, c->stream, S.d_n, n);
  S.n_host = n;
  return hipGetLastError();
}

// ---- block-level exclusive scan helper --------------------------------------
// Exclusive scan of one int per thread over a 256-thread block.
__device__ __forceinline__ int block_excl_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) sh[wid] = inc;
  __syncthreads();
  int wbase = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    if (w < wid) wbase += sh[w];
    tot += sh[w];
  }
  __syncthreads();
  *total = tot;
  return wbase + inc - v;
}

// ---- compaction -------------------------------------------------------------
template <typename P>
__global__ __launch_bounds__(AUX_BLOCK) void live_count_kernel(const P* x, const int64_t* d_n, int* blk) {
  __shared__ int sh[AUX_BLOCK / 64];
  const int64_t n = *d_n;
  const int64_t base = (int64_t)blockIdx.x * CMP_CH + (int64_t)threadIdx.x * CMP_PPT;
  int c = 0;
  for (int k = 0; k < CMP_PPT; k++) {
    const int64_t i = base + k;
    if (i < n && !isnan(x[pslot<P>(i)])) c++;
  }
  int tot;
  block_excl_scan(c, sh, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// Exclusive scan of m ints in place; a[m] receives the total.  Three
// kernels: per-tile sums, a scan of the tile sums (one workgroup), and the
// per-tile scan with its offset.  Tiles of 1024 elements.
constexpr int SCAN_TILE = 1024;

__device__ __forceinline__ long long block_scan_ll(long long v, long long* sh, long long* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  long long inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const long long t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) sh[wid] = inc;
  __syncthreads();
  long long base = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    if (w < wid) base += sh[w];
    tot += sh[w];
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

__global__ __launch_bounds__(SCAN_TILE) void scan_tile_sums_kernel(const int* a, int64_t m, long long* tiles) {
  __shared__ long long sh[SCAN_TILE / 64];
  const int64_t k = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  long long tot;
  block_scan_ll(k < m ? a[k] : 0, sh, &tot);
  if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_TILE) void scan_tiles_kernel(long long* tiles, int nt) {
  __shared__ long long sh[SCAN_TILE / 64];
  long long run = 0;
  for (int b0 = 0; b0 < nt; b0 += SCAN_TILE) {
    const int k = b0 + threadIdx.x;
    const long long v = k < nt ? tiles[k] : 0;
    long long tot;
    const long long ex = block_scan_ll(v, sh, &tot);
    if (k < nt) tiles[k] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) tiles[nt] = run;
}

__global__ __launch_bounds__(SCAN_TILE) void scan_apply_kernel(int* a, int64_t m, const long long* tiles, int nt) {
  __shared__ long long sh[SCAN_TILE / 64];
  const int64_t k = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  const int v = k < m ? a[k] : 0;
  long long tot;
  const long long ex = block_scan_ll(v, sh, &tot);
  if (k < m) a[k] = (int)(tiles[blockIdx.x] + ex);
  if (blockIdx.x == 0 && threadIdx.x == 0) a[m] = (int)tiles[nt];
}

static hipError_t scan_ints(Ctx* c, int* a, int64_t m) {
  const int nt = (int)((m + SCAN_TILE - 1) / SCAN_TILE);
  long long* tiles = c->scan_tiles;
  if (nt + 1 > c->scan_tiles_cap) return hipErrorInvalidValue;
  hipLaunchKernelGGL(scan_tile_sums_kernel, dim3(nt), dim3(SCAN_TILE), 0, c->stream, a, m, tiles);
  hipLaunchKernelGGL(scan_tiles_kernel, dim3(1), dim3(SCAN_TILE), 0, c->stream, tiles, nt);
  hipLaunchKernelGGL(scan_apply_kernel, dim3(nt), dim3(SCAN_TILE), 0, c->stream, a, m, tiles, nt);
  return hipGetLastError();
}

template <typename P>
__global__ __launch_bounds__(AUX_BLOCK) void compact_scatter_kernel(const P* x, const P* y, const P* vx, const P* vy,
                                                                    const P* vz, const uint32_t* id, P* x2, P* y2,
                                                                    P* vx2, P* vy2, P* vz2, uint32_t* id2,
                                                                    const int64_t* d_n, const int* blk_off) {
  __shared__ int sh[AUX_BLOCK / 64];
  const int64_t n = *d_n;
  const int64_t base = (int64_t)blockIdx.x * CMP_CH + (int64_t)threadIdx.x * CMP_PPT;
  int c = 0;
  for (int k = 0; k < CMP_PPT; k++) {
    const int64_t i = base + k;
    if (i < n && !isnan(x[pslot<P>(i)])) c++;
  }
  int tot;
  int off = blk_off[blockIdx.x] + block_excl_scan(c, sh, &tot);
  for (int k = 0; k < CMP_PPT; k++) {
    const int64_t i = base + k;
    const int64_t si = pslot<P>(i);
    if (i < n && !isnan(x[si])) {
      const int64_t so = pslot<P>(off);
      x2[so] = x[si]; y2[so] = y[si]; vx2[so] = vx[si]; vy2[so] = vy[si]; vz2[off] = vz[i]; id2[off] = id[i];
      off++;
    }
  }
}

__global__ void set_count_from_int_kernel(int64_t* d_n, const int* v) { d_n[0] = (int64_t)*v; }

static void swap_buffers(SpeciesDev& S) {
  std::swap(S.x, S.x2); std::swap(S.y, S.y2); std::swap(S.vx, S.vx2);
  std::swap(S.vy, S.vy2); std::swap(S.vz, S.vz2); std::swap(S.id, S.id2);
}

template <typename P>
static hipError_t compact_t(Ctx* c, int s) {
  SpeciesDev& S = c->sp[s];
  const int64_t n = S.n_host;
  const int64_t nb = (n + CMP_CH - 1) / CMP_CH;
  if (nb == 0) return hipSuccess;
  if (nb + 1 > c->blk_cap) return hipErrorInvalidValue;
  hipLaunchKernelGGL(live_count_kernel<P>, dim3(nb), dim3(AUX_BLOCK), 0, c->stream, (const P*)S.x, S.d_n, c->blk_count);
  EPIC_HIP_TRY(hipGetLastError());
  EPIC_HIP_TRY(scan_ints(c, c->blk_count, nb));
  hipLaunchKernelGGL(compact_scatter_kernel<P>, dim3(nb), dim3(AUX_BLOCK), 0, c->stream, (const P*)S.x,
                     (const P*)S.y, (const P*)S.vx, (const P*)S.vy, (const P*)S.vz, S.id, (P*)S.x2, (P*)S.y2,
                     (P*)S.vx2, (P*)S.vy2, (P*)S.vz2, S.id2, S.d_n, c->blk_count);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(set_count_from_int_kernel, dim3(1), dim3(1), 0, c->stream, S.d_n, c->blk_count + nb);
  EPIC_HIP_TRY(hipGetLastError());
  swap_buffers(S);
  int64_t nn;
  EPIC_HIP_TRY(hipMemcpyAsync(&nn, S.d_n, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
  S.n_host = nn;
  S.maybe_dead = false;  // compaction removed every NaN slot
  return hipSuccess;
}

hipError_t launch_compact(Ctx* c, int s) {
  if (c->dom.precision == EPIC_FP64) return compact_t<double>(c, s);
  return compact_t<float>(c, s);
}

// ---- counting sort by cell (sort_ops.hpp) ------------------------------------
int64_t sort_key_space(int nx, int ny) { return key_map(nx - 1, ny - 1).K; }

template <typename P>
__global__ __launch_bounds__(256) void sort_hist_kernel(const P* x, const P* y, const int64_t* d_n, int* count, P x0,
                                                        P y0, P dx, P dy, KeyMap m, uint32_t* pks) {
  __shared__ int win[HWIN];
  __shared__ int wbase_s;
  const int64_t n = *d_n;
  const int64_t stride = (int64_t)gridDim.x * SORT_CHUNK;
  for (int64_t base = blockIdx.x * (int64_t)SORT_CHUNK; base < n; base += stride)
    hist_chunk(x, y, n, base, count, x0, y0, dx, dy, m, win, &wbase_s, pks);
}

__global__ __launch_bounds__(256) void sort_fill_kernel(const uint32_t* pks, const int64_t* d_n, const int* start,
                                                        int* fill, uint32_t* perm, KeyMap m) {
  __shared__ int win[HWIN];
  __shared__ int wbase_s;
  const int64_t n = *d_n;
  const int64_t stride = (int64_t)gridDim.x * SORT_CHUNK;
  for (int64_t base = blockIdx.x * (int64_t)SORT_CHUNK; base < n; base += stride)
    fill_chunk(pks, n, base, start, fill, perm, m, win, &wbase_s);
}

// One wavefront per FIX_GROUP cells.
__global__ __launch_bounds__(256) void sort_fix_kernel(const int* start, uint32_t* perm, uint32_t* scratch,
                                                        int ncell) {
  __shared__ __attribute__((aligned(16))) uint32_t buf[4][FIX_LDS + 4];
  const int w = threadIdx.x >> 6;
  for (int c0 = (blockIdx.x * 4 + w) * FIX_GROUP; c0 < ncell; c0 += gridDim.x * 4 * FIX_GROUP)
    fix_cells(start, perm, scratch, c0, ncell, buf[w]);
}

// Each XCD walks its own contiguous range of chunks, its workgroups side by
// side: the source slots of neighbouring chunks overlap and stay in that
// XCD's L2.  Launched with one workgroup per CU, so that the chunks in flight
// on an XCD (32 x 2048 slots, ~1.5 MB of source) fit its 4 MB L2 while their
// neighbours read the same lines: 16 workgroups per CU keep 12 MB in flight
// and read each moved particle's lines again from HBM (gather 556-640 us per
// 26.2M-particle species against 495-536 us at one, 469-510 us with the
// nontemporal streams).
template <typename P>
__global__ __launch_bounds__(256) void sort_gather_kernel(const uint32_t* perm, const int* total, const P* x,
                                                          const P* y, const P* vx, const P* vy, const P* vz,
                                                          const uint32_t* id, P* x2, P* y2, P* vx2, P* vy2, P* vz2,
                                                          uint32_t* id2) {
  const int64_t n = *total;
  int64_t cb, ce, cs;
  xcd_chunk_range((n + SORT_CHUNK - 1) / SORT_CHUNK, &cb, &ce, &cs);
  uint32_t p[SORT_U], pn[SORT_U];
  if (cb < ce) gather_sources(perm, n, cb * SORT_CHUNK, pn);
  for (int64_t chunk = cb; chunk < ce; chunk += cs) {
#pragma unroll
    for (int u = 0; u < SORT_U; u++) p[u] = pn[u];
    const int64_t next = chunk + cs < ce ? (chunk + cs) * SORT_CHUNK : -1;
    gather_move(p, n, chunk * SORT_CHUNK, x, y, vx, vy, vz, id, x2, y2, vx2, vy2, vz2, id2, perm, next, pn);
  }
}

__global__ void zero_int_kernel(int* p, int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) p[k] = 0;
}

template <typename P>
static hipError_t sort_t(Ctx* c, int s) {
  SpeciesDev& S = c->sp[s];
  const epic_domain_desc& d = c->dom;
  const KeyMap km = key_map(d.nx - 1, d.ny - 1);
  const int ncell = km.K;  // key space (tile-padded cells)
  const int64_t n = S.n_host;
  if (n == 0) return hipSuccess;
  const P x0 = (P)d.x0, y0 = (P)d.y0, dx = (P)d.dx, dy = (P)d.dy;
  const int g = std::max(1, (int)std::min<int64_t>((n + 256 * SORT_U - 1) / (256 * SORT_U), (int64_t)c->num_cu * 16));
  const int gc = std::max(1, std::min((ncell + 2 + 255) / 256, 4096));
  hipLaunchKernelGGL(zero_int_kernel, dim3(gc), dim3(256), 0, c->stream, c->sort_count, (int64_t)ncell + 2);
  hipLaunchKernelGGL(zero_int_kernel, dim3(gc), dim3(256), 0, c->stream, c->sort_fill, (int64_t)ncell + 2);
  // the histogram's sorted (key, lane) words go to perm's second half (the
  // per-cell fix's scratch, which is not used before the fix)
  uint32_t* pks = c->sort_perm + c->sort_cap;
  hipLaunchKernelGGL(sort_hist_kernel<P>, dim3(g), dim3(256), 0, c->stream, (const P*)S.x, (const P*)S.y, S.d_n,
                     c->sort_count, x0, y0, dx, dy, km, pks);
  EPIC_HIP_TRY(hipGetLastError());
  // start[k] = exclusive scan; start[ncell] = live total; start[ncell+1] = all
  EPIC_HIP_TRY(scan_ints(c, c->sort_count, (int64_t)ncell + 1));
  hipLaunchKernelGGL(sort_fill_kernel, dim3(g), dim3(256), 0, c->stream, (const uint32_t*)pks, S.d_n,
                     c->sort_count, c->sort_fill, c->sort_perm, km);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(sort_fix_kernel, dim3(std::min((ncell + 4 * FIX_GROUP - 1) / (4 * FIX_GROUP), 65536)), dim3(256),
                     0, c->stream,
                     c->sort_count, c->sort_perm, c->sort_perm + c->sort_cap, ncell);
  EPIC_HIP_TRY(hipGetLastError());
  const int gg =
      std::max(1, (int)std::min<int64_t>((n + SORT_CHUNK - 1) / SORT_CHUNK, (int64_t)c->num_cu * c->sort_gather_wg));
  hipLaunchKernelGGL(sort_gather_kernel<P>, dim3(gg), dim3(256), 0, c->stream, c->sort_perm, c->sort_count + ncell,
                     (const P*)S.x, (const P*)S.y, (const P*)S.vx, (const P*)S.vy, (const P*)S.vz, S.id, (P*)S.x2,
                     (P*)S.y2, (P*)S.vx2, (P*)S.vy2, (P*)S.vz2, S.id2);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(set_count_from_int_kernel, dim3(1), dim3(1), 0, c->stream, S.d_n, c->sort_count + ncell);
  EPIC_HIP_TRY(hipGetLastError());
  swap_buffers(S);
  int64_t nn;
  EPIC_HIP_TRY(hipMemcpyAsync(&nn, S.d_n, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
  S.n_host = nn;
  S.maybe_dead = false;  // the sort dropped every NaN slot
  return hipSuccess;
}

hipError_t launch_sort(Ctx* c, int s) {
  if (c->dom.precision == EPIC_FP64) return sort_t<double>(c, s);
  return sort_t<float>(c, s);
}

// ---- emitter injection (Emitter::inject, emitter.cpp:5-36) -------------------
// Per emitter and step, numparticle particles at (unirand(x0,x1),
// unirand(y0,y1)) for species_idx1 and, when different, the same positions for
// species_idx2; velocity SampleVel(temp) + vd*vel_norm, normalised, vz = 0.
// Positions and velocities come from Philox keyed by (emitter, step, k) in
// place of the reference's random_device-seeded draws.
template <typename P>
__global__ void inject_kernel(P* x, P* y, P* vx, P* vy, P* vz, uint32_t* id, int64_t* d_n, int64_t cap, int k, epic_emitter_desc e, double vth_sqrt2, double vel_norm,
                              uint32_t step, uint32_t emitter, uint32_t which, uint32_t k0, uint32_t k1,
                              unsigned long long* Q, int nx, int ny, double x0, double y0, double dx, double dy,
                              int fix_bits, unsigned long long* errc) {
  const int64_t n = d_n[0];
  const uint32_t id0 = (uint32_t)d_n[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < k; t += gridDim.x * blockDim.x) {
    const int64_t slot = n + t;
    if (slot >= cap) {
      atomicAdd(&errc[EPIC_ERRC_CAPACITY], 1ull);
      continue;
    }
    const u32x4 rp = philox4x32_10({(uint32_t)t, step, RNG_INJECT, emitter}, k0, k1);
    const double px = e.x0 + (e.x1 - e.x0) * u01_53(rp.a, rp.b);
    const double py = e.y0 + (e.y1 - e.y0) * u01_53(rp.c, rp.d);
    // Init::SampleVel's three uniforms per component (init.cpp:110-114),
    // six independent 53-bit draws from three Philox blocks
    const u32x4 r1 = philox4x32_10({(uint32_t)t, step, RNG_INJECT2 + 2 * which, emitter}, k0, k1);
    const u32x4 r2 = philox4x32_10({(uint32_t)t, step, RNG_INJECT2 + 2 * which + 1, emitter}, k0, k1);
    const u32x4 r3 = philox4x32_10({(uint32_t)t, step, RNG_INJECT3 + which, emitter}, k0, k1);
    const double sx = u01_53(r1.a, r1.b) + u01_53(r1.c, r1.d) + u01_53(r2.a, r2.b) - 1.5;
    const double sy = u01_53(r2.c, r2.d) + u01_53(r3.a, r3.b) + u01_53(r3.c, r3.d) - 1.5;
    const double vxp = (vth_sqrt2 * sx + e.vdx * vel_norm) / vel_norm;
    const double vyp = (vth_sqrt2 * sy + e.vdy * vel_norm) / vel_norm;
    const int64_t q = pslot<P>(slot);
    x[q] = (P)px; y[q] = (P)py; vx[q] = (P)vxp; vy[q] = (P)vyp; vz[slot] = (P)0;
    id[slot] = id0 + (uint32_t)t;
    if (Q) {
      const double lx = (px - x0) / dx, ly = (py - y0) / dy;
      const int i = (int)lx, j = (int)ly;
      if (i < 0 || i + 1 >= nx || j < 0 || j + 1 >= ny) {
        atomicAdd(&errc[EPIC_ERRC_DEPOSIT_OOB], 1ull);
        continue;
      }
      const double di = lx - i, dj = ly - j, sc = ldexp(1.0, fix_bits);
      const long long one = 1ll << fix_bits;
      const long long q00 = (long long)rint((1 - di) * (1 - dj) * sc);
      const long long q10 = (long long)rint(di * (1 - dj) * sc);
      const long long q01 = (long long)rint((1 - di) * dj * sc);
      const long long q11 = one - q00 - q10 - q01;
      unsigned long long* b = Q + (size_t)i * ny + j;
      atomicAdd(b, (unsigned long long)q00);
      atomicAdd(b + ny, (unsigned long long)q10);
      atomicAdd(b + 1, (unsigned long long)q01);
      atomicAdd(b + ny + 1, (unsigned long long)q11);
    }
  }
}

__global__ void bump_count_kernel(int64_t* d_n, int64_t k, int64_t cap) {
  d_n[0] = std::min(d_n[0] + k, cap);
  d_n[1] += k;
}

hipError_t launch_inject(Ctx* c, int step) {
  const epic_domain_desc& d = c->dom;
  const uint32_t k0 = (uint32_t)d.seed, k1 = (uint32_t)(d.seed >> 32);
  for (size_t ei = 0; ei < c->emitters.size(); ei++) {
    const epic_emitter_desc& e = c->emitters[ei].d;
    const int targets[2] = {e.species_idx1, e.species_idx2};
    const int nt = (e.species_idx1 != e.species_idx2) ? 2 : 1;
    for (int w = 0; w < nt; w++) {
      const int s = targets[w];
      SpeciesDev& S = c->sp[s];
      // Init::SampleVel(species, temp): v_th = sqrt(2 K_b T EV_to_K / m)
      const double v_th = sqrt(2 * K_b * e.temp * (QE / K_b) / S.desc.mass);
      const double vts = v_th * sqrt(2);
      const int k = e.numparticle;
      const int grid = std::max(1, std::min((k + 255) / 256, 1024));
      unsigned long long* Q = c->deposit_valid ? S.Q : nullptr;
      if (d.precision == EPIC_FP64)
        hipLaunchKernelGGL(inject_kernel<double>, dim3(grid), dim3(256), 0, c->stream, (double*)S.x, (double*)S.y,
                           (double*)S.vx, (double*)S.vy, (double*)S.vz, S.id, S.d_n, S.cap, k, e, vts,
                           d.vel_norm, (uint32_t)step, (uint32_t)ei, (uint32_t)w, k0, k1, Q, d.nx, d.ny, d.x0, d.y0,
                           d.dx, d.dy, c->fix_bits, c->errc);
      else
        hipLaunchKernelGGL(inject_kernel<float>, dim3(grid), dim3(256), 0, c->stream, (float*)S.x, (float*)S.y,
                           (float*)S.vx, (float*)S.vy, (float*)S.vz, S.id, S.d_n, S.cap, k, e, vts,
                           d.vel_norm, (uint32_t)step, (uint32_t)ei, (uint32_t)w, k0, k1, Q, d.nx, d.ny, d.x0, d.y0,
                           d.dx, d.dy, c->fix_bits, c->errc);
      EPIC_HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(bump_count_kernel, dim3(1), dim3(1), 0, c->stream, S.d_n, (int64_t)k, S.cap);
      EPIC_HIP_TRY(hipGetLastError());
      S.n_host = std::min<int64_t>(S.n_host + k, S.cap);
    }
  }
  return hipSuccess;
}

}  // namespace epic
