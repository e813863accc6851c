// Batched ensemble: many independent simulations stepped together on one GPU.
//
// The reference runs one simulation per process (main.cpp:24-36, 395-455).
// Here a batch owns any number of member simulations (each an ordinary
// context on the batch's stream, so every per-simulation entry point works
// on a member) and advances all of them with one launch per stage:
//
//   finalize   Q -> den, rho, Poisson right-hand side b (FP64), X = (float) b
//   solve      per correction: X <- A^-1 X by sine transforms (rocBLAS FP32
//              GEMMs, one call per distinct grid size for the whole group),
//              phi (+)= X, r = b - A phi in FP64 and X = (float) r
//   norm       final ||r|| / ||b|| per member (not-converged counter)
//   efield     E = -grad phi, the FP32 gather copy, and Q = 0
//   push       gather + leapfrog + walls + deposit over a (segment, chunk)
//              work list: every species of every member in one persistent
//              launch, through the same wavefront routine as the
//              single-simulation FP32 push
//   collide    every member's null-collision passes in one launch, births in
//              a second
//
// Node stages walk a (member, tile) list: each workgroup owns NB consecutive
// nodes of one member.  Every kernel reads its per-member and per-species
// descriptors from device arrays, so a buffer swap (a member's sort) only
// re-uploads descriptors; the step itself can be captured once in a HIP
// graph and replayed.  Per-member step counters (the MCC random-stream key)
// live on the device and are advanced by the last kernel of the step.
//
// Per-node and per-particle arithmetic is shared with the single-simulation
// kernels (node_ops.hpp, push_f32.hpp).  Kernel shapes (particles per thread
// and occupancy target of the push, NB of the node stages) do not change any
// result: the deposit is an integer sum and each node value is computed by
// one thread.  They are chosen by measurement (epic_batch_tune).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gemm_tile.hpp"
#include "node_ops.hpp"
#include "push_f32.hpp"
#include "sort_ops.hpp"

namespace epic {

struct BSim {
  double* rho;
  double* phi;
  double* efx;
  double* efy;
  double* b;
  float2* E2;
  float4* Ec;      // FP32 gather table (cell_coeffs)
  float* X;        // this member's P*Q block of the group's transform workspace
  const float* G;  // group's spectral scaling c / (lam_k + mu_l), P*Q
  SpecLayout L;
  int nx, ny, bc, op;
  int sp0, nsp;      // species records [sp0, sp0 + nsp)
  int tile0, ntile;  // node tiles [tile0, tile0 + ntile)
  double dx, dy, dx2, s, norm, tol;
  unsigned long long* errc;
};

struct BSpecies {
  unsigned long long* Q;
  double* den;
  double scale;   // 2^-B spwt
  double factor;  // charge / QE
};

struct BTile {
  int sim, k0;
};

// ---- node stages -----------------------------------------------------------
__global__ void fill_f32_kernel(float* p, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0.f;
}

__global__ void bfinalize_kernel(const BSim* sims, const BSpecies* sps, const BTile* tiles, double* part_b) {
  __shared__ double sh[16];
  const BTile t = tiles[blockIdx.x];
  const BSim& S = sims[t.sim];
  const int nn = S.nx * S.ny;
  const int k = t.k0 + threadIdx.x;
  double bb = 0;
  if (k < nn) {
    const int i = k / S.ny, j = k % S.ny;
    double rho = 0;
    for (int s = S.sp0; s < S.sp0 + S.nsp; s++) {
      const BSpecies& P = sps[s];
      const long long q = (long long)fold_q(P.Q, i, j, S.nx, S.ny, S.bc);
      if (q < 0) atomicAdd(&S.errc[EPIC_ERRC_FIXED_OVF], 1ull);
      const double den = den_from_q(q, P.scale, S.norm, i, j, S.nx, S.ny, S.bc);
      P.den[k] = den;
      rho += den * P.factor;
    }
    S.rho[k] = rho;
    const double b = rhs_at(rho, i, j, S.nx, S.ny, S.dx2, S.s);
    S.b[k] = b;
    const int x = spec_index(S.L, k);
    if (x >= 0) S.X[x] = (float)b;
    bb = b * b;
  }
  double v[1] = {bb};
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part_b[blockIdx.x] = v[0];
}

__global__ void bunpack_kernel(const BSim* sims, const BTile* tiles, int first) {
  const BTile t = tiles[blockIdx.x];
  const BSim& S = sims[t.sim];
  const int k = t.k0 + threadIdx.x;
  if (k >= S.nx * S.ny) return;
  const int x = spec_index(S.L, k);
  if (first) S.phi[k] = x >= 0 ? (double)S.X[x] : 0.0;
  else if (x >= 0) S.phi[k] = S.phi[k] + (double)S.X[x];
}

__global__ void bresidual_kernel(const BSim* sims, const BTile* tiles, double* part_r) {
  __shared__ double sh[16];
  const BTile t = tiles[blockIdx.x];
  const BSim& S = sims[t.sim];
  const int k = t.k0 + threadIdx.x;
  double rr = 0;
  if (k < S.nx * S.ny) {
    const double r = residual_at(S.op, S.b, S.phi, k, S.nx, S.ny);
    const int x = spec_index(S.L, k);
    if (x >= 0) S.X[x] = (float)r;
    rr = r * r;
  }
  double v[1] = {rr};
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part_r[blockIdx.x] = v[0];
}

// One wavefront per member: ||r|| / ||b|| from the tile partials.
__global__ void bnorm_kernel(const BSim* sims, const double* part_b, const double* part_r, double* relres) {
  const BSim& S = sims[blockIdx.x];
  double b = 0, r = 0;
  for (int t = threadIdx.x; t < S.ntile; t += 64) {
    b += part_b[S.tile0 + t];
    r += part_r[S.tile0 + t];
  }
  b = wave_sum(b);
  r = wave_sum(r);
  if (threadIdx.x == 0) {
    const double rel = b > 0 ? sqrt(r / b) : 0.0;
    relres[blockIdx.x] = rel;
    if (!(rel <= S.tol)) atomicAdd(&S.errc[EPIC_ERRC_NOCONV], 1ull);  // NaN counts too
  }
}

__global__ void befield_kernel(const BSim* sims, const BSpecies* sps, const BTile* tiles) {
  const BTile t = tiles[blockIdx.x];
  const BSim& S = sims[t.sim];
  const int k = t.k0 + threadIdx.x;
  if (k >= S.nx * S.ny) return;
  double ex, ey;
  efield_at(S.phi, k / S.ny, k % S.ny, S.nx, S.ny, S.dx, S.dy, S.bc, &ex, &ey);
  S.efx[k] = ex;
  S.efy[k] = ey;
  S.E2[k] = make_float2((float)ex, (float)ey);
  for (int s = S.sp0; s < S.sp0 + S.nsp; s++) sps[s].Q[k] = 0ull;
}

// The FP32 gather table of every member (the node tile's cells; the last
// node row and column own no cell).
__global__ void bcell_kernel(const BSim* sims, const BTile* tiles) {
  const BTile t = tiles[blockIdx.x];
  const BSim& S = sims[t.sim];
  const int k = t.k0 + threadIdx.x;
  if (k >= S.nx * S.ny) return;
  const int i = k / S.ny, j = k % S.ny;
  if (i >= S.nx - 1 || j >= S.ny - 1) return;
  cell_coeffs(S.E2, i, j, S.ny, S.Ec + 2 * (size_t)(i * (S.ny - 1) + j));
}

__global__ void bstep_kernel(int* steps, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) steps[i] += 1;
}

// ---- particle stage --------------------------------------------------------
// Work item = one workgroup chunk (PART_BLOCK * PPT particles) of one
// segment (a species of a member); wavefront w of the workgroup takes the
// w-th 64*PPT run of it.  Chunks past a segment's live slot count are
// skipped, so the list can cover capacity (births, injection).  The
// descriptors are read-only and unaliased (__restrict__): that lets the
// compiler fetch the work item and its segment with scalar loads after the
// kernel's own particle stores; without it they arrive in vector registers
// (92 VGPRs for the 4-particle shape, 5 waves per SIMD, instead of 64 and 8).
template <int PPT, int WAVES, int BC>
__global__ __launch_bounds__(PART_BLOCK, WAVES) void push_batch_kernel(const PushF32* __restrict__ segs,
                                                                        const int2* __restrict__ chunks, int nchunks) {
  __shared__ unsigned long long lds[(PART_BLOCK / 64) * WAVE_WIN];
  constexpr int CHW = 64 * PPT;
  constexpr int CH = PART_BLOCK * PPT;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long* wl = lds + wid * WAVE_WIN;
  // round-robin walk, software-pipelined as the single-simulation kernel:
  // the next item's particles (full chunks) are loaded before this item's
  // deposit
  int c = blockIdx.x;
  if (c >= nchunks) return;
  int2 ch = chunks[c];
  PushF32 a = segs[ch.x];
  int64_t n = *a.d_n;
  int64_t wb = (int64_t)ch.y * CH + (int64_t)wid * CHW;
  ChunkRegs<PPT> r;
  if (wb < n) load_chunk<PPT>(a, wb, n, lane, r);
  for (; c < nchunks; c += gridDim.x) {
    ChunkDep<PPT> dp;
    const bool active = wb < n;
    if (active) push_chunk_regs<true, PPT, false, BC>(a, wb, n, lane, r, dp);
    const int cn = c + (int)gridDim.x;
    PushF32 an = a;
    int64_t nn = 0, wbn = 0;
    bool full = false;
    if (cn < nchunks) {
      const int2 chn = chunks[cn];
      an = segs[chn.x];
      nn = *an.d_n;
      wbn = (int64_t)chn.y * CH + (int64_t)wid * CHW;
      full = wbn + 64 * PPT <= nn;
      if (full) load_chunk_full<PPT>(an, wbn, lane, r);
    }
    if (active) deposit_chunk<PPT, BC == EPIC_BC_PBC>(a, lane, wl, dp);
    if (cn < nchunks && !full && wbn < nn) load_chunk<PPT>(an, wbn, nn, lane, r);
    a = an;
    n = nn;
    wb = wbn;
  }
}

// ---- batched counting sort ------------------------------------------------------
// One species of each member whose sort is due, all sorted together with the
// routines of the single-simulation sort (sort_ops.hpp); members' own sort
// workspaces, so species go in rounds (species 0 of every due member, then
// species 1, ...).  Work items (particle chunks of SORT_CHUNK over capacity,
// groups of 4 cells) find their segment by binary search over per-segment
// prefix offsets, so no host synchronisation sizes any launch.
struct SortSeg {
  const float *x, *y, *vx, *vy, *vz;
  const uint32_t* id;
  float *x2, *y2, *vx2, *vy2, *vz2;
  uint32_t* id2;
  int64_t* d_n;
  int* count;  // K + 2: histogram, then start
  int* fill;   // K + 2
  uint32_t* perm;
  uint32_t* scratch;
  float x0, y0, dx, dy;
  KeyMap km;
  int64_t chunk0;  // first particle chunk of this segment
  int64_t item0;   // first group of 4 * FIX_GROUP cells of this segment
};

__device__ __forceinline__ int find_seg(const SortSeg* segs, int nseg, int64_t v, bool cells) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t b = cells ? segs[mid].item0 : segs[mid].chunk0;
    if (b <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void bsort_zero_kernel(const SortSeg* segs) {
  const SortSeg& S = segs[blockIdx.y];
  const int n = S.km.K + 2;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    S.count[k] = 0;
    S.fill[k] = 0;
  }
}

__global__ __launch_bounds__(256) void bsort_hist_kernel(const SortSeg* segs, int nseg, int64_t nchunks) {
  __shared__ int win[HWIN];
  __shared__ int wbase_s;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const SortSeg& S = segs[find_seg(segs, nseg, c, false)];
    const int64_t n = *S.d_n, base = (c - S.chunk0) * SORT_CHUNK;
    if (base < n) hist_chunk(S.x, S.y, n, base, S.count, S.x0, S.y0, S.dx, S.dy, S.km, win, &wbase_s, S.scratch);
  }
}

// Exclusive scan of count[0..K] in place, one workgroup per segment:
// count[k] = start of key k, count[K] = live total, count[K+1] = all.
__global__ __launch_bounds__(1024) void bsort_scan_kernel(const SortSeg* segs) {
  __shared__ int sh[16];
  __shared__ int carry_s;
  const SortSeg& S = segs[blockIdx.x];
  const int m = S.km.K + 1;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int t0 = 0; t0 < m; t0 += 4 * 1024) {
    int v[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = t0 + threadIdx.x * 4 + q;
      v[q] = k < m ? S.count[k] : 0;
      sum += v[q];
    }
    int inc = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(inc, off, 64);
      if (lane >= off) inc += o;
    }
    if (lane == 63) sh[wid] = inc;
    __syncthreads();
    int wb = 0, tot = 0;
    for (int w = 0; w < 16; w++) {
      if (w < wid) wb += sh[w];
      tot += sh[w];
    }
    int run = carry_s + wb + inc - sum;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = t0 + threadIdx.x * 4 + q;
      if (k < m) S.count[k] = run;
      run += v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) carry_s += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) S.count[m] = carry_s;
}

__global__ __launch_bounds__(256) void bsort_fill_kernel(const SortSeg* segs, int nseg, int64_t nchunks) {
  __shared__ int win[HWIN];
  __shared__ int wbase_s;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const SortSeg& S = segs[find_seg(segs, nseg, c, false)];
    const int64_t n = *S.d_n, base = (c - S.chunk0) * SORT_CHUNK;
    if (base < n) fill_chunk(S.scratch, n, base, S.count, S.fill, S.perm, S.km, win, &wbase_s);
  }
}

__global__ __launch_bounds__(256) void bsort_fix_kernel(const SortSeg* segs, int nseg, int64_t nitems) {
  __shared__ __attribute__((aligned(16))) uint32_t buf[4][FIX_LDS + 4];
  const int w = threadIdx.x >> 6;
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const SortSeg& S = segs[find_seg(segs, nseg, it, true)];
    const int c0 = ((int)(it - S.item0) * 4 + w) * FIX_GROUP;
    if (c0 < S.km.K) fix_cells(S.count, S.perm, S.scratch, c0, S.km.K, buf[w]);
  }
}

// Each XCD walks a contiguous range of chunks with one workgroup per CU, as
// the single-simulation gather does (sort_gather_kernel in aux.hip).
__global__ __launch_bounds__(256) void bsort_gather_kernel(const SortSeg* segs, int nseg, int64_t nchunks) {
  int64_t cb, ce, cs;
  xcd_chunk_range(nchunks, &cb, &ce, &cs);
  for (int64_t c = cb; c < ce; c += cs) {
    const SortSeg& S = segs[find_seg(segs, nseg, c, false)];
    const int64_t n = S.count[S.km.K], base = (c - S.chunk0) * SORT_CHUNK;
    if (base < n)
      gather_chunk(S.perm, n, base, S.x, S.y, S.vx, S.vy, S.vz, S.id, S.x2, S.y2, S.vx2, S.vy2, S.vz2, S.id2);
  }
}

__global__ void bsort_setn_kernel(const SortSeg* segs, int nseg) {
  for (int k = threadIdx.x; k < nseg; k += blockDim.x) segs[k].d_n[0] = (int64_t)segs[k].count[segs[k].km.K];
}

struct PushShape {
  int ppt, waves;
};
static const PushShape kPushShapes[EPIC_BATCH_NPUSH] = {{4, 8}, {8, 1}, {16, 1}, {4, 1}, {8, 5}};
static const int kNodeBlocks[EPIC_BATCH_NNODE] = {64, 128, 256, 512};

template <int PPT, int W, int BC>
static hipError_t launch_push_batch_t(hipStream_t st, int num_cu, const PushF32* segs, const int2* chunks,
                                      int nchunks) {
  auto k = push_batch_kernel<PPT, W, BC>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, PART_BLOCK, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  int grid = num_cu * per_cu;
  // no more workgroups than chunks, rounded to whole XCD rounds
  if (grid > nchunks) grid = std::max(8, (nchunks + 7) / 8 * 8);
  hipLaunchKernelGGL(k, dim3(grid), dim3(PART_BLOCK), 0, st, segs, chunks, nchunks);
  return hipGetLastError();
}

template <int BC>
static hipError_t launch_push_batch_bc(int variant, hipStream_t st, int num_cu, const PushF32* segs,
                                       const int2* chunks, int nchunks) {
  switch (variant) {
    case 1: return launch_push_batch_t<8, 1, BC>(st, num_cu, segs, chunks, nchunks);
    case 2: return launch_push_batch_t<16, 1, BC>(st, num_cu, segs, chunks, nchunks);
    case 3: return launch_push_batch_t<4, 1, BC>(st, num_cu, segs, chunks, nchunks);
    case 4: return launch_push_batch_t<8, 5, BC>(st, num_cu, segs, chunks, nchunks);
    default: return launch_push_batch_t<4, 8, BC>(st, num_cu, segs, chunks, nchunks);
  }
}

// bc: the wall kind every member shares, or BC_ANY when they differ (the
// kernel then reads it per chunk)
static hipError_t launch_push_batch(int variant, int bc, hipStream_t st, int num_cu, const PushF32* segs,
                                    const int2* chunks, int nchunks) {
  if (nchunks <= 0) return hipSuccess;
  switch (bc) {
    case EPIC_BC_PBC: return launch_push_batch_bc<EPIC_BC_PBC>(variant, st, num_cu, segs, chunks, nchunks);
    case EPIC_BC_OPEN: return launch_push_batch_bc<EPIC_BC_OPEN>(variant, st, num_cu, segs, chunks, nchunks);
    case EPIC_BC_REFLECTIVE:
      return launch_push_batch_bc<EPIC_BC_REFLECTIVE>(variant, st, num_cu, segs, chunks, nchunks);
    default: return launch_push_batch_bc<BC_ANY>(variant, st, num_cu, segs, chunks, nchunks);
  }
}

// ---- host side ---------------------------------------------------------------
void spec_destroy(Ctx* c);

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes, bool* grown = nullptr) {
    if (grown) *grown = false;
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 2, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    if (grown) *grown = e == hipSuccess;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Sine-transform matrices of one distinct (p, q), zero-padded to P x Q
// (multiples of 64, the tile products' shapes), and the workspace of the
// members that share it (P*Q per member, its padding kept zero).
struct Group {
  int p = 0, q = 0;
  int P = 0, Q = 0;
  int nb = 0;  // members in the current layout
  float* Sp = nullptr;
  float* Sq = nullptr;
  float* G = nullptr;
  DevBuf X, Y;
};

struct Member {
  epic_ctx* h = nullptr;
  int id = -1;
  int64_t steps = 0;
  bool init_pending = false;  // next step is the initial field + half-step rewind
};

struct Batch {
  hipStream_t stream = nullptr;
  int device = 0, num_cu = 0;
  epic_batch_opts opts{};
  int push_variant = 0;
  int push_bc = BC_ANY;  // the members' common wall kind (BC_ANY: mixed)
  int node_block = 256;
  std::vector<Member> members;  // slot order
  int next_id = 0;
  std::vector<Group> groups;

  // descriptors (host copies of what is on the device)
  std::vector<BSim> h_sims;
  std::vector<BSpecies> h_sps;
  std::vector<BTile> h_tiles;
  std::vector<PushF32> h_segs;
  std::vector<int2> h_chunks;
  std::vector<unsigned char> h_mcc, h_app;
  std::vector<int> h_steps;
  std::vector<unsigned char> uploaded;  // concatenation last uploaded
  DevBuf d_sims, d_sps, d_tiles, d_segs, d_chunks, d_mcc, d_app, d_steps, d_part_b, d_part_r, d_relres;
  int n_mcc = 0, n_app = 0;
  bool layout_dirty = true;
  int built_variant = -1, built_nb = -1;  // shapes the tile and chunk lists were built for
  std::vector<long long> shape_key;  // launch dimensions the graph was captured with

  DevBuf d_sortsegs[4];                    // one per species round
  std::vector<SortSeg> h_sortsegs[4];

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_ok = false;
  bool warm = false;  // one eager step ran with the current layout

  // statistics
  int64_t steps = 0, replays = 0, captures = 0;
  bool timing = false;
  hipEvent_t ev[6] = {};
  double stage_ms[4] = {};
  double ms_total = 0;
  std::vector<double> relres_host;

  std::string err;
};

static Ctx* member_ctx(const Member& m) { return ctx_of(m.h); }

// The single-simulation solver's layout (spectral.hip), padded the same way.
static SpecLayout layout_of(const epic_domain_desc& d) {
  if (d.poisson_op == EPIC_OP_DIRICHLET_WALL)
    return SpecLayout{d.ny - 2, d.nx - 2, d.ny, d.ny + 1, spec_pad32(d.ny - 2), spec_pad32(d.nx - 2)};
  return SpecLayout{d.nx, d.ny, d.nx, 0, spec_pad32(d.nx), spec_pad32(d.ny)};
}

// FP32 sine matrices and scaling, computed in FP64 and rounded (the FP32
// copies of the single-simulation solver), zero outside p x p, q x q, p x q.
static hipError_t group_init(Group& g) {
  const int p = g.p, q = g.q, P = g.P, Q = g.Q;
  std::vector<float> sp((size_t)P * P, 0.f), sq((size_t)Q * Q, 0.f), gg((size_t)P * Q, 0.f);
  for (int a = 0; a < p; a++)
    for (int k = 0; k < p; k++) sp[a + (size_t)k * P] = (float)sin(M_PI * (double)(a + 1) * (k + 1) / (p + 1));
  for (int a = 0; a < q; a++)
    for (int k = 0; k < q; k++) sq[a + (size_t)k * Q] = (float)sin(M_PI * (double)(a + 1) * (k + 1) / (q + 1));
  const double cc = (2.0 / (p + 1)) * (2.0 / (q + 1));
  for (int k = 0; k < p; k++)
    for (int l = 0; l < q; l++) {
      const double lam = 2 * cos(M_PI * (k + 1) / (p + 1)) - 2 + 2 * cos(M_PI * (l + 1) / (q + 1)) - 2;
      gg[k + (size_t)l * P] = (float)(cc / lam);
    }
  EPIC_HIP_TRY(hipMalloc(&g.Sp, sp.size() * 4));
  EPIC_HIP_TRY(hipMalloc(&g.Sq, sq.size() * 4));
  EPIC_HIP_TRY(hipMalloc(&g.G, gg.size() * 4));
  EPIC_HIP_TRY(hipMemcpy(g.Sp, sp.data(), sp.size() * 4, hipMemcpyHostToDevice));
  EPIC_HIP_TRY(hipMemcpy(g.Sq, sq.data(), sq.size() * 4, hipMemcpyHostToDevice));
  EPIC_HIP_TRY(hipMemcpy(g.G, gg.data(), gg.size() * 4, hipMemcpyHostToDevice));
  return hipSuccess;
}

static void group_free(Group& g) {
  if (g.Sp) (void)hipFree(g.Sp);
  if (g.Sq) (void)hipFree(g.Sq);
  if (g.G) (void)hipFree(g.G);
  g.Sp = g.Sq = g.G = nullptr;
  g.X.release();
  g.Y.release();
}

// Assign members to transform groups and size the group workspaces.
static hipError_t build_groups(Batch* B, std::vector<int>& grp, std::vector<int>& slot_in_grp) {
  for (Group& g : B->groups) g.nb = 0;
  grp.assign(B->members.size(), -1);
  slot_in_grp.assign(B->members.size(), 0);
  for (size_t m = 0; m < B->members.size(); m++) {
    const SpecLayout L = layout_of(member_ctx(B->members[m])->dom);
    int gi = -1;
    for (size_t k = 0; k < B->groups.size(); k++)
      if (B->groups[k].p == L.p && B->groups[k].q == L.q) gi = (int)k;
    if (gi < 0) {
      Group g;
      g.p = L.p;
      g.q = L.q;
      g.P = L.pp;
      g.Q = L.qq;
      EPIC_HIP_TRY(group_init(g));
      B->groups.push_back(g);
      gi = (int)B->groups.size() - 1;
    }
    grp[m] = gi;
    slot_in_grp[m] = B->groups[gi].nb++;
  }
  for (Group& g : B->groups) {
    if (g.nb == 0) continue;
    const size_t bytes = (size_t)g.P * g.Q * g.nb * sizeof(float);
    // growing frees the old workspace: the previous step must be done with it
    if (bytes > g.X.cap || bytes > g.Y.cap) EPIC_HIP_TRY(hipStreamSynchronize(B->stream));
    bool gx = false, gy = false;
    EPIC_HIP_TRY(g.X.ensure(bytes, &gx));
    EPIC_HIP_TRY(g.Y.ensure(bytes, &gy));
    // a new workspace starts zero (its padding stays zero through the
    // products); a fill kernel, not a memset, in case a step is captured
    if (gx) {
      hipLaunchKernelGGL(fill_f32_kernel, dim3(1024), dim3(256), 0, B->stream, (float*)g.X.p, g.X.cap / 4);
      EPIC_HIP_TRY(hipGetLastError());
    }
    if (gy) {
      hipLaunchKernelGGL(fill_f32_kernel, dim3(1024), dim3(256), 0, B->stream, (float*)g.Y.p, g.Y.cap / 4);
      EPIC_HIP_TRY(hipGetLastError());
    }
  }
  return hipSuccess;
}

// Host descriptors of the current members (pointers are re-read every step:
// a member's sort swaps its particle buffers).
static hipError_t build_descriptors(Batch* B, bool full) {
  std::vector<int> grp, sig;
  EPIC_HIP_TRY(build_groups(B, grp, sig));
  const int NB = B->node_block;
  const int ppt = kPushShapes[B->push_variant].ppt;
  const int64_t CH = (int64_t)PART_BLOCK * ppt;
  B->h_sims.clear();
  B->h_sps.clear();
  B->h_segs.clear();
  if (full) {
    B->h_tiles.clear();
    B->h_chunks.clear();
  }
  B->h_mcc.clear();
  B->h_app.clear();
  B->push_bc = B->members.empty() ? BC_ANY : member_ctx(B->members[0])->dom.bc;
  for (size_t m = 0; m < B->members.size(); m++) {
    Ctx* c = member_ctx(B->members[m]);
    const epic_domain_desc& d = c->dom;
    const Group& g = B->groups[grp[m]];
    if (d.bc != B->push_bc) B->push_bc = BC_ANY;
    BSim S{};
    S.rho = c->rho; S.phi = c->phi; S.efx = c->efx; S.efy = c->efy; S.b = c->b64;
    S.E2 = (float2*)c->E2;
    S.Ec = (float4*)c->Ec;
    S.X = (float*)g.X.p + (size_t)sig[m] * g.P * g.Q;
    S.G = g.G;
    S.L = layout_of(d);
    S.nx = d.nx; S.ny = d.ny; S.bc = d.bc; S.op = d.poisson_op;
    S.sp0 = (int)B->h_sps.size();
    S.nsp = c->nsp;
    const int nn = d.nx * d.ny;
    S.tile0 = m == 0 ? 0 : B->h_sims[m - 1].tile0 + B->h_sims[m - 1].ntile;
    S.ntile = (nn + NB - 1) / NB;
    S.dx = d.dx; S.dy = d.dy;
    S.dx2 = d.dx * d.dx;
    S.s = (d.L * d.L) / (d.LDe * d.LDe);
    const double cell_area = d.dx * d.dy;
    const double L2 = d.L * d.L;
    S.norm = (cell_area * L2 * d.density);
    S.tol = d.tol_relative ? d.tolerance : 1e-12;
    S.errc = c->errc;
    B->h_sims.push_back(S);
    if (full)
      for (int k0 = 0; k0 < nn; k0 += NB) B->h_tiles.push_back(BTile{(int)m, k0});
    for (int s = 0; s < c->nsp; s++) {
      SpeciesDev& P = c->sp[s];
      B->h_sps.push_back(BSpecies{P.Q, P.den, ldexp(1.0, -c->fix_bits) * P.desc.spwt, P.desc.charge * (1 / QE)});
      double kv = 0, kx = 0;
      if (B->members[m].init_pending) {
        // Rewind_species (species.cpp:229-246, every species): v -= 1/2 (q/m) E DT,
        // positions kept (and deposited again for the first step)
        push_f32_consts(c, s, &kv, &kx);
        kv = -0.5 * kv;
        kx = 0.0;
      } else if (P.desc.pushed) {
        push_f32_consts(c, s, &kv, &kx);
      }
      const int seg = (int)B->h_segs.size();
      // in place: the descriptors (and the captured step) keep their pointers
      B->h_segs.push_back(make_push_f32(c, s, kv, kx, false));
      if (full) {
        const int64_t nch = (P.cap + CH - 1) / CH;
        for (int64_t k = 0; k < nch; k++) B->h_chunks.push_back(make_int2(seg, (int)k));
      }
    }
    if (!B->members[m].init_pending) mcc_batch_collect(c, (int)m, B->h_mcc, B->h_app);
  }
  B->n_mcc = (int)(B->h_mcc.size() / mcc_args_bytes());
  B->n_app = (int)(B->h_app.size() / append_args_bytes());
  return hipSuccess;
}

template <typename T>
static void append_bytes(std::vector<unsigned char>& v, const std::vector<T>& a) {
  const unsigned char* p = (const unsigned char*)a.data();
  v.insert(v.end(), p, p + a.size() * sizeof(T));
}

template <typename T>
static hipError_t upload(DevBuf& d, const std::vector<T>& h, hipStream_t st) {
  EPIC_HIP_TRY(d.ensure(std::max<size_t>(h.size() * sizeof(T), 16)));
  if (h.empty()) return hipSuccess;
  return hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

// Rebuild descriptors; upload what changed.  The graph stays valid unless a
// launch dimension changed.
static hipError_t refresh(Batch* B) {
  const bool full = B->layout_dirty || B->built_variant != B->push_variant || B->built_nb != B->node_block;
  EPIC_HIP_TRY(build_descriptors(B, full));
  std::vector<long long> key = {(long long)B->members.size(), (long long)B->h_tiles.size(),
                                (long long)B->h_segs.size(), (long long)B->h_chunks.size(), B->n_mcc, B->n_app,
                                B->push_variant, B->node_block, B->opts.refine};
  for (const Group& g : B->groups) {
    key.push_back(g.nb);
    key.push_back((long long)(uintptr_t)g.X.p);
  }
  // per-step check of what can change between steps: buffer pointers
  std::vector<unsigned char> small;
  append_bytes(small, B->h_sims);
  append_bytes(small, B->h_sps);
  append_bytes(small, B->h_segs);
  small.insert(small.end(), B->h_mcc.begin(), B->h_mcc.end());
  small.insert(small.end(), B->h_app.begin(), B->h_app.end());
  const bool dims_changed = key != B->shape_key;
  if (!full && !dims_changed && small == B->uploaded) return hipSuccess;
  // the previous step may still read the descriptor arrays
  EPIC_HIP_TRY(hipStreamSynchronize(B->stream));
  const void* before[] = {B->d_sims.p, B->d_sps.p, B->d_tiles.p, B->d_segs.p, B->d_chunks.p,
                          B->d_mcc.p, B->d_app.p, B->d_steps.p, B->d_part_b.p, B->d_part_r.p, B->d_relres.p};
  EPIC_HIP_TRY(upload(B->d_sims, B->h_sims, B->stream));
  EPIC_HIP_TRY(upload(B->d_sps, B->h_sps, B->stream));
  EPIC_HIP_TRY(upload(B->d_segs, B->h_segs, B->stream));
  EPIC_HIP_TRY(upload(B->d_mcc, B->h_mcc, B->stream));
  EPIC_HIP_TRY(upload(B->d_app, B->h_app, B->stream));
  if (full) {
    EPIC_HIP_TRY(upload(B->d_tiles, B->h_tiles, B->stream));
    EPIC_HIP_TRY(upload(B->d_chunks, B->h_chunks, B->stream));
  }
  if (B->layout_dirty) {
    B->h_steps.clear();
    // the step's last kernel advances every counter; an initial-field step
    // must leave its member's counter where it was
    for (const Member& m : B->members) B->h_steps.push_back(member_ctx(m)->step_count - (m.init_pending ? 1 : 0));
    EPIC_HIP_TRY(upload(B->d_steps, B->h_steps, B->stream));
  }
  EPIC_HIP_TRY(B->d_part_b.ensure(std::max<size_t>(B->h_tiles.size(), 1) * sizeof(double)));
  EPIC_HIP_TRY(B->d_part_r.ensure(std::max<size_t>(B->h_tiles.size(), 1) * sizeof(double)));
  EPIC_HIP_TRY(B->d_relres.ensure(std::max<size_t>(B->members.size(), 1) * sizeof(double)));
  EPIC_HIP_TRY(hipStreamSynchronize(B->stream));
  const void* after[] = {B->d_sims.p, B->d_sps.p, B->d_tiles.p, B->d_segs.p, B->d_chunks.p,
                         B->d_mcc.p, B->d_app.p, B->d_steps.p, B->d_part_b.p, B->d_part_r.p, B->d_relres.p};
  const bool moved = memcmp(before, after, sizeof(before)) != 0;
  B->uploaded.swap(small);
  if (dims_changed || moved || full) {
    B->graph_ok = false;
    B->warm = false;
  }
  B->shape_key = key;
  B->layout_dirty = false;
  B->built_variant = B->push_variant;
  B->built_nb = B->node_block;
  return hipSuccess;
}

// X <- S_p X S_q (times G elementwise when scale) for every member of group
// g (in place through Y), on the tile products: the same arithmetic for a
// member whatever the group's size, and the same as the single-simulation
// solver's.
static hipError_t group_transform(Batch* B, Group& g, bool scale) {
  const int P = g.P, Q = g.Q;
  float* X = (float*)g.X.p;
  float* Y = (float*)g.Y.p;
  const SpecLayout none{};
  // Y = S_p [X_1 ... X_nb]: one product over the members' concatenated columns
  const TileGemm t1{g.Sp, P, X, P, Y, P, nullptr, nullptr, none, 0, 0, 0, 0};
  EPIC_HIP_TRY(gemm_tile32(B->stream, P, Q * g.nb, P, t1));
  // X_m = Y_m S_q (o G), one launch for the group's members (blockIdx.z)
  const TileGemm t2{Y, P, g.Sq, Q, X, P, scale ? g.G : nullptr, nullptr, none, 0, (size_t)P * Q, 0, (size_t)P * Q};
  EPIC_HIP_TRY(gemm_tile32(B->stream, P, Q, Q, t2, g.nb));
  return hipSuccess;
}

static void rec(Batch* B, int k) {
  if (B->timing) (void)hipEventRecord(B->ev[k], B->stream);
}

// One step of every member, enqueued on the batch stream (eager or under
// capture).  Stage events only when timing eager steps.
static hipError_t enqueue_step(Batch* B, bool events) {
  const int nt = (int)B->h_tiles.size();
  const int NB = B->node_block;
  const BSim* sims = (const BSim*)B->d_sims.p;
  const BSpecies* sps = (const BSpecies*)B->d_sps.p;
  const BTile* tiles = (const BTile*)B->d_tiles.p;
  double* pb = (double*)B->d_part_b.p;
  double* pr = (double*)B->d_part_r.p;
  const bool ev = events && B->timing;
  if (ev) rec(B, 0);
  if (nt > 0) {
    hipLaunchKernelGGL(bfinalize_kernel, dim3(nt), dim3(NB), 0, B->stream, sims, sps, tiles, pb);
    EPIC_HIP_TRY(hipGetLastError());
    const int K = B->opts.refine;
    for (int k = 0; k < K; k++) {
      for (Group& g : B->groups)
        if (g.nb) EPIC_HIP_TRY(group_transform(B, g, true));
      for (Group& g : B->groups)
        if (g.nb) EPIC_HIP_TRY(group_transform(B, g, false));
      hipLaunchKernelGGL(bunpack_kernel, dim3(nt), dim3(NB), 0, B->stream, sims, tiles, k == 0 ? 1 : 0);
      EPIC_HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(bresidual_kernel, dim3(nt), dim3(NB), 0, B->stream, sims, tiles, pr);
      EPIC_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(bnorm_kernel, dim3((int)B->members.size()), dim3(64), 0, B->stream, sims, pb, pr,
                       (double*)B->d_relres.p);
    EPIC_HIP_TRY(hipGetLastError());
    if (ev) rec(B, 1);
    hipLaunchKernelGGL(befield_kernel, dim3(nt), dim3(NB), 0, B->stream, sims, sps, tiles);
    EPIC_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(bcell_kernel, dim3(nt), dim3(NB), 0, B->stream, sims, tiles);
    EPIC_HIP_TRY(hipGetLastError());
  }
  if (ev) rec(B, 2);
  EPIC_HIP_TRY(launch_push_batch(B->push_variant, B->push_bc, B->stream, B->num_cu, (const PushF32*)B->d_segs.p,
                                 (const int2*)B->d_chunks.p, (int)B->h_chunks.size()));
  if (ev) rec(B, 3);
  EPIC_HIP_TRY(launch_mcc_batch(B->stream, B->num_cu, B->d_mcc.p, B->n_mcc, B->d_app.p, B->n_app,
                                (const int*)B->d_steps.p));
  hipLaunchKernelGGL(bstep_kernel, dim3(1), dim3(256), 0, B->stream, (int*)B->d_steps.p, (int)B->members.size());
  EPIC_HIP_TRY(hipGetLastError());
  if (ev) rec(B, 4);
  return hipSuccess;
}

static void destroy_graph(Batch* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  B->graph_ok = false;
}

static hipError_t capture(Batch* B) {
  destroy_graph(B);
  EPIC_HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue_step(B, false);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(B->stream, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return e;
  }
  EPIC_HIP_TRY(e2);
  B->graph = g;
  EPIC_HIP_TRY(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  B->graph_ok = true;
  B->captures++;
  return hipSuccess;
}

// Per-member work outside the batched step: first deposit after a load,
// emitter injection, and (after the step) sorts on each member's schedule.
static hipError_t pre_step(Batch* B) {
  for (Member& m : B->members) {
    Ctx* c = member_ctx(m);
    if (c->blas) {
      // the initial-field solver's matrices are not needed again
      EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
      spec_destroy(c);
    }
    if (!c->emitters.empty()) EPIC_HIP_TRY(launch_inject(c, c->step_count));
    if (!c->deposit_valid) {
      EPIC_HIP_TRY(launch_zero_Q(c));
      for (int s = 0; s < c->nsp; s++) EPIC_HIP_TRY(launch_deposit(c, s));
      c->deposit_valid = true;
    }
  }
  return hipSuccess;
}

// Sorts species `round` of the due members' (one launch per sort stage).
static hipError_t batch_sort_round(Batch* B, const std::vector<Ctx*>& due, int round) {
  std::vector<SortSeg>& hs = B->h_sortsegs[round];
  hs.clear();
  int64_t chunks = 0, items = 0;
  int kmax = 0;
  for (Ctx* c : due) {
    if (round >= c->nsp) continue;
    SpeciesDev& P = c->sp[round];
    const epic_domain_desc& d = c->dom;
    SortSeg g{};
    g.x = (const float*)P.x; g.y = (const float*)P.y; g.vx = (const float*)P.vx; g.vy = (const float*)P.vy;
    g.vz = (const float*)P.vz; g.id = P.id;
    g.x2 = (float*)P.x2; g.y2 = (float*)P.y2; g.vx2 = (float*)P.vx2; g.vy2 = (float*)P.vy2;
    g.vz2 = (float*)P.vz2; g.id2 = P.id2;
    g.d_n = P.d_n;
    g.count = c->sort_count;
    g.fill = c->sort_fill;
    g.perm = c->sort_perm;
    g.scratch = c->sort_perm + c->sort_cap;
    g.x0 = (float)d.x0; g.y0 = (float)d.y0; g.dx = (float)d.dx; g.dy = (float)d.dy;
    g.km = key_map(d.nx - 1, d.ny - 1);
    g.chunk0 = chunks;
    g.item0 = items;
    chunks += (P.cap + SORT_CHUNK - 1) / SORT_CHUNK;
    items += (g.km.K + 4 * FIX_GROUP - 1) / (4 * FIX_GROUP);
    kmax = std::max(kmax, g.km.K);
    hs.push_back(g);
  }
  const int nseg = (int)hs.size();
  if (nseg == 0) return hipSuccess;
  DevBuf& db = B->d_sortsegs[round];
  EPIC_HIP_TRY(db.ensure(hs.size() * sizeof(SortSeg)));
  EPIC_HIP_TRY(hipMemcpyAsync(db.p, hs.data(), hs.size() * sizeof(SortSeg), hipMemcpyHostToDevice, B->stream));
  const SortSeg* S = (const SortSeg*)db.p;
  const int gpart = (int)std::min<int64_t>(chunks, (int64_t)B->num_cu * 16);
  hipLaunchKernelGGL(bsort_zero_kernel, dim3(std::min((kmax + 2 + 255) / 256, 256), nseg), dim3(256), 0, B->stream, S);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bsort_hist_kernel, dim3(gpart), dim3(256), 0, B->stream, S, nseg, chunks);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bsort_scan_kernel, dim3(nseg), dim3(1024), 0, B->stream, S);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bsort_fill_kernel, dim3(gpart), dim3(256), 0, B->stream, S, nseg, chunks);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bsort_fix_kernel, dim3((int)std::min<int64_t>(items, 65536)), dim3(256), 0, B->stream, S, nseg,
                     items);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bsort_gather_kernel, dim3((int)std::min<int64_t>(chunks, (int64_t)B->num_cu)), dim3(256), 0,
                     B->stream, S, nseg, chunks);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bsort_setn_kernel, dim3(1), dim3(256), 0, B->stream, S, nseg);
  EPIC_HIP_TRY(hipGetLastError());
  for (Ctx* c : due) {
    if (round >= c->nsp) continue;
    SpeciesDev& P = c->sp[round];
    std::swap(P.x, P.x2); std::swap(P.y, P.y2); std::swap(P.vx, P.vx2);
    std::swap(P.vy, P.vy2); std::swap(P.vz, P.vz2); std::swap(P.id, P.id2);
  }
  return hipSuccess;
}

static hipError_t post_step(Batch* B) {
  std::vector<Ctx*> due;
  int nsp_max = 0;
  for (Member& m : B->members) {
    Ctx* c = member_ctx(m);
    c->deposit_valid = true;
    if (m.init_pending) {
      m.init_pending = false;
      continue;
    }
    c->step_count++;
    m.steps++;
    bool any = false;
    for (int s = 0; s < c->nsp; s++) any = any || c->sort_due(s);
    if (any) {
      due.push_back(c);
      nsp_max = std::max(nsp_max, c->nsp);
    }
  }
  if (!due.empty()) {
    // the previous round's host records may still be read by a pending copy
    EPIC_HIP_TRY(hipStreamSynchronize(B->stream));
    std::vector<Ctx*> due_r;
    for (int r = 0; r < nsp_max && r < 4; r++) {
      due_r.clear();
      for (Ctx* c : due)
        if (r < c->nsp && c->sort_due(r)) due_r.push_back(c);
      EPIC_HIP_TRY(batch_sort_round(B, due_r, r));
    }
    for (Ctx* c : due)
      if (c->nsp > 4) {  // beyond four species: one by one
        EPIC_HIP_TRY(sync_counts(c));
        for (int r = 4; r < c->nsp; r++)
          if (c->sort_due(r)) EPIC_HIP_TRY(launch_sort(c, r));
      }
  }
  B->steps++;
  return hipSuccess;
}

static hipError_t resolve_stage_events(Batch* B) {
  EPIC_HIP_TRY(hipEventSynchronize(B->ev[4]));
  float t[4];
  for (int k = 0; k < 4; k++) EPIC_HIP_TRY(hipEventElapsedTime(&t[k], B->ev[k], B->ev[k + 1]));
  // stage 0: finalize + solve + norm; 1: efield; 2: push; 3: collide
  for (int k = 0; k < 4; k++) B->stage_ms[k] += t[k];
  B->ms_total += t[0] + t[1] + t[2] + t[3];
  return hipSuccess;
}

static hipError_t one_step(Batch* B, bool force_eager, double* push_ms, double* node_ms) {
  EPIC_HIP_TRY(pre_step(B));
  EPIC_HIP_TRY(refresh(B));
  if (B->members.empty()) return hipSuccess;
  const bool graph = B->opts.use_graph && !force_eager;
  if (graph && B->warm) {
    if (!B->graph_ok) EPIC_HIP_TRY(capture(B));
    if (B->timing) EPIC_HIP_TRY(hipEventRecord(B->ev[0], B->stream));
    EPIC_HIP_TRY(hipGraphLaunch(B->exec, B->stream));
    if (B->timing) {
      EPIC_HIP_TRY(hipEventRecord(B->ev[5], B->stream));
      EPIC_HIP_TRY(hipEventSynchronize(B->ev[5]));
      float t = 0;
      EPIC_HIP_TRY(hipEventElapsedTime(&t, B->ev[0], B->ev[5]));
      B->ms_total += t;
    }
    B->replays++;
  } else {
    const bool want = B->timing || push_ms || node_ms;
    const bool saved = B->timing;
    B->timing = want;
    hipError_t e = enqueue_step(B, want);
    if (e == hipSuccess && want) {
      e = hipEventSynchronize(B->ev[4]);
      float t[4] = {0, 0, 0, 0};
      for (int k = 0; k < 4 && e == hipSuccess; k++) e = hipEventElapsedTime(&t[k], B->ev[k], B->ev[k + 1]);
      if (e == hipSuccess) {
        if (saved) {
          for (int k = 0; k < 4; k++) B->stage_ms[k] += t[k];
          B->ms_total += t[0] + t[1] + t[2] + t[3];
        }
        if (push_ms) *push_ms = t[2];
        if (node_ms) *node_ms = t[0] + t[1];
      }
    }
    B->timing = saved;
    EPIC_HIP_TRY(e);
    B->warm = true;
  }
  return post_step(B);
}

}  // namespace epic

struct epic_batch {
  epic::Batch b;
};

using namespace epic;

static int bfail(Batch* B, int code, const std::string& msg) {
  if (B) B->err = msg;
  return code;
}

#define BCK(expr, where)                                                                         \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) return bfail(B, EPIC_ERR_HIP, std::string(where) + ": " + hipGetErrorString(_e)); \
  } while (0)

static int find_member(Batch* B, int id) {
  for (size_t m = 0; m < B->members.size(); m++)
    if (B->members[m].id == id) return (int)m;
  return -1;
}

extern "C" {

int epic_batch_create(const epic_batch_opts* opts, epic_batch** out) {
  if (!out) return EPIC_ERR_ARG;
  *out = nullptr;
  epic_batch_opts o{};
  o.refine = 2;
  o.use_graph = 1;
  o.push_variant = -1;
  o.node_block = 0;
  if (opts) o = *opts;
  if (o.refine <= 0) o.refine = 2;
  if (o.refine > 8) return EPIC_ERR_ARG;
  if (o.push_variant >= EPIC_BATCH_NPUSH) return EPIC_ERR_ARG;
  if (o.node_block != 0 && o.node_block != 64 && o.node_block != 128 && o.node_block != 256 && o.node_block != 512)
    return EPIC_ERR_ARG;
  epic_batch* h = new (std::nothrow) epic_batch();
  if (!h) return EPIC_ERR_CAPACITY;
  Batch* B = &h->b;
  B->opts = o;
  B->push_variant = o.push_variant >= 0 ? o.push_variant : 0;
  B->node_block = o.node_block ? o.node_block : 256;
  hipDeviceProp_t prop;
  if (hipGetDevice(&B->device) != hipSuccess || hipGetDeviceProperties(&prop, B->device) != hipSuccess ||
      hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return EPIC_ERR_HIP;
  }
  B->num_cu = prop.multiProcessorCount;
  for (auto& e : B->ev) (void)hipEventCreate(&e);
  *out = h;
  return EPIC_OK;
}

int epic_batch_destroy(epic_batch* h) {
  if (!h) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  (void)hipStreamSynchronize(B->stream);
  destroy_graph(B);
  for (Member& m : B->members) epic_destroy(m.h);
  B->members.clear();
  for (Group& g : B->groups) group_free(g);
  DevBuf* bufs[] = {&B->d_sims, &B->d_sps, &B->d_tiles, &B->d_segs, &B->d_chunks, &B->d_mcc,
                    &B->d_app, &B->d_steps, &B->d_part_b, &B->d_part_r, &B->d_relres};
  for (DevBuf* d : bufs) d->release();
  for (DevBuf& d : B->d_sortsegs) d.release();
  for (auto& e : B->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(B->stream);
  delete h;
  return EPIC_OK;
}

const char* epic_batch_last_error(const epic_batch* h) { return h ? h->b.err.c_str() : "null batch"; }

int epic_batch_add(epic_batch* h, const epic_domain_desc* dom, const epic_species_desc* sp, int nsp, int* id) {
  if (!h || !dom || !sp || !id) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  if (dom->precision != EPIC_FP32) return bfail(B, EPIC_ERR_ARG, "batch members store particles in FP32");
  if (dom->shape != EPIC_SHAPE_CIC) return bfail(B, EPIC_ERR_ARG, "batch members use CIC weighting");
  if (dom->solver != EPIC_SOLVER_PCG_MIXED && dom->solver != EPIC_SOLVER_SPECTRAL)
    return bfail(B, EPIC_ERR_ARG, "batch members solve with the mixed-precision sine-transform solver");
  if (dom->poisson_op == EPIC_OP_DIRICHLET_WALL && (dom->nx < 4 || dom->ny < 4))
    return bfail(B, EPIC_ERR_ARG, "grid too small");
  if (dom->nx > 65535 || dom->ny > 65535) return bfail(B, EPIC_ERR_ARG, "grid sides up to 65535 nodes");
  epic_ctx* m = nullptr;
  const int rc = ctx_create(dom, sp, nsp, B->stream, &m);
  if (rc != EPIC_OK) return bfail(B, rc, "member creation failed");
  Member mem;
  mem.h = m;
  mem.id = B->next_id++;
  B->members.push_back(mem);
  B->layout_dirty = true;
  *id = mem.id;
  return EPIC_OK;
}

int epic_batch_member(epic_batch* h, int id, epic_ctx** ctx) {
  if (!h || !ctx) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  const int m = find_member(B, id);
  if (m < 0) return bfail(B, EPIC_ERR_ARG, "no such member");
  *ctx = B->members[m].h;
  return EPIC_OK;
}

int epic_batch_remove(epic_batch* h, int id) {
  if (!h) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  const int m = find_member(B, id);
  if (m < 0) return bfail(B, EPIC_ERR_ARG, "no such member");
  BCK(hipStreamSynchronize(B->stream), "sync");
  epic_destroy(B->members[m].h);
  B->members.erase(B->members.begin() + m);
  B->layout_dirty = true;
  return EPIC_OK;
}

int epic_batch_init(epic_batch* h, int id) {
  if (!h) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  const int m = find_member(B, id);
  if (m < 0) return bfail(B, EPIC_ERR_ARG, "no such member");
  if (member_ctx(B->members[m])->step_count != 0 || B->members[m].steps != 0)
    return bfail(B, EPIC_ERR_STATE, "member has already stepped");
  B->members[m].init_pending = true;
  member_ctx(B->members[m])->deposit_valid = false;
  B->layout_dirty = true;
  return EPIC_OK;
}

int epic_batch_step(epic_batch* h, int nsteps) {
  if (!h || nsteps < 0) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  for (int k = 0; k < nsteps; k++) BCK(one_step(B, false, nullptr, nullptr), "batch step");
  return EPIC_OK;
}

int epic_batch_sync(epic_batch* h) {
  if (!h) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  BCK(hipStreamSynchronize(B->stream), "sync");
  return EPIC_OK;
}

int epic_batch_tune(epic_batch* h, int steps_per_candidate, epic_batch_tuning* out) {
  if (!h || steps_per_candidate < 1) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  if (B->members.empty()) return bfail(B, EPIC_ERR_STATE, "empty batch");
  for (const Member& m : B->members)
    if (m.init_pending) BCK(one_step(B, true, nullptr, nullptr), "initial field");
  epic_batch_tuning t{};
  for (int v = 0; v < EPIC_BATCH_NPUSH; v++) t.push_ms[v] = -1;
  for (int v = 0; v < EPIC_BATCH_NNODE; v++) t.node_ms[v] = -1;
  // every candidate runs real steps: kernel shapes do not change results
  int best_v = B->push_variant;
  double best = 1e300;
  for (int v = 0; v < EPIC_BATCH_NPUSH; v++) {
    B->push_variant = v;
    double sum = 0, ms = 0;
    BCK(one_step(B, true, &ms, nullptr), "tune");  // first launch of the variant
    for (int k = 0; k < steps_per_candidate; k++) {
      BCK(one_step(B, true, &ms, nullptr), "tune");
      sum += ms;
    }
    t.push_ms[v] = sum / steps_per_candidate;
    if (t.push_ms[v] < best) {
      best = t.push_ms[v];
      best_v = v;
    }
  }
  B->push_variant = best_v;
  int best_nb = B->node_block;
  best = 1e300;
  for (int v = 0; v < EPIC_BATCH_NNODE; v++) {
    B->node_block = kNodeBlocks[v];
    double sum = 0, ms = 0;
    BCK(one_step(B, true, nullptr, &ms), "tune");
    for (int k = 0; k < steps_per_candidate; k++) {
      BCK(one_step(B, true, nullptr, &ms), "tune");
      sum += ms;
    }
    t.node_ms[v] = sum / steps_per_candidate;
    if (t.node_ms[v] < best) {
      best = t.node_ms[v];
      best_nb = kNodeBlocks[v];
    }
  }
  B->node_block = best_nb;
  t.push_variant = best_v;
  t.node_block = best_nb;
  t.steps_used = (1 + steps_per_candidate) * (EPIC_BATCH_NPUSH + EPIC_BATCH_NNODE);
  if (out) *out = t;
  return EPIC_OK;
}

int epic_batch_relres(epic_batch* h, int id, double* relres) {
  if (!h || !relres) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  const int m = find_member(B, id);
  if (m < 0) return bfail(B, EPIC_ERR_ARG, "no such member");
  if (B->steps == 0 || B->layout_dirty) return bfail(B, EPIC_ERR_STATE, "no step since the last membership change");
  BCK(hipMemcpyAsync(relres, (double*)B->d_relres.p + m, sizeof(double), hipMemcpyDeviceToHost, B->stream), "relres");
  BCK(hipStreamSynchronize(B->stream), "relres");
  return EPIC_OK;
}

int epic_batch_member_steps(epic_batch* h, int id, int64_t* steps) {
  if (!h || !steps) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  const int m = find_member(B, id);
  if (m < 0) return bfail(B, EPIC_ERR_ARG, "no such member");
  *steps = B->members[m].steps;
  return EPIC_OK;
}

int epic_batch_get_stats(epic_batch* h, epic_batch_stats* out) {
  if (!h || !out) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  epic_batch_stats s{};
  s.members = (int)B->members.size();
  for (const Group& g : B->groups) s.groups += g.nb > 0 ? 1 : 0;
  s.tiles = (int)B->h_tiles.size();
  s.segments = (int)B->h_segs.size();
  s.chunks = (int)B->h_chunks.size();
  s.mcc_passes = B->n_mcc;
  s.push_variant = B->push_variant;
  s.node_block = B->node_block;
  s.steps = B->steps;
  s.graph_replays = B->replays;
  s.graph_captures = B->captures;
  s.ms_total = B->ms_total;
  for (int k = 0; k < 4; k++) s.stage_ms[k] = B->stage_ms[k];
  int64_t bytes = 0;
  for (const Member& m : B->members) bytes += (int64_t)member_ctx(m)->bytes;
  for (const Group& g : B->groups)
    bytes += (int64_t)(g.X.cap + g.Y.cap) + 4ll * ((int64_t)g.P * g.P + (int64_t)g.Q * g.Q + (int64_t)g.P * g.Q);
  s.device_bytes = bytes;
  s.relres_max = 0;
  if (!B->members.empty() && B->steps > 0 && !B->layout_dirty) {
    std::vector<double> r(B->members.size());
    BCK(hipMemcpyAsync(r.data(), B->d_relres.p, r.size() * sizeof(double), hipMemcpyDeviceToHost, B->stream),
        "stats");
    BCK(hipStreamSynchronize(B->stream), "stats");
    for (double v : r) s.relres_max = std::max(s.relres_max, v);
  }
  *out = s;
  return EPIC_OK;
}

int epic_batch_configure(epic_batch* h, const epic_batch_opts* opts) {
  if (!h || !opts) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  if (opts->refine < 0 || opts->refine > 8 || opts->push_variant >= EPIC_BATCH_NPUSH) return EPIC_ERR_ARG;
  if (opts->node_block != 0 && opts->node_block != 64 && opts->node_block != 128 && opts->node_block != 256 &&
      opts->node_block != 512)
    return EPIC_ERR_ARG;
  B->opts.refine = opts->refine > 0 ? opts->refine : 2;
  B->opts.use_graph = opts->use_graph;
  if (opts->push_variant >= 0) B->push_variant = opts->push_variant;
  if (opts->node_block) B->node_block = opts->node_block;
  B->opts.push_variant = opts->push_variant;
  B->opts.node_block = opts->node_block;
  B->graph_ok = false;  // launch shapes may have changed
  return EPIC_OK;
}

int epic_batch_timing_enable(epic_batch* h, int on) {
  if (!h) return EPIC_ERR_ARG;
  Batch* B = &h->b;
  B->timing = on != 0;
  for (int k = 0; k < 4; k++) B->stage_ms[k] = 0;
  B->ms_total = 0;
  return EPIC_OK;
}

}  // extern "C"
