// Counting sort by cell: the per-chunk and per-cell routines, shared by the
// single-simulation kernels (aux.hip) and the batched ones (batch.hip).
//
// The sort orders particles by a tile-major cell key (16 x 16-cell tiles in
// row-major order, cells row-major inside a tile) and, inside a cell, by
// their previous storage index: a stable sort, equal element for element to
// std::stable_sort by key on the CPU.  Tiles keep a particle that moved a
// cell or two since the previous sort within a few thousand slots of where it
// was, so the permutation is local.  Absorbed particles get the key K (the
// key space size) and are dropped, so a sort also compacts.
//
// Stages: histogram of keys -> exclusive scan (start) -> slot reservation
// (perm: destination slot -> source index; order inside a wavefront kept,
// order between wavefronts arbitrary) -> per-cell fix (each cell's sources
// ascending) -> gather into the second buffer.  The histogram keeps its
// wavefront-sorted (key, lane) words, from which the reservation works
// without computing the keys or sorting again.
#pragma once

#include "device_common.hpp"
#include "epic_internal.hpp"

namespace epic {

constexpr int SORT_TILE = 16;
// Particles per thread per pass of the sort kernels: the loads of SORT_U
// particles are in flight together; a workgroup chunk is 256 * SORT_U.
constexpr int SORT_U = 8;
constexpr int SORT_CHUNK = 256 * SORT_U;

struct KeyMap {
  int ncx, ncy;  // cells
  int nty;       // tiles along y
  int K;         // key space: tiles * SORT_TILE^2; K is the key of absorbed particles
};

inline KeyMap key_map(int ncx, int ncy) {
  KeyMap m;
  m.ncx = ncx;
  m.ncy = ncy;
  const int ntx = (ncx + SORT_TILE - 1) / SORT_TILE;
  m.nty = (ncy + SORT_TILE - 1) / SORT_TILE;
  m.K = ntx * m.nty * SORT_TILE * SORT_TILE;
  return m;
}

// A live position maps to a cell; one whose logical coordinate rounds onto
// the upper boundary node belongs to the last cell (as in the push's
// deposit), so the sort never drops it.  Positions outside the grid and dead
// (NaN) ones get the key K and are dropped.
template <typename P>
__device__ __forceinline__ int cell_key(P x, P y, P x0, P y0, P dx, P dy, const KeyMap& m) {
  if (isnan(x)) return m.K;
  const P lx = (x - x0) / dx, ly = (y - y0) / dy;
  if (!(lx >= 0) || !(ly >= 0) || !(lx <= (P)m.ncx) || !(ly <= (P)m.ncy)) return m.K;
  const int i = min((int)lx, m.ncx - 1), j = min((int)ly, m.ncy - 1);
  return ((i / SORT_TILE) * m.nty + j / SORT_TILE) * (SORT_TILE * SORT_TILE) + (i % SORT_TILE) * SORT_TILE +
         j % SORT_TILE;
}

// Runs of equal keys among consecutive lanes: the run's first lane does one
// atomic for the whole run.
struct LaneRun {
  int head;  // lane that starts this lane's run
  int len;   // run length (valid in the head lane)
};
__device__ __forceinline__ LaneRun lane_run(int key) {
  const int lane = threadIdx.x & 63;
  // the previous lane's key by a DPP whole-wave shift (wave_shr:1); lane 0
  // receives a value different from its own key
  const int prev = __builtin_amdgcn_update_dpp(~key, key, 0x138, 0xf, 0xf, false);
  const bool h = lane == 0 || key != prev;
  const unsigned long long heads = __ballot(h);
  const unsigned long long upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
  LaneRun r;
  r.head = 63 - __builtin_clzll(upto);
  const unsigned long long after = lane == 63 ? 0ull : (heads >> (lane + 1));
  r.len = (after ? __builtin_ctzll(after) + 1 : 64 - lane);
  return r;
}

// Lane-xor exchange.  Inside a 16-lane row the exchange is a DPP operand
// (no LDS round trip): xor 1 and xor 2 are quad permutations, xor 8 a
// rotation of the row by 8, xor 4 the row shifted down 4 for lanes with bit
// 2 clear and up 4 for the others; xor 16 is ds_swizzle's xor mode inside
// 32-lane halves, xor 32 a permute across them.
template <int J>
__device__ __forceinline__ int xor_lane(int v) {
  if (J == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, true);   // quad_perm [1,0,3,2]
  if (J == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, true);   // quad_perm [2,3,0,1]
  if (J == 8) return __builtin_amdgcn_mov_dpp(v, 0x128, 0xf, 0xf, true);  // row_ror:8
  if (J == 4) {
    // row_shl:4 reads lane + 4 (0 past the row's end: those lanes take the other)
    const int from_above = __builtin_amdgcn_mov_dpp(v, 0x104, 0xf, 0xf, true);
    const int from_below = __builtin_amdgcn_mov_dpp(v, 0x114, 0xf, 0xf, true);  // row_shr:4, lane - 4
    return (threadIdx.x & 4) ? from_below : from_above;
  }
  if (J < 32) return __builtin_amdgcn_ds_swizzle(v, (J << 10) | 0x1F);
  return __shfl_xor(v, J, 64);
}

// Ascending bitonic sort of one int per lane across the 64-lane wavefront.
__device__ __forceinline__ int wave_sort(int v) {
  const int lane = threadIdx.x & 63;
#define EPIC_BSTEP(K, J)                                       \
  {                                                            \
    const int o = xor_lane<J>(v);                              \
    const bool asc = (lane & K) == 0, lower = (lane & J) == 0; \
    v = (lower == asc) ? min(v, o) : max(v, o);                \
  }
  EPIC_BSTEP(2, 1)
  EPIC_BSTEP(4, 2) EPIC_BSTEP(4, 1)
  EPIC_BSTEP(8, 4) EPIC_BSTEP(8, 2) EPIC_BSTEP(8, 1)
  EPIC_BSTEP(16, 8) EPIC_BSTEP(16, 4) EPIC_BSTEP(16, 2) EPIC_BSTEP(16, 1)
  EPIC_BSTEP(32, 16) EPIC_BSTEP(32, 8) EPIC_BSTEP(32, 4) EPIC_BSTEP(32, 2) EPIC_BSTEP(32, 1)
  EPIC_BSTEP(64, 32) EPIC_BSTEP(64, 16) EPIC_BSTEP(64, 8) EPIC_BSTEP(64, 4) EPIC_BSTEP(64, 2) EPIC_BSTEP(64, 1)
#undef EPIC_BSTEP
  return v;
}

// Histogram of one workgroup chunk (256*SORT_U particles from base;
// block-uniform call).  The chunk counts into an LDS window of HWIN keys
// centred on its first key (after a wavefront sort, equal keys add once),
// then adds the window's non-zero counts to the global histogram: particles
// that moved less than a tile since the previous sort cost no global atomic.
// Keys outside the window go straight to the global histogram.
//
// The wavefront sorts its (key, lane) words (key * 64 + lane; key spaces
// below 2^25, multiples of 256, so no word reaches PK_NONE; slots past the
// end as PK_NONE) and stores them at the chunk's own slots of pks (those
// below n), for the slot reservation.
// HWIN covers a chunk's keys (2048 slots span a few dozen cells, and a
// particle that changed cell within its tile moved its key by at most a
// tile's 256) with a window that is cheap to clear and flush per chunk.
constexpr int HWIN = 1024;
constexpr int PK_NONE = 0x7fffffff;

template <typename P>
__device__ __forceinline__ void hist_chunk(const P* x, const P* y, int64_t n, int64_t base, int* count, P x0, P y0,
                                           P dx, P dy, const KeyMap& m, int* win, int* wbase_s, uint32_t* pks) {
  const int lane = threadIdx.x & 63;
  // every load first, from in-range slots (past the end: the last live one,
  // not used), so that all 2 SORT_U are in flight together
  P xv[SORT_U], yv[SORT_U];
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int64_t p = base + (int64_t)u * blockDim.x + threadIdx.x;
    const int64_t q = pslot<P>(p < n ? p : n - 1);
    xv[u] = x[q];
    yv[u] = y[q];
  }
  int key[SORT_U];
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int64_t p = base + (int64_t)u * blockDim.x + threadIdx.x;
    key[u] = p < n ? cell_key(xv[u], yv[u], x0, y0, dx, dy, m) : -1;
  }
  if (threadIdx.x == 0) *wbase_s = key[0] >= 0 && key[0] < m.K ? max(0, key[0] - HWIN / 2) : 0;
  for (int t = threadIdx.x; t < HWIN; t += blockDim.x) win[t] = 0;
  __syncthreads();
  const int wbase = *wbase_s;
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int pk = wave_sort(key[u] >= 0 ? key[u] * 64 + lane : PK_NONE);
    const int64_t p = base + (int64_t)u * blockDim.x + threadIdx.x;
    // a group's slots past the end hold its PK_NONE words (they sort last)
    if (p < n) pks[p] = (uint32_t)pk;
    const int ks = pk == PK_NONE ? -1 : pk >> 6;
    const LaneRun r = lane_run(ks);
    if (r.head == lane && ks >= 0) {
      const int loc = ks - wbase;
      if (loc >= 0 && loc < HWIN) atomicAdd(&win[loc], r.len);
      else atomicAdd(&count[ks], r.len);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < HWIN; t += blockDim.x) {
    const int v = win[t];
    if (v) atomicAdd(&count[wbase + t], v);
  }
  __syncthreads();
}

// Slots of one chunk from the histogram's sorted (key, lane) words: equal
// keys form one run in lane (= source) order, which takes len consecutive
// slots.  The runs of keys inside an LDS window of HWIN keys (centred on the
// chunk's first key, as the histogram's) take their offsets inside the
// chunk's block of that key with an LDS atomic; one global atomic per key
// then reserves the chunk's block.  Runs of keys outside the window reserve
// their slots with a global atomic each.  Absorbed particles (key K) are
// not placed.  (Block-uniform call; win and wbase_s in LDS.)
__device__ __forceinline__ void fill_chunk(const uint32_t* pks, int64_t n, int64_t base, const int* start, int* fill,
                                           uint32_t* perm, const KeyMap& m, int* win, int* wbase_s) {
  const int dead = m.K;
  const int lane = threadIdx.x & 63;
  int pk[SORT_U];
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int64_t p = base + (int64_t)u * blockDim.x + threadIdx.x;
    const int v = (int)pks[p < n ? p : n - 1];
    pk[u] = p < n ? v : PK_NONE;
  }
  if (threadIdx.x == 0) {
    const int k0 = pk[0] == PK_NONE ? -1 : pk[0] >> 6;
    *wbase_s = k0 >= 0 && k0 < m.K ? max(0, k0 - HWIN / 2) : 0;
  }
  for (int t = threadIdx.x; t < HWIN; t += blockDim.x) win[t] = 0;
  __syncthreads();
  const int wbase = *wbase_s;
  int off[SORT_U];
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int key = pk[u] == PK_NONE ? -1 : pk[u] >> 6;
    const int ks = key != dead ? key : -1;
    const LaneRun r = lane_run(ks);
    int o = 0;
    if (r.head == lane && ks >= 0) {
      const int loc = ks - wbase;
      o = loc >= 0 && loc < HWIN ? atomicAdd(&win[loc], r.len) : atomicAdd(&fill[ks], r.len);
    }
    off[u] = o;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < HWIN; t += blockDim.x) {
    const int c = win[t];
    if (c) win[t] = atomicAdd(&fill[wbase + t], c);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int key = pk[u] == PK_NONE ? -1 : pk[u] >> 6;
    const int ks = key != dead ? key : -1;
    const LaneRun r = lane_run(ks);
    const int o = __shfl(off[u], r.head, 64);
    if (ks >= 0) {
      const int loc = ks - wbase;
      const int blk = loc >= 0 && loc < HWIN ? win[loc] : 0;
      const int64_t p = base + (int64_t)u * blockDim.x + (threadIdx.x & ~63) + (pk[u] & 63);
      perm[start[ks] + blk + o + (lane - r.head)] = (uint32_t)p;
    }
  }
  __syncthreads();
}

// One cell's sources ascending (one wavefront): rank counting (rank = number
// of smaller indices; indices are unique).  Cells of up to FIX_LDS particles
// are staged in LDS (buf), larger ones in scratch (same offsets as perm).
constexpr int FIX_LDS = 1024;
constexpr int FIX_Q = FIX_LDS / 64;

// Ranks of a staged cell of m <= 64 * CNT indices; four staged indices per
// LDS read (the staging pads to a multiple of 4 with all-ones, which is
// never smaller).
template <int CNT>
__device__ __forceinline__ void fix_rank_lds(uint32_t* perm, int b, int m, const uint32_t* buf) {
  const int lane = threadIdx.x & 63;
  uint32_t mine[CNT];
  int rank[CNT];
#pragma unroll
  for (int q = 0; q < CNT; q++) {
    const int k = lane + q * 64;
    mine[q] = k < m ? buf[k] : 0xFFFFFFFFu;
    rank[q] = 0;
  }
  for (int t = 0; t < m; t += 4) {
    const uint4 o = *reinterpret_cast<const uint4*>(buf + t);
#pragma unroll
    for (int q = 0; q < CNT; q++)
      rank[q] += (o.x < mine[q] ? 1 : 0) + (o.y < mine[q] ? 1 : 0) + (o.z < mine[q] ? 1 : 0) + (o.w < mine[q] ? 1 : 0);
  }
#pragma unroll
  for (int q = 0; q < CNT; q++) {
    const int k = lane + q * 64;
    if (k < m) perm[b + rank[q]] = mine[q];
  }
}

// Sort one cell known to be out of order (buf: FIX_LDS + 4 entries, 16-byte
// aligned).
__device__ __forceinline__ void fix_unordered(uint32_t* perm, uint32_t* scratch, int b, int m, uint32_t* buf) {
  const int lane = threadIdx.x & 63;
  if (m <= FIX_LDS) {
    for (int k = lane; k < m; k += 64) buf[k] = perm[b + k];
    if (lane < 4) buf[m + lane] = 0xFFFFFFFFu;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int cnt = (m + 63) / 64;
    if (cnt == 1) fix_rank_lds<1>(perm, b, m, buf);
    else if (cnt == 2) fix_rank_lds<2>(perm, b, m, buf);
    else if (cnt <= 4) fix_rank_lds<4>(perm, b, m, buf);
    else if (cnt <= 8) fix_rank_lds<8>(perm, b, m, buf);
    else fix_rank_lds<FIX_Q>(perm, b, m, buf);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    for (int k = lane; k < m; k += 64) scratch[b + k] = perm[b + k];
    __threadfence();
    __builtin_amdgcn_wave_barrier();
    for (int k0 = 0; k0 < m; k0 += 64) {
      const int k = k0 + lane;
      const uint32_t me = k < m ? __builtin_nontemporal_load(&scratch[b + k]) : 0xFFFFFFFFu;
      int rk = 0;
      for (int t = 0; t < m; t++) rk += (__builtin_nontemporal_load(&scratch[b + t]) < me) ? 1 : 0;
      if (k < m) perm[b + rk] = me;
    }
    __threadfence();
    __builtin_amdgcn_wave_barrier();
  }
}

// Ascending bitonic sort of one unsigned per lane across the wavefront.
__device__ __forceinline__ uint32_t wave_sort_u(uint32_t v) {
  const int lane = threadIdx.x & 63;
#define EPIC_BSTEP_U(K, J)                                     \
  {                                                            \
    const uint32_t o = (uint32_t)xor_lane<J>((int)v);          \
    const bool asc = (lane & K) == 0, lower = (lane & J) == 0; \
    v = (lower == asc) ? min(v, o) : max(v, o);                \
  }
  EPIC_BSTEP_U(2, 1)
  EPIC_BSTEP_U(4, 2) EPIC_BSTEP_U(4, 1)
  EPIC_BSTEP_U(8, 4) EPIC_BSTEP_U(8, 2) EPIC_BSTEP_U(8, 1)
  EPIC_BSTEP_U(16, 8) EPIC_BSTEP_U(16, 4) EPIC_BSTEP_U(16, 2) EPIC_BSTEP_U(16, 1)
  EPIC_BSTEP_U(32, 16) EPIC_BSTEP_U(32, 8) EPIC_BSTEP_U(32, 4) EPIC_BSTEP_U(32, 2) EPIC_BSTEP_U(32, 1)
  EPIC_BSTEP_U(64, 32) EPIC_BSTEP_U(64, 16) EPIC_BSTEP_U(64, 8) EPIC_BSTEP_U(64, 4) EPIC_BSTEP_U(64, 2)
  EPIC_BSTEP_U(64, 1)
#undef EPIC_BSTEP_U
  return v;
}

// Set bits of m strictly below position p (0 <= p <= 64).
__device__ __forceinline__ int popc_below(unsigned long long m, int p) {
  return p >= 64 ? (int)__popcll(m) : (int)__popcll(m & ((1ull << p) - 1));
}

// Ranks of one cell of m <= 64 * CNT entries (v: entry t in lane t % 64,
// register t / 64; entries past m all-ones) from its runs.  An entry's source
// lies in the 64-slot group (source / 64) of the wavefront that reserved its
// slot; a group's entries of one cell are one run of consecutive slots in
// ascending order (one reservation), and groups are disjoint ranges of
// sources.  So an entry's rank in the sorted cell is the length of every run
// of a smaller group plus its index in its own run.  A few runs are walked
// one by one; more (up to 64) are gathered one per lane, sorted by group
// with a bitonic sort across the wavefront, and their lengths scanned in
// that order, which gives each run's first rank.  Writes only when some
// entry moves.
template <int CNT>
__device__ __forceinline__ void fix_runs_regs(uint32_t* perm, int b, int m, const uint32_t (&v)[CNT]) {
  const int lane = threadIdx.x & 63;
  // run heads: the first entry and every entry whose group differs from the
  // previous entry's
  unsigned long long H[CNT];
  int cb[CNT + 1];  // heads in the registers before q
  cb[0] = 0;
#pragma unroll
  for (int q = 0; q < CNT; q++) {
    const int t = lane + 64 * q;
    uint32_t prev = (uint32_t)__shfl_up((int)v[q], 1, 64);
    if (lane == 0) prev = q > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)v[q > 0 ? q - 1 : 0], 63) : 0xFFFFFFFFu;
    H[q] = __builtin_amdgcn_ballot_w64(t < m && (t == 0 || (v[q] >> 6) != (prev >> 6)));
    cb[q + 1] = cb[q] + (int)__popcll(H[q]);
  }
  const int R = cb[CNT];
  // the groups' spread, for the packed sort keys (group offset << 8 | start)
  uint32_t gmin = 0xFFFFFFFFu, gmax = 0;
#pragma unroll
  for (int q = 0; q < CNT; q++) {
    const bool in = lane + 64 * q < m;
    gmin = min(gmin, in ? v[q] >> 6 : 0xFFFFFFFFu);
    gmax = max(gmax, in ? v[q] >> 6 : 0u);
  }
  gmin = (uint32_t)wave_min((int)min(gmin, 0x7FFFFFFFu));
  gmax = (uint32_t)wave_max((int)min(gmax, 0x7FFFFFFFu));
  int rank[CNT];
  if (R > 8 && R <= 64 && gmax - gmin < (1u << 24) && 64 * CNT <= 256) {
    // lane r < R: run r (in slot order) -- its head's register and lane by a
    // binary search for the (r - cb[q])-th set bit of H[q]
    int qr = 0;
#pragma unroll
    for (int q = 1; q < CNT; q++) qr += lane >= cb[q] ? 1 : 0;
    unsigned long long hm = H[0];
    int kb = 0;
#pragma unroll
    for (int q = 1; q < CNT; q++)
      if (qr == q) {
        hm = H[q];
        kb = cb[q];
      }
    const int k = lane - kb;
    int lo = 0;
#pragma unroll
    for (int st = 32; st >= 1; st >>= 1)
      if (popc_below(hm, lo + st) <= k) lo += st;
    // the head's group (read from its lane of every register, the right one kept)
    uint32_t gh = 0;
#pragma unroll
    for (int q = 0; q < CNT; q++) {
      const uint32_t x = (uint32_t)__shfl((int)v[q], lo, 64);
      gh = qr == q ? x : gh;
    }
    const int t_r = 64 * qr + lo;  // run r's first entry (lanes r < R)
    const int t_next = __shfl_down(t_r, 1, 64);
    const int len_r = lane + 1 < R ? t_next - t_r : m - t_r;
    const uint32_t key = lane < R ? (((gh >> 6) - gmin) << 8) | (uint32_t)t_r : 0xFFFFFFFFu;
    const uint32_t sk = wave_sort_u(key);
    // sorted position k: its run's index in slot order (from its start) and length
    const int ts = (int)(sk & 255u);
    int rs = 0;
#pragma unroll
    for (int q = 0; q < CNT; q++) rs += (ts >> 6) > q ? cb[q + 1] - cb[q] : ((ts >> 6) == q ? popc_below(H[q], ts & 63) : 0);
    const int lk = __shfl(len_r, lane < R ? rs : 0, 64);
    // exclusive scan of the lengths in group order
    int inc = lane < R ? lk : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(inc, off, 64);
      inc += lane >= off ? o : 0;
    }
    const int base_k = inc - (lane < R ? lk : 0);
    // each run's first rank back to lane r (a permutation: sorted lanes k < R
    // send to their run's lane, the rest to themselves)
    const int base_r = __builtin_amdgcn_ds_permute((lane < R ? rs : lane) * 4, base_k);
#pragma unroll
    for (int q = 0; q < CNT; q++) {
      // entry t's run: the heads at or before it, minus one
      const int r = cb[q] + popc_below(H[q], lane + 1) - 1;
      const int rr = max(r, 0);
      rank[q] = __shfl(base_r, rr, 64) + (lane + 64 * q) - __shfl(t_r, rr, 64);
    }
  } else {
#pragma unroll
    for (int q = 0; q < CNT; q++) rank[q] = 0;
    auto close_run = [&](int th, uint32_t g, int end) {
#pragma unroll
      for (int q = 0; q < CNT; q++) {
        const int t = lane + 64 * q;
        rank[q] += (v[q] >> 6) > g ? end - th : 0;
        rank[q] += (t >= th && t < end) ? t - th : 0;
      }
    };
    int prev_th = -1;
    uint32_t prev_g = 0;
#pragma unroll
    for (int qh = 0; qh < CNT; qh++) {
      for (unsigned long long h = H[qh]; h; h &= h - 1) {
        const int lh = (int)__builtin_ctzll(h);
        const int th = lh + 64 * qh;
        const uint32_t g = (uint32_t)__builtin_amdgcn_readlane((int)v[qh], lh) >> 6;
        if (prev_th >= 0) close_run(prev_th, prev_g, th);
        prev_th = th;
        prev_g = g;
      }
    }
    if (prev_th >= 0) close_run(prev_th, prev_g, m);
  }
  bool moved = false;
#pragma unroll
  for (int q = 0; q < CNT; q++) moved = moved || (lane + 64 * q < m && rank[q] != lane + 64 * q);
  if (!__builtin_amdgcn_ballot_w64(moved)) return;
#pragma unroll
  for (int q = 0; q < CNT; q++)
    if (lane + 64 * q < m) perm[b + rank[q]] = v[q];
}

template <int CNT>
__device__ __forceinline__ void fix_by_runs(uint32_t* perm, int b, int m) {
  const int lane = threadIdx.x & 63;
  uint32_t v[CNT];
#pragma unroll
  for (int q = 0; q < CNT; q++) {
    const int t = lane + 64 * q;
    v[q] = t < m ? perm[b + t] : 0xFFFFFFFFu;
  }
  fix_runs_regs<CNT>(perm, b, m, v);
}

// FIX_GROUP consecutive cells per wavefront, each read once.  When every cell
// of the group has at most 128 entries (the common case), all the group's
// entries are loaded at once, so the wavefront waits for memory once per
// group rather than once per cell.  Otherwise cell by cell: cells of up to
// 256 entries are ordered by their runs, larger ones checked for a descent
// and sorted by rank counting when they have one.
constexpr int FIX_GROUP = 8;

__device__ __forceinline__ void fix_cells(const int* start, uint32_t* perm, uint32_t* scratch, int c0, int ncell,
                                          uint32_t* buf) {
  const int lane = threadIdx.x & 63;
  const int nc = min(FIX_GROUP, ncell - c0);
  const int sv = lane <= nc ? start[c0 + lane] : 0;
  // the group's largest cell: lanes q < nc hold cell q's size (the shuffle
  // runs on every lane: lane nc - 1 reads lane nc's start)
  const int sv_next = __shfl_down(sv, 1, 64);
  const int msz = lane < nc ? sv_next - sv : 0;
  if (wave_max(msz) <= 128) {
    uint32_t v[FIX_GROUP][2];
#pragma unroll
    for (int q = 0; q < FIX_GROUP; q++) {
      const int b = __builtin_amdgcn_readlane(sv, q);
      const int m = q < nc ? __builtin_amdgcn_readlane(sv, q + 1) - b : 0;
      v[q][0] = lane < m ? perm[b + lane] : 0xFFFFFFFFu;
      v[q][1] = lane + 64 < m ? perm[b + 64 + lane] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int q = 0; q < FIX_GROUP; q++) {
      const int b = __builtin_amdgcn_readlane(sv, q);
      const int m = q < nc ? __builtin_amdgcn_readlane(sv, q + 1) - b : 0;
      if (m > 1) fix_runs_regs<2>(perm, b, m, v[q]);
    }
    return;
  }
  for (int q = 0; q < nc; q++) {
    const int b = __builtin_amdgcn_readlane(sv, q);
    const int m = __builtin_amdgcn_readlane(sv, q + 1) - b;
    if (m <= 1) continue;
    if (m <= 64) {
      fix_by_runs<1>(perm, b, m);
    } else if (m <= 128) {
      fix_by_runs<2>(perm, b, m);
    } else if (m <= 256) {
      fix_by_runs<4>(perm, b, m);
    } else {
      bool desc = false;
      for (int t = b + lane; t + 1 < b + m; t += 64) desc = desc || perm[t] > perm[t + 1];
      if (__builtin_amdgcn_ballot_w64(desc)) fix_unordered(perm, scratch, b, m, buf);
    }
  }
}

// Gather of one chunk into the second buffer (n = live total).  The source
// indices and the destination are streamed with nontemporal loads and
// stores, so that the L2 keeps the source lines, which neighbouring chunks
// read again (a particle that changed cell is read from its old cell's run):
// 5 % faster than plain stores on a 26.2M-particle species.
// Source indices of one chunk.
// Source indices of one chunk (base < n); a slot past the end takes the
// last live slot's source, so that every load of the move is in range and
// none needs a branch (a branch per particle makes each wait for the last).
__device__ __forceinline__ void gather_sources(const uint32_t* perm, int64_t n, int64_t base, uint32_t (&p)[SORT_U]) {
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
    p[u] = __builtin_nontemporal_load(&perm[i < n ? i : n - 1]);
  }
}

// Moves one chunk's particles from their sources p.  When nbase >= 0 the
// next chunk's source indices are loaded into pn after this chunk's reads
// are issued, so that their latency overlaps this chunk's.
template <typename P>
__device__ __forceinline__ void gather_move(const uint32_t (&p)[SORT_U], int64_t n, int64_t base, const P* x,
                                            const P* y, const P* vx, const P* vy, const P* vz, const uint32_t* id,
                                            P* x2, P* y2, P* vx2, P* vy2, P* vz2, uint32_t* id2,
                                            const uint32_t* perm, int64_t nbase, uint32_t (&pn)[SORT_U]) {
  P a[SORT_U], b[SORT_U], c[SORT_U], d[SORT_U], e[SORT_U];
  uint32_t f[SORT_U];
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int64_t q = pslot<P>(p[u]);
    a[u] = x[q]; b[u] = y[q]; c[u] = vx[q]; d[u] = vy[q]; e[u] = vz[p[u]]; f[u] = id[p[u]];
  }
  if (nbase >= 0) gather_sources(perm, n, nbase, pn);
#pragma unroll
  for (int u = 0; u < SORT_U; u++) {
    const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
    if (i < n) {
      const int64_t q = pslot<P>(i);
      __builtin_nontemporal_store(a[u], &x2[q]);
      __builtin_nontemporal_store(b[u], &y2[q]);
      __builtin_nontemporal_store(c[u], &vx2[q]);
      __builtin_nontemporal_store(d[u], &vy2[q]);
      __builtin_nontemporal_store(e[u], &vz2[i]);
      __builtin_nontemporal_store(f[u], &id2[i]);
    }
  }
}

template <typename P>
__device__ __forceinline__ void gather_chunk(const uint32_t* perm, int64_t n, int64_t base, const P* x, const P* y,
                                             const P* vx, const P* vy, const P* vz, const uint32_t* id, P* x2,
                                             P* y2, P* vx2, P* vy2, P* vz2, uint32_t* id2) {
  uint32_t p[SORT_U];
  gather_sources(perm, n, base, p);
  gather_move(p, n, base, x, y, vx, vy, vz, id, x2, y2, vx2, vy2, vz2, id2, perm, -1, p);
}

}  // namespace epic
