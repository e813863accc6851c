// Device building blocks of the particle kernels, shared by the
// single-simulation kernels (particles.hip) and the batched ensemble
// (batch.hip).
#pragma once

#include "device_common.hpp"
#include "epic_internal.hpp"
#include <type_traits>

namespace epic {

typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));

template <typename P>
struct alignas(2 * sizeof(P)) Vec2 {
  P a, b;
};

// Particle k of a lane within its wavefront chunk of 64*PPT particles:
// V = 16/sizeof(P) consecutive particles per 16-byte vector, group g = k/V
// at offset g*64*V.  Load g of a wavefront is then one contiguous 1 KiB
// block (fully coalesced), and each lane owns runs of V consecutive
// particles (what the gather cell reuse and the run-length deposit rely on).
// This is the slot number; its address in an FP32 array is pslot<float>.
// Chunks start at multiples of 64*PPT slots, so for FP32 (PPT a multiple of
// 4) every group of 64*V = 256 slots is one tile's run of one coordinate.
template <typename P, int PPT>
__device__ __forceinline__ int64_t part_index(int64_t wbase, int lane, int k) {
  static_assert(sizeof(P) == 8 || PPT % 4 == 0, "FP32 chunks are whole tiles");
  constexpr int V = 16 / (int)sizeof(P);
  return wbase + (int64_t)(k / V) * (64 * V) + (int64_t)lane * V + (k % V);
}

template <typename P, int PPT>
__device__ __forceinline__ void ld_run(const P* __restrict__ a, int64_t wbase, int lane, int64_t n, P (&o)[PPT]) {
  constexpr int V = 16 / (int)sizeof(P);
  constexpr int NG = PPT / V;
  if (wbase + 64 * PPT <= n) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
      const v4u32* p =
          reinterpret_cast<const v4u32*>(a + pslot<P>(wbase + (int64_t)g * 64 * V) + (int64_t)lane * V);
      const v4u32 t = __builtin_nontemporal_load(p);
      __builtin_memcpy(&o[g * V], &t, 16);
    }
  } else {
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int64_t i = part_index<P, PPT>(wbase, lane, k);
      o[k] = i < n ? a[pslot<P>(i)] : P(0);
    }
  }
}

template <typename P, int PPT>
__device__ __forceinline__ void st_run(P* __restrict__ a, int64_t wbase, int lane, int64_t n, const P (&o)[PPT]) {
  constexpr int V = 16 / (int)sizeof(P);
  constexpr int NG = PPT / V;
  if (wbase + 64 * PPT <= n) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
      v4u32 t;
      __builtin_memcpy(&t, &o[g * V], 16);
      __builtin_nontemporal_store(
          t, reinterpret_cast<v4u32*>(a + pslot<P>(wbase + (int64_t)g * 64 * V) + (int64_t)lane * V));
    }
  } else {
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int64_t i = part_index<P, PPT>(wbase, lane, k);
      if (i < n) a[pslot<P>(i)] = o[k];
    }
  }
}

// Node-window accumulator flush: the 2x2 node block whose lower corner is
// (bi, bj); inside the LDS window -> LDS atomics, else global atomics.
struct Window {
  int i0, j0, wi, wj;
};

// Atomic adds with the address space spelled out.  Through a generic
// pointer the compiler may merge an LDS and a global atomic into one FLAT
// atomic, which counts against both the vector-memory and the LDS counters
// and makes every later wait drain all outstanding loads.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
typedef __attribute__((address_space(1))) unsigned long long global_u64;

__device__ __forceinline__ void lds_add(unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_add((lds_u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void global_add(unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_add((global_u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void flush_block(int bi, int bj, int ny, const unsigned long long (&acc)[4],
                                            int ncorner, const Window& w, unsigned long long* lds,
                                            unsigned long long* Q) {
  const int li = bi - w.i0, lj = bj - w.j0;
  const bool in = li >= 0 && lj >= 0 && li + (ncorner > 1 ? 1 : 0) < w.wi && lj + (ncorner > 1 ? 1 : 0) < w.wj;
  if (in) {
    unsigned long long* b = lds + li * w.wj + lj;
    if (acc[0]) lds_add(b, acc[0]);
    if (ncorner > 1) {
      if (acc[1]) lds_add(b + w.wj, acc[1]);
      if (acc[2]) lds_add(b + 1, acc[2]);
      if (acc[3]) lds_add(b + w.wj + 1, acc[3]);
    }
  } else {
    unsigned long long* b = Q + (size_t)bi * ny + bj;
    if (acc[0]) global_add(b, acc[0]);
    if (ncorner > 1) {
      if (acc[1]) global_add(b + ny, acc[1]);
      if (acc[2]) global_add(b + 1, acc[2]);
      if (acc[3]) global_add(b + ny + 1, acc[3]);
    }
  }
}

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int WAVE_WIN = 512;


// ---------------------------------------------------------------------------
// FP32 hot path: fused gather + leapfrog push + boundary + deposit of the
// new position, for the storage precision of the large workloads.  Same
// structure as particle_kernel, with the arithmetic cut to what an FP32
// particle needs:
//   logical coordinates by multiplication with 1/dx (no division),
//   v += (c1 c2) E and x += (c3 c2) v as fused multiply-adds,
//   bilinear gather from per-cell coefficients (cell_coeffs): one 32-byte
//   fetch when the cell changes, three FMAs per component,
//   integer CIC weights from 14-bit offsets dI = rint(di 2^14):
//     w00 = (2^14-dI)(2^14-dJ), w10 = dI(2^14-dJ), w01 = (2^14-dI)dJ, w11 = dI dJ
//   which sum to exactly 2^28 per particle (FIX_BITS_FP32).
// The deposit stays an integer sum, so it is deterministic; the 2^-14 cell
// resolution matches the FP32 position resolution at 512 cells.
// x, xo: a species' tiled particle storage (x, y, vx, vy of each 256-slot
// tile at +0, +256, +512, +768) read and written; xo is the other buffer of
// the pair (the push streams from one into the other, which the memory
// system serves faster than an in-place update) or x itself.
struct PushF32 {
  const float* x;
  float* xo;
  const int64_t* d_n;
  const float4* Ec;  // per cell c = i*(ny-1) + j: bilinear coefficients of ex, then of ey
  float x0, y0, idx, idy, xhi, yhi, Lx, Ly;
  float kv, kx;  // c1*c2, c3*c2
  int nx, ny, bc;
  unsigned long long* Q;
  unsigned long long* errc;
};

// Bilinear interpolation of node values e00, e10 (i+1), e01 (j+1), e11 over
// a cell as a0 + a1 di + dj (a2 + a3 di).
__device__ __forceinline__ float4 bilinear_coeffs(float e00, float e10, float e01, float e11) {
  const float d10 = e10 - e00, d01 = e01 - e00;
  return make_float4(e00, d10, d01, (e11 - e10) - d01);
}

// Cell table of the FP32 gather from the interleaved node field E2.
__device__ __forceinline__ void cell_coeffs(const float2* E2, int i, int j, int ny, float4* out2) {
  const float2 e00 = E2[(size_t)i * ny + j], e10 = E2[(size_t)(i + 1) * ny + j];
  const float2 e01 = E2[(size_t)i * ny + j + 1], e11 = E2[(size_t)(i + 1) * ny + j + 1];
  out2[0] = bilinear_coeffs(e00.x, e10.x, e01.x, e11.x);
  out2[1] = bilinear_coeffs(e00.y, e10.y, e01.y, e11.y);
}

constexpr unsigned NO_DEPOSIT = 0xffffffffu;

// The particle registers of one wavefront chunk (64*PPT particles).
template <int PPT>
struct ChunkRegs {
  float x[PPT], y[PPT], vx[PPT], vy[PPT];
};

// Deposit information of a pushed chunk: packed cells (i << 16 | j) and
// 14-bit offsets (dI << 16 | dJ) of the new positions (grids up to 65535
// nodes per side), and the lane's bounding box of cells.
template <int PPT>
struct ChunkDep {
  unsigned cp[PPT], wp[PPT];
  int imin, imax, jmin, jmax;
};

// Slots past the end of the array (a partial last chunk) load x = NaN, the
// mark of an absent particle, so the push needs no separate range test.
template <int PPT>
__device__ __forceinline__ void load_chunk(const PushF32& a, int64_t wb, int64_t n, int lane, ChunkRegs<PPT>& r) {
  ld_run<float, PPT>(a.x, wb, lane, n, r.x);
  ld_run<float, PPT>(a.x + PTILE, wb, lane, n, r.y);
  ld_run<float, PPT>(a.x + 2 * PTILE, wb, lane, n, r.vx);
  ld_run<float, PPT>(a.x + 3 * PTILE, wb, lane, n, r.vy);
  if (wb + 64 * PPT > n) {
#pragma unroll
    for (int k = 0; k < PPT; k++)
      if (part_index<float, PPT>(wb, lane, k) >= n) r.x[k] = __builtin_nanf("");
  }
}

// A chunk known to be full (wb + 64*PPT <= n): the vector loads only.
template <int PPT>
__device__ __forceinline__ void load_chunk_full(const PushF32& a, int64_t wb, int lane, ChunkRegs<PPT>& r) {
  constexpr int V = 4;
  float* const dst[4] = {r.x, r.y, r.vx, r.vy};
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int g = 0; g < PPT / V; g++) {
      const v4u32 t = __builtin_nontemporal_load(
          reinterpret_cast<const v4u32*>(a.x + q * PTILE + ptile(wb + (int64_t)g * 64 * V) + (int64_t)lane * V));
      __builtin_memcpy(&dst[q][g * V], &t, 16);
    }
}

// Lanes of the wavefront where c holds (a ballot and a scalar popcount).
__device__ __forceinline__ unsigned wave_count(bool c) {
  return (unsigned)__builtin_popcountll(__builtin_amdgcn_ballot_w64(c));
}

// Gather, push, walls and store of a loaded chunk, for wall kind BC; fills
// the deposit information of the new positions (DEP).  The per-particle
// code is straight-line: the wall kind is resolved at compile time, the rare
// cases (absent or absorbed particles, positions outside the grid) are
// selects, and a particle's index values (cells, offsets) are used only
// where its range test holds.  COND: the walls are followed by the conductor
// test against the first ncond shapes of cond, in the cells that cmask flags
// (a separate instantiation, so the kernels without conductors keep their
// code; the table is an argument of its own, not part of PushF32, whose size
// the batched kernel's descriptor loads depend on).
template <int BC, bool DEP, int PPT, bool COND>
__device__ __forceinline__ void push_chunk_bc(const PushF32& a, int64_t wb, int64_t n, int lane,
                                              ChunkRegs<PPT>& r, ChunkDep<PPT>& dp, const CondDev* cond,
                                              int ncond, const uint32_t* cmask) {
  constexpr int V = 4;
  constexpr int G = PPT / V;
  const int nyc = a.ny - 1;
  const float fmax_i = (float)(a.nx - 1), fmax_j = (float)(a.ny - 1);
  const int imx = a.nx - 2, jmx = a.ny - 2;
  // reflecting walls: the highest FP32 position below each upper wall
  const float xtop = BC == EPIC_BC_REFLECTIVE ? nextafterf(a.xhi, a.x0) : 0.f;
  const float ytop = BC == EPIC_BC_REFLECTIVE ? nextafterf(a.yhi, a.y0) : 0.f;
  const float qnan = __builtin_nanf("");
  int imin = 1 << 30, imax = -1, jmin = 1 << 30, jmax = -1;
  // error counts of the chunk, wavefront-uniform (ballots in scalar
  // registers: a per-lane count would hold a vector register all along)
  unsigned n_oob = 0, n_nan = 0;
  // the cell of each run's first particle: its coefficients are fetched for
  // all runs at once, so the fetches overlap
  int gkey[G];
  float4 gx[G], gy[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    gkey[g] = -1;
    gx[g] = gy[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    const float lx = (r.x[g * V] - a.x0) * a.idx;
    const float ly = (r.y[g * V] - a.y0) * a.idy;
    if (lx >= 0.f && ly >= 0.f && lx <= fmax_i && ly <= fmax_j) {
      gkey[g] = min((int)lx, imx) * nyc + min((int)ly, jmx);
      gx[g] = a.Ec[2 * (size_t)gkey[g]];
      gy[g] = a.Ec[2 * (size_t)gkey[g] + 1];
    }
  }
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    const float x = r.x[k], y = r.y[k], vx = r.vx[k], vy = r.vy[k];
    // NaN x: an absorbed particle or a slot past the end.  It stays NaN (its
    // other coordinates are not used again: every consumer skips NaN x)
    const bool live = !isnan(x);
    const float lx = (x - a.x0) * a.idx;
    const float ly = (y - a.y0) * a.idy;
    const bool inr = lx >= 0.f && ly >= 0.f && lx <= fmax_i && ly <= fmax_j;
    n_oob += wave_count(live && !inr);
    const int i = min((int)lx, imx), j = min((int)ly, jmx);
    const int key = i * nyc + j;
    float4& cx = gx[k / V];
    float4& cy = gy[k / V];
    if (inr && key != gkey[k / V]) {
      cx = a.Ec[2 * (size_t)key];
      cy = a.Ec[2 * (size_t)key + 1];
      gkey[k / V] = key;
    }
    const float di = lx - (float)i, dj = ly - (float)j;
    float ex = fmaf(dj, fmaf(di, cx.w, cx.z), fmaf(di, cx.y, cx.x));
    float ey = fmaf(dj, fmaf(di, cy.w, cy.z), fmaf(di, cy.y, cy.x));
    ex = inr ? ex : 0.f;
    ey = inr ? ey : 0.f;
    float nvx = fmaf(a.kv, ex, vx);
    float nvy = fmaf(a.kv, ey, vy);
    const float px = fmaf(a.kx, nvx, x);
    const float py = fmaf(a.kx, nvy, y);
    // a non-finite new position: counted, the particle is dropped
    const bool nanp = live & (isnan(px) | isnan(py));
    n_nan += wave_count(nanp);
    float xo, yo;
    bool alive;
    if (BC == EPIC_BC_PBC) {
      float wx = px < a.x0 ? px + a.Lx : (px >= a.xhi ? px - a.Lx : px);
      float wy = py < a.y0 ? py + a.Ly : (py >= a.yhi ? py - a.Ly : py);
      // x0 - tiny + Lx can round up to xhi itself: the periodic image of x0
      wx = wx >= a.xhi ? wx - a.Lx : wx;
      wy = wy >= a.yhi ? wy - a.Ly : wy;
      alive = live && !nanp;
      xo = nanp ? qnan : wx;
      yo = nanp ? py : wy;
    } else if (BC == EPIC_BC_OPEN) {
      // false for a NaN coordinate
      alive = live && px >= a.x0 && px < a.xhi && py >= a.y0 && py < a.yhi;
      xo = alive ? px : qnan;
      yo = py;
    } else {
      const bool lox = px < a.x0, hix = px >= a.xhi;
      const bool loy = py < a.y0, hiy = py >= a.yhi;
      float rx = lox ? a.x0 + (a.x0 - px) : (hix ? a.xhi - (px - a.xhi) : px);
      float ry = loy ? a.y0 + (a.y0 - py) : (hiy ? a.yhi - (py - a.yhi) : py);
      // a position exactly on the upper wall mirrors onto itself: keep live
      // positions below it
      rx = rx >= a.xhi ? xtop : rx;
      ry = ry >= a.yhi ? ytop : ry;
      const bool flipx = !nanp && (lox || hix), flipy = !nanp && (loy || hiy);
      nvx = flipx ? -nvx : nvx;
      nvy = flipy ? -nvy : nvy;
      alive = live && !nanp;
      xo = nanp ? qnan : rx;
      yo = nanp ? py : ry;
    }
    if (COND && alive) {
      // the cell's flag first (cmask: cells a shape's bounding box reaches,
      // widened by one), the exact FP64 test only in flagged cells
      const int ci = min(max((int)((xo - a.x0) * a.idx), 0), a.nx - 2);
      const int cj = min(max((int)((yo - a.y0) * a.idy), 0), a.ny - 2);
      const int cell = ci * nyc + cj;
      if (((cmask[cell >> 5] >> (cell & 31)) & 1u) && in_conductor(cond, ncond, (double)xo, (double)yo)) {
        alive = false;
        xo = qnan;
      }
    }
    r.x[k] = xo;
    r.y[k] = yo;
    r.vx[k] = nvx;
    r.vy[k] = nvy;
    if (DEP) {
      const float nlx = (xo - a.x0) * a.idx, nly = (yo - a.y0) * a.idy;
      // a wrapped position can round onto the upper boundary node: it
      // deposits into the last cell
      const bool inr2 = nlx >= 0.f && nly >= 0.f && nlx <= fmax_i && nly <= fmax_j;
      n_oob += wave_count(alive && !inr2);
      const bool dep = alive && inr2;
      const int ni = min((int)nlx, imx), nj = min((int)nly, jmx);
      const int dI = min((int)rintf((nlx - (float)ni) * 16384.0f), 16384);
      const int dJ = min((int)rintf((nly - (float)nj) * 16384.0f), 16384);
      dp.cp[k] = dep ? (((unsigned)ni << 16) | (unsigned)nj) : NO_DEPOSIT;
      dp.wp[k] = ((unsigned)dI << 16) | (unsigned)dJ;
      imin = min(imin, dep ? ni : 1 << 30);
      imax = max(imax, dep ? ni : -1);
      jmin = min(jmin, dep ? nj : 1 << 30);
      jmax = max(jmax, dep ? nj : -1);
    }
    // one particle at a time: interleaving the particles' straight-line code
    // (eight waves per SIMD hide the latency anyway) needs more registers
    // than the 64 of that occupancy
    __builtin_amdgcn_sched_barrier(0);
  }
  dp.imin = imin;
  dp.imax = imax;
  dp.jmin = jmin;
  dp.jmax = jmax;
  // Every gather fetch has been consumed here, so this wait costs nothing;
  // it tells the compiler that no fetch is outstanding, which otherwise
  // (its analysis is path-insensitive) drains all vector memory -- the next
  // chunk's loads included -- before registers are reused in the deposit.
  // s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15)
  __builtin_amdgcn_s_waitcnt(0x0F70);
  st_run<float, PPT>(a.xo + 2 * PTILE, wb, lane, n, r.vx);
  st_run<float, PPT>(a.xo + 3 * PTILE, wb, lane, n, r.vy);
  st_run<float, PPT>(a.xo, wb, lane, n, r.x);
  st_run<float, PPT>(a.xo + PTILE, wb, lane, n, r.y);
  // one lane adds the wavefront's counts
  if ((n_oob | n_nan) && lane == __builtin_amdgcn_readfirstlane(lane)) {
    if (n_oob) atomicAdd(&a.errc[EPIC_ERRC_GATHER_OOB], (unsigned long long)n_oob);
    if (n_nan) atomicAdd(&a.errc[EPIC_ERRC_NAN], (unsigned long long)n_nan);
  }
}

// Gather, push, walls and store of a loaded chunk.  BC >= 0: the wall kind
// of every chunk the kernel sees (a.bc == BC), fixed at compile time; BC < 0:
// read from a.bc, one branch per chunk (three copies of the particle code,
// which need more registers: for the measurement shapes and batches whose
// members mix wall kinds).
template <bool DEP, int PPT, bool COND, int BC>
__device__ __forceinline__ void push_chunk_regs(const PushF32& a, int64_t wb, int64_t n, int lane,
                                                ChunkRegs<PPT>& r, ChunkDep<PPT>& dp,
                                                const CondDev* cond = nullptr, int ncond = 0,
                                                const uint32_t* cmask = nullptr) {
  if constexpr (BC >= 0) {
    push_chunk_bc<BC, DEP, PPT, COND>(a, wb, n, lane, r, dp, cond, ncond, cmask);
  } else {
    if (a.bc == EPIC_BC_OPEN)
      push_chunk_bc<EPIC_BC_OPEN, DEP, PPT, COND>(a, wb, n, lane, r, dp, cond, ncond, cmask);
    else if (a.bc == EPIC_BC_PBC)
      push_chunk_bc<EPIC_BC_PBC, DEP, PPT, COND>(a, wb, n, lane, r, dp, cond, ncond, cmask);
    else
      push_chunk_bc<EPIC_BC_REFLECTIVE, DEP, PPT, COND>(a, wb, n, lane, r, dp, cond, ncond, cmask);
  }
}

// Wall kind of a kernel instantiation: -1 reads it per chunk.
constexpr int BC_ANY = -1;

// Periodic walls: offset d of a cell index from the anchor's, taken modulo
// the period C (cells) into [-C/2, C/2], and a node index brought back into
// [0, C) (node C is node 0's periodic image; the finalisation folds the two).
__device__ __forceinline__ int per_off(int d, int C) { return d > C / 2 ? d - C : (d < -(C / 2) ? d + C : d); }
__device__ __forceinline__ int per_wrap(int g, int C) { return g < 0 ? g + C : (g >= C ? g - C : g); }

// Node-window accumulator flush of a run under periodic walls: (ri, rj) is
// the run's cell as offsets from the anchor (ai, aj); the global fallback
// adds to the wrapped nodes.
__device__ __forceinline__ void flush_block_per(int ri, int rj, int ny, const unsigned long long (&acc)[4],
                                                const Window& w, unsigned long long* lds, unsigned long long* Q,
                                                int ai, int aj, int Ci, int Cj) {
  const int li = ri - w.i0, lj = rj - w.j0;
  if (li >= 0 && lj >= 0 && li + 1 < w.wi && lj + 1 < w.wj) {
    unsigned long long* b = lds + li * w.wj + lj;
    if (acc[0]) lds_add(b, acc[0]);
    if (acc[1]) lds_add(b + w.wj, acc[1]);
    if (acc[2]) lds_add(b + 1, acc[2]);
    if (acc[3]) lds_add(b + w.wj + 1, acc[3]);
  } else {
    const int gi = per_wrap(ai + ri, Ci), gj = per_wrap(aj + rj, Cj);
    const int gi1 = per_wrap(gi + 1, Ci), gj1 = per_wrap(gj + 1, Cj);
    if (acc[0]) global_add(Q + (size_t)gi * ny + gj, acc[0]);
    if (acc[1]) global_add(Q + (size_t)gi1 * ny + gj, acc[1]);
    if (acc[2]) global_add(Q + (size_t)gi * ny + gj1, acc[2]);
    if (acc[3]) global_add(Q + (size_t)gi1 * ny + gj1, acc[3]);
  }
}

// Deposit of a pushed chunk through the wavefront's LDS window wl: each
// lane sums runs of equal cells in registers, runs go to the window (or
// straight to Q outside it), the window goes to Q.  PER (periodic walls):
// the window is placed in cell offsets from an anchor cell (the first
// deposit of the first lane that has one), each offset taken modulo the
// period, so that a particle that has just wrapped across the boundary stays
// next to its chunk's other particles instead of stretching the window
// across the grid; nodes are mapped back modulo the period when they reach Q
// (node C then lands on its image, node 0, which changes no folded density).
template <int PPT, bool PER = false>
__device__ __forceinline__ void deposit_chunk(const PushF32& a, int lane, unsigned long long* wl,
                                              const ChunkDep<PPT>& dp) {
  const int ny = a.ny;
  const int Ci = a.nx - 1, Cj = a.ny - 1;
  int ai = 0, aj = 0;
  int gi0, gi1, gj0, gj1;
  if (PER) {
    unsigned first = NO_DEPOSIT;
#pragma unroll
    for (int k = 0; k < PPT; k++) first = first == NO_DEPOSIT ? dp.cp[k] : first;
    const unsigned long long m = __builtin_amdgcn_ballot_w64(first != NO_DEPOSIT);
    if (!m) return;
    const unsigned anc = (unsigned)__builtin_amdgcn_readlane((int)first, (int)__builtin_ctzll(m));
    ai = (int)(anc >> 16);
    aj = (int)(anc & 0xffffu);
    int imin = 1 << 30, imax = -(1 << 30), jmin = 1 << 30, jmax = -(1 << 30);
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      if (dp.cp[k] == NO_DEPOSIT) continue;
      const int ri = per_off((int)(dp.cp[k] >> 16) - ai, Ci), rj = per_off((int)(dp.cp[k] & 0xffffu) - aj, Cj);
      imin = min(imin, ri); imax = max(imax, ri);
      jmin = min(jmin, rj); jmax = max(jmax, rj);
    }
    gi0 = wave_min_dpp(imin); gi1 = wave_max_dpp(imax);
    gj0 = wave_min_dpp(jmin); gj1 = wave_max_dpp(jmax);
  } else {
    gi0 = wave_min_dpp(dp.imin); gi1 = wave_max_dpp(dp.imax);
    if (gi1 < 0) return;
    gj0 = wave_min_dpp(dp.jmin); gj1 = wave_max_dpp(dp.jmax);
  }
  const int wi = gi1 - gi0 + 2, wj = gj1 - gj0 + 2;
  Window win;
  if (wi * wj <= WAVE_WIN) {
    win = Window{gi0, gj0, wi, wj};
  } else {
    const int wjc = min(wj, 32);
    win = Window{gi0, gj0, WAVE_WIN / wjc, wjc};
  }
  for (int k = lane; k < win.wi * win.wj; k += 64) wl[k] = 0ull;
  wave_fence();

  unsigned rk = NO_DEPOSIT;
  // A particle adds at most 16384^2 = 2^28 units to a corner, so a run of
  // PPT particles in one cell sums to at most PPT * 2^28: 32-bit sums hold
  // up to 15 particles, longer chunks sum in 64 bits.
  using acc_t = typename std::conditional<((unsigned long long)PPT << 28) < (1ull << 32), unsigned,
                                          unsigned long long>::type;
  static_assert((unsigned long long)PPT * (1ull << 28) <= (unsigned long long)(acc_t)~(acc_t)0,
                "run accumulator too narrow for PPT particles");
  acc_t acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
  auto flush = [&]() {
    const unsigned long long acc[4] = {acc0, acc1, acc2, acc3};
    if (PER)
      flush_block_per(per_off((int)(rk >> 16) - ai, Ci), per_off((int)(rk & 0xffffu) - aj, Cj), ny, acc, win, wl,
                      a.Q, ai, aj, Ci, Cj);
    else
      flush_block((int)(rk >> 16), (int)(rk & 0xffffu), ny, acc, 4, win, wl, a.Q);
  };
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    if (dp.cp[k] == NO_DEPOSIT) continue;
    if (dp.cp[k] != rk) {
      if (rk != NO_DEPOSIT) flush();
      rk = dp.cp[k];
      acc0 = acc1 = acc2 = acc3 = 0;
    }
    const unsigned dI = dp.wp[k] >> 16, dJ = dp.wp[k] & 0xffffu;
    const unsigned oI = 16384u - dI, oJ = 16384u - dJ;
    acc0 += (acc_t)(oI * oJ);
    acc1 += (acc_t)(dI * oJ);
    acc2 += (acc_t)(oI * dJ);
    acc3 += (acc_t)(dI * dJ);
  }
  if (rk != NO_DEPOSIT) flush();
  wave_fence();
  const int W = win.wi * win.wj;
  // row of window slot k as floor((k + 1/2) / wj) in FP32: exact, since
  // k < 512 and wj <= 512 keep (k + 1/2) / wj at least 1/1024 from an integer
  const float rwj = 1.0f / (float)win.wj;
  for (int k = lane; k < W; k += 64) {
    const unsigned long long v = wl[k];
    if (v) {
      const int ri = (int)(((float)k + 0.5f) * rwj);
      int gi = win.i0 + ri, gj = win.j0 + (k - ri * win.wj);
      if (PER) {
        atomicAdd(&a.Q[(size_t)per_wrap(ai + gi, Ci) * ny + per_wrap(aj + gj, Cj)], v);
      } else if (gi >= 0 && gj >= 0 && gi < a.nx && gj < a.ny) {
        atomicAdd(&a.Q[(size_t)gi * ny + gj], v);
      }
    }
  }
  wave_fence();
}

}  // namespace epic
