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
template <typename P, int PPT>
__device__ __forceinline__ int64_t part_index(int64_t wbase, int lane, int k) {
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
      const v4u32* p = reinterpret_cast<const v4u32*>(a + wbase + (int64_t)g * 64 * V + (int64_t)lane * V);
      const v4u32 t = __builtin_nontemporal_load(p);
      __builtin_memcpy(&o[g * V], &t, 16);
    }
  } else {
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int64_t i = part_index<P, PPT>(wbase, lane, k);
      o[k] = i < n ? a[i] : P(0);
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
      __builtin_nontemporal_store(t, reinterpret_cast<v4u32*>(a + wbase + (int64_t)g * 64 * V + (int64_t)lane * V));
    }
  } else {
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int64_t i = part_index<P, PPT>(wbase, lane, k);
      if (i < n) a[i] = o[k];
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
struct PushF32 {
  float* x;
  float* y;
  float* vx;
  float* vy;
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

template <int PPT>
__device__ __forceinline__ void load_chunk(const PushF32& a, int64_t wb, int64_t n, int lane, ChunkRegs<PPT>& r) {
  ld_run<float, PPT>(a.x, wb, lane, n, r.x);
  ld_run<float, PPT>(a.y, wb, lane, n, r.y);
  ld_run<float, PPT>(a.vx, wb, lane, n, r.vx);
  ld_run<float, PPT>(a.vy, wb, lane, n, r.vy);
}

// A chunk known to be full (wb + 64*PPT <= n): the vector loads only.
template <int PPT>
__device__ __forceinline__ void load_chunk_full(const PushF32& a, int64_t wb, int lane, ChunkRegs<PPT>& r) {
  constexpr int V = 4;
  float* const arr[4] = {a.x, a.y, a.vx, a.vy};
  float* const dst[4] = {r.x, r.y, r.vx, r.vy};
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int g = 0; g < PPT / V; g++) {
      const v4u32 t = __builtin_nontemporal_load(
          reinterpret_cast<const v4u32*>(arr[q] + wb + (int64_t)g * 64 * V + (int64_t)lane * V));
      __builtin_memcpy(&dst[q][g * V], &t, 16);
    }
}

// Gather, push, walls and store of a loaded chunk; fills the deposit
// information of the new positions (DEP).  COND: the walls are followed by
// the conductor test against the first ncond shapes of cond, in the cells
// that cmask flags (a separate
// instantiation, so the kernels without conductors keep their code; the
// table is an argument of its own, not part of PushF32, whose size the
// batched kernel's descriptor loads depend on).
template <bool DEP, int PPT, bool COND = false>
__device__ __forceinline__ void push_chunk_regs(const PushF32& a, int64_t wb, int64_t n, int lane,
                                                ChunkRegs<PPT>& r, ChunkDep<PPT>& dp,
                                                const CondDev* cond = nullptr, int ncond = 0,
                                                const uint32_t* cmask = nullptr) {
  constexpr int V = 4;
  const int nyc = a.ny - 1;
  const float fmax_i = (float)(a.nx - 1), fmax_j = (float)(a.ny - 1);
  const int rem = (int)(n - wb < (int64_t)(64 * PPT) ? n - wb : (int64_t)(64 * PPT));
  dp.imin = 1 << 30;
  dp.imax = -1;
  dp.jmin = 1 << 30;
  dp.jmax = -1;
  // the cell of each run's first particle: its coefficients are fetched for
  // all runs at once, so the fetches overlap
  constexpr int G = PPT / V;
  int gkey[G];
  float4 gx[G], gy[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    gkey[g] = -1;
    gx[g] = gy[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int off = g * (64 * V) + lane * V;
    const float lx = (r.x[g * V] - a.x0) * a.idx;
    const float ly = (r.y[g * V] - a.y0) * a.idy;
    if (off < rem && lx >= 0.f && ly >= 0.f && lx <= fmax_i && ly <= fmax_j) {
      gkey[g] = min((int)lx, a.nx - 2) * nyc + min((int)ly, a.ny - 2);
      gx[g] = a.Ec[2 * (size_t)gkey[g]];
      gy[g] = a.Ec[2 * (size_t)gkey[g] + 1];
    }
  }
  int bad = 0;
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    dp.cp[k] = NO_DEPOSIT;
    dp.wp[k] = 0;
    const int off = (k / V) * (64 * V) + lane * V + (k % V);
    if (off >= rem || isnan(r.x[k])) continue;
    const float lx = (r.x[k] - a.x0) * a.idx;
    const float ly = (r.y[k] - a.y0) * a.idy;
    float ex = 0.f, ey = 0.f;
    if (lx >= 0.f && ly >= 0.f && lx <= fmax_i && ly <= fmax_j) {
      const int i = min((int)lx, a.nx - 2), j = min((int)ly, a.ny - 2);
      const int key = i * nyc + j;
      float4& cx = gx[k / V];
      float4& cy = gy[k / V];
      if (key != gkey[k / V]) {
        cx = a.Ec[2 * (size_t)key];
        cy = a.Ec[2 * (size_t)key + 1];
        gkey[k / V] = key;
      }
      const float di = lx - (float)i, dj = ly - (float)j;
      ex = fmaf(dj, fmaf(di, cx.w, cx.z), fmaf(di, cx.y, cx.x));
      ey = fmaf(dj, fmaf(di, cy.w, cy.z), fmaf(di, cy.y, cy.x));
    } else {
      bad++;
    }
    r.vx[k] = fmaf(a.kv, ex, r.vx[k]);
    r.vy[k] = fmaf(a.kv, ey, r.vy[k]);
    float px = fmaf(a.kx, r.vx[k], r.x[k]);
    float py = fmaf(a.kx, r.vy[k], r.y[k]);
    bool alive = true;
    if (isnan(px) || isnan(py)) {
      bad += NAN_UNIT;
      alive = false;
      px = __builtin_nanf("");
    } else if (a.bc == EPIC_BC_PBC) {
      px = px < a.x0 ? px + a.Lx : (px >= a.xhi ? px - a.Lx : px);
      py = py < a.y0 ? py + a.Ly : (py >= a.yhi ? py - a.Ly : py);
      // x0 - tiny + Lx can round up to xhi itself: the periodic image of x0
      px = px >= a.xhi ? px - a.Lx : px;
      py = py >= a.yhi ? py - a.Ly : py;
    } else if (a.bc == EPIC_BC_OPEN) {
      alive = !(px < a.x0 || px >= a.xhi || py < a.y0 || py >= a.yhi);
      if (!alive) px = __builtin_nanf("");
    } else {
      if (px < a.x0) { px = a.x0 + (a.x0 - px); r.vx[k] = -r.vx[k]; }
      else if (px >= a.xhi) { px = a.xhi - (px - a.xhi); r.vx[k] = -r.vx[k]; }
      if (py < a.y0) { py = a.y0 + (a.y0 - py); r.vy[k] = -r.vy[k]; }
      else if (py >= a.yhi) { py = a.yhi - (py - a.yhi); r.vy[k] = -r.vy[k]; }
      // a position exactly on the upper wall mirrors onto itself: keep live
      // positions below it
      if (px >= a.xhi) px = nextafterf(a.xhi, a.x0);
      if (py >= a.yhi) py = nextafterf(a.yhi, a.y0);
    }
    if (COND && alive) {
      // the cell's flag first (cmask: cells a shape's bounding box reaches,
      // widened by one), the exact FP64 test only in flagged cells
      const int ci = min(max((int)((px - a.x0) * a.idx), 0), a.nx - 2);
      const int cj = min(max((int)((py - a.y0) * a.idy), 0), a.ny - 2);
      const int cell = ci * nyc + cj;
      if (((cmask[cell >> 5] >> (cell & 31)) & 1u) && in_conductor(cond, ncond, (double)px, (double)py)) {
        alive = false;
        px = __builtin_nanf("");
      }
    }
    r.x[k] = px;
    r.y[k] = py;
    if (DEP && alive) {
      const float nlx = (px - a.x0) * a.idx, nly = (py - a.y0) * a.idy;
      // a wrapped position can round onto the upper boundary node: it
      // deposits into the last cell
      if (nlx >= 0.f && nly >= 0.f && nlx <= fmax_i && nly <= fmax_j) {
        const int ni = min((int)nlx, a.nx - 2), nj = min((int)nly, a.ny - 2);
        const int dI = min((int)rintf((nlx - (float)ni) * 16384.0f), 16384);
        const int dJ = min((int)rintf((nly - (float)nj) * 16384.0f), 16384);
        dp.cp[k] = ((unsigned)ni << 16) | (unsigned)nj;
        dp.wp[k] = ((unsigned)dI << 16) | (unsigned)dJ;
        dp.imin = min(dp.imin, ni); dp.imax = max(dp.imax, ni);
        dp.jmin = min(dp.jmin, nj); dp.jmax = max(dp.jmax, nj);
      } else {
        bad++;
      }
    }
  }
  // Every gather fetch has been consumed here, so this wait costs nothing;
  // it tells the compiler that no fetch is outstanding, which otherwise
  // (its analysis is path-insensitive) drains all vector memory -- the next
  // chunk's loads included -- before registers are reused in the deposit.
  // s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15)
  __builtin_amdgcn_s_waitcnt(0x0F70);
  st_run<float, PPT>(a.vx, wb, lane, n, r.vx);
  st_run<float, PPT>(a.vy, wb, lane, n, r.vy);
  st_run<float, PPT>(a.x, wb, lane, n, r.x);
  st_run<float, PPT>(a.y, wb, lane, n, r.y);
  if (bad) flush_push_errors(a.errc, bad);
}

// Deposit of a pushed chunk through the wavefront's LDS window wl: each
// lane sums runs of equal cells in registers, runs go to the window (or
// straight to Q outside it), the window goes to Q.
template <int PPT>
__device__ __forceinline__ void deposit_chunk(const PushF32& a, int lane, unsigned long long* wl,
                                              const ChunkDep<PPT>& dp) {
  const int ny = a.ny;
  const int gi0 = wave_min_dpp(dp.imin), gi1 = wave_max_dpp(dp.imax);
  if (gi1 < 0) return;
  const int gj0 = wave_min_dpp(dp.jmin), gj1 = wave_max_dpp(dp.jmax);
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
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    if (dp.cp[k] == NO_DEPOSIT) continue;
    if (dp.cp[k] != rk) {
      if (rk != NO_DEPOSIT) {
        const unsigned long long acc[4] = {acc0, acc1, acc2, acc3};
        flush_block((int)(rk >> 16), (int)(rk & 0xffffu), ny, acc, 4, win, wl, a.Q);
      }
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
  if (rk != NO_DEPOSIT) {
    const unsigned long long acc[4] = {acc0, acc1, acc2, acc3};
    flush_block((int)(rk >> 16), (int)(rk & 0xffffu), ny, acc, 4, win, wl, a.Q);
  }
  wave_fence();
  const int W = win.wi * win.wj;
  for (int k = lane; k < W; k += 64) {
    const unsigned long long v = wl[k];
    if (v) {
      const int gi = win.i0 + k / win.wj, gj = win.j0 + k % win.wj;
      if (gi >= 0 && gj >= 0 && gi < a.nx && gj < a.ny) atomicAdd(&a.Q[(size_t)gi * ny + gj], v);
    }
  }
  wave_fence();
}

// One wavefront's chunk of 64*PPT particles starting at wb (n = particles in
// the array): push, walls, store, and (DEP) deposit of the new positions.
// Shared by the single-simulation kernel and the batched one, so both do
// identical arithmetic.
template <bool DEP, int PPT>
__device__ __forceinline__ void push_f32_chunk(const PushF32& a, int64_t wb, int64_t n, int lane,
                                               unsigned long long* wl) {
  ChunkRegs<PPT> r;
  ChunkDep<PPT> dp;
  load_chunk<PPT>(a, wb, n, lane, r);
  push_chunk_regs<DEP, PPT>(a, wb, n, lane, r, dp);
  if (DEP) deposit_chunk<PPT>(a, lane, wl, dp);
}

}  // namespace epic
