// Device building blocks of the particle kernels, shared by the
// single-simulation kernels (particles.hip) and the batched ensemble
// (batch.hip).
#pragma once

#include "device_common.hpp"
#include "epic_internal.hpp"

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

__device__ __forceinline__ void flush_block(int bi, int bj, int ny, const unsigned long long (&acc)[4],
                                            int ncorner, const Window& w, unsigned long long* lds,
                                            unsigned long long* Q) {
  const int li = bi - w.i0, lj = bj - w.j0;
  const bool in = li >= 0 && lj >= 0 && li + (ncorner > 1 ? 1 : 0) < w.wi && lj + (ncorner > 1 ? 1 : 0) < w.wj;
  if (in) {
    unsigned long long* b = lds + li * w.wj + lj;
    if (acc[0]) atomicAdd(b, acc[0]);
    if (ncorner > 1) {
      if (acc[1]) atomicAdd(b + w.wj, acc[1]);
      if (acc[2]) atomicAdd(b + 1, acc[2]);
      if (acc[3]) atomicAdd(b + w.wj + 1, acc[3]);
    }
  } else {
    unsigned long long* b = Q + (size_t)bi * ny + bj;
    if (acc[0]) atomicAdd(b, acc[0]);
    if (ncorner > 1) {
      if (acc[1]) atomicAdd(b + ny, acc[1]);
      if (acc[2]) atomicAdd(b + 1, acc[2]);
      if (acc[3]) atomicAdd(b + ny + 1, acc[3]);
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
//   bilinear gather as two lerps per component,
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
  const float2* E;
  float x0, y0, idx, idy, xhi, yhi, Lx, Ly;
  float kv, kx;  // c1*c2, c3*c2
  int nx, ny, bc;
  unsigned long long* Q;
  unsigned long long* errc;
};

__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

// One wavefront's chunk of 64*PPT particles starting at wb (n = particles in
// the array): push, walls, store, and (DEP) deposit of the new positions
// through the wavefront's LDS window wl.  Shared by the single-simulation
// kernel and the batched one, so both do identical arithmetic.
template <bool DEP, int PPT>
__device__ __forceinline__ void push_f32_chunk(const PushF32& a, int64_t wb, int64_t n, int lane,
                                               unsigned long long* wl) {
  const int ny = a.ny;
  const float fmax_i = (float)(a.nx - 1), fmax_j = (float)(a.ny - 1);
  float x[PPT], y[PPT], vx[PPT], vy[PPT];
  ld_run<float, PPT>(a.x, wb, lane, n, x);
  ld_run<float, PPT>(a.y, wb, lane, n, y);
  ld_run<float, PPT>(a.vx, wb, lane, n, vx);
  ld_run<float, PPT>(a.vy, wb, lane, n, vy);
  unsigned live = 0;
  int imin = 1 << 30, imax = -1, jmin = 1 << 30, jmax = -1;
  int ckey = -1;
  float c00x = 0, c10x = 0, c01x = 0, c11x = 0, c00y = 0, c10y = 0, c01y = 0, c11y = 0;
  int bad = 0;
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    if (!(part_index<float, PPT>(wb, lane, k) < n) || isnan(x[k])) continue;
    const float lx = (x[k] - a.x0) * a.idx;
    const float ly = (y[k] - a.y0) * a.idy;
    const int i = (int)lx, j = (int)ly;
    float ex = 0.f, ey = 0.f;
    if (i >= 0 && i + 1 < a.nx && j >= 0 && j + 1 < a.ny) {
      const int key = i * ny + j;
      if (key != ckey) {
        const float2 e00 = a.E[key], e01 = a.E[key + 1], e10 = a.E[key + ny], e11 = a.E[key + ny + 1];
        c00x = e00.x; c01x = e01.x; c10x = e10.x; c11x = e11.x;
        c00y = e00.y; c01y = e01.y; c10y = e10.y; c11y = e11.y;
        ckey = key;
      }
      const float di = lx - (float)i, dj = ly - (float)j;
      ex = lerpf(lerpf(c00x, c10x, di), lerpf(c01x, c11x, di), dj);
      ey = lerpf(lerpf(c00y, c10y, di), lerpf(c01y, c11y, di), dj);
    } else {
      bad++;
    }
    vx[k] = fmaf(a.kv, ex, vx[k]);
    vy[k] = fmaf(a.kv, ey, vy[k]);
    float px = fmaf(a.kx, vx[k], x[k]);
    float py = fmaf(a.kx, vy[k], y[k]);
    bool alive = true;
    if (a.bc == EPIC_BC_PBC) {
      px = px < a.x0 ? px + a.Lx : (px >= a.xhi ? px - a.Lx : px);
      py = py < a.y0 ? py + a.Ly : (py >= a.yhi ? py - a.Ly : py);
      // x0 - tiny + Lx can round up to xhi itself: the periodic image of x0
      px = px >= a.xhi ? px - a.Lx : px;
      py = py >= a.yhi ? py - a.Ly : py;
    } else if (a.bc == EPIC_BC_OPEN) {
      alive = !(px < a.x0 || px >= a.xhi || py < a.y0 || py >= a.yhi);
      if (!alive) px = __builtin_nanf("");
    } else {
      if (px < a.x0) { px = a.x0 + (a.x0 - px); vx[k] = -vx[k]; }
      else if (px >= a.xhi) { px = a.xhi - (px - a.xhi); vx[k] = -vx[k]; }
      if (py < a.y0) { py = a.y0 + (a.y0 - py); vy[k] = -vy[k]; }
      else if (py >= a.yhi) { py = a.yhi - (py - a.yhi); vy[k] = -vy[k]; }
    }
    x[k] = px;
    y[k] = py;
    if (DEP && alive) {
      const float nlx = (px - a.x0) * a.idx, nly = (py - a.y0) * a.idy;
      // a wrapped position can round onto the upper boundary node: it
      // deposits into the last cell
      if (nlx >= 0.f && nly >= 0.f && nlx <= fmax_i && nly <= fmax_j) {
        live |= 1u << k;
        const int ni = min((int)nlx, a.nx - 2), nj = min((int)nly, a.ny - 2);
        imin = min(imin, ni); imax = max(imax, ni);
        jmin = min(jmin, nj); jmax = max(jmax, nj);
      } else {
        bad++;
      }
    }
  }
  st_run<float, PPT>(a.vx, wb, lane, n, vx);
  st_run<float, PPT>(a.vy, wb, lane, n, vy);
  st_run<float, PPT>(a.x, wb, lane, n, x);
  st_run<float, PPT>(a.y, wb, lane, n, y);
  if (bad) atomicAdd(&a.errc[EPIC_ERRC_GATHER_OOB], (unsigned long long)bad);
  if (!DEP) return;

  const int gi0 = wave_min(imin), gi1 = wave_max(imax);
  if (gi1 < 0) return;
  const int gj0 = wave_min(jmin), gj1 = wave_max(jmax);
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

  int rk_i = -1, rk_j = -1;
  unsigned acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    if (!(live & (1u << k))) continue;
    const float nlx = (x[k] - a.x0) * a.idx, nly = (y[k] - a.y0) * a.idy;
    const int i = min((int)nlx, a.nx - 2), j = min((int)nly, a.ny - 2);
    const int dI = min((int)rintf((nlx - (float)i) * 16384.0f), 16384);
    const int dJ = min((int)rintf((nly - (float)j) * 16384.0f), 16384);
    if (i != rk_i || j != rk_j) {
      if (rk_i >= 0) {
        const unsigned long long acc[4] = {acc0, acc1, acc2, acc3};
        flush_block(rk_i, rk_j, ny, acc, 4, win, wl, a.Q);
      }
      rk_i = i;
      rk_j = j;
      acc0 = acc1 = acc2 = acc3 = 0;
    }
    const unsigned oI = 16384u - (unsigned)dI, oJ = 16384u - (unsigned)dJ;
    acc0 += oI * oJ;
    acc1 += (unsigned)dI * oJ;
    acc2 += oI * (unsigned)dJ;
    acc3 += (unsigned)dI * (unsigned)dJ;
  }
  if (rk_i >= 0) {
    const unsigned long long acc[4] = {acc0, acc1, acc2, acc3};
    flush_block(rk_i, rk_j, ny, acc, 4, win, wl, a.Q);
  }
  wave_fence();
  const int W = win.wi * win.wj;
  for (int k = lane; k < W; k += 64) {
    const unsigned long long v = wl[k];
    if (v) {
      const int gi = win.i0 + k / win.wj, gj = win.j0 + k % win.wj;
      if (gi < a.nx && gj < a.ny) atomicAdd(&a.Q[(size_t)gi * ny + gj], v);
    }
  }
  wave_fence();
}

}  // namespace epic
