// Particle kernels for gfx950: fused CIC/NGP gather + leapfrog push + boundary
// handling + fixed-point deposit of the new position, the stand-alone deposit
// (Species::ScatterSpecies_serial), and the half-step rewind.
//
// Layout: structure-of-arrays x, y, vx, vy, vz (FP64 or FP32) per species.
// Each thread owns PPT consecutive particles (32 bytes per array, two 16-byte
// loads), so a wavefront streams 2 KiB per array per load pair, and, when the
// species is sorted by cell, a thread's particles share a cell: the gather
// reuses its four node values and the deposit keeps a run-length register
// accumulator that touches LDS only when the cell changes.
//
// Deposit: int64 fixed point (weight * 2^B rounded; the four CIC weights of a
// particle sum to exactly 2^B), accumulated per workgroup in an LDS window of
// node accumulators and flushed to the global int64 node array with atomics.
// Integer addition is associative, so the deposited density does not depend
// on thread scheduling: runs are bit-identical.
#include "push_f32.hpp"

namespace epic {

enum { MODE_PUSH = 0, MODE_REWIND = 1, MODE_DEPOSIT = 2 };

template <typename P>
struct PartArgs {
  P* x;
  P* y;
  P* vx;
  P* vy;
  const int64_t* d_n;  // device slot count
  int64_t n_fixed;     // >= 0: process [0, n_fixed) instead of [0, *d_n)
  const Vec2<P>* E;    // (ex, ey) per node, i*ny + j
  P x0, y0, dx, dy, xhi, yhi, Lx, Ly;
  P idx, idy;          // 1/dx, 1/dy (FP32 deposit coordinates)
  int nx, ny;
  int bc, shape;
  P c1, c2, c3;        // dv = (c1*E)*c2 ; dpos = (c3*v)*c2
  unsigned long long* Q;
  int fix_bits;
  unsigned long long* errc;
  int rr;              // chunk walk: 0 per-XCD contiguous ranges, 1 round-robin
  int fastdiv;         // FP64: dx, dy in [2^-400, 2^400], so x_to_l may divide by reciprocal + correction
  const CondDev* cond; // absorbing conductors (the first ncond entries are active)
  int ncond;
};

template <typename P>
__device__ __forceinline__ P p_nan();
template <>
__device__ __forceinline__ double p_nan<double>() {
  return __builtin_nan("");
}
template <>
__device__ __forceinline__ float p_nan<float>() {
  return __builtin_nanf("");
}

// rint(w * scale) as an integer.  For |w * scale| < 2^51 the sum
// w*scale + 1.5*2^52 lies in [2^52, 2^53), where doubles are exactly the
// integers: the addition rounds w*scale to the nearest integer, ties to even
// (1.5*2^52 is even), and the bit pattern minus that of 1.5*2^52 is it.
// Deposit weights are below 4 in magnitude (bounds-checked cells) and the
// scale is 2^44.
static_assert(FIX_BITS_FP64 <= 48, "fix_round needs |w| * 2^fix_bits < 2^51");
__device__ __forceinline__ long long fix_round(double w, double scale) {
  const double t = w * scale + 0x1.8p52;
  return __double_as_longlong(t) - 0x4338000000000000ll;
}
__device__ __forceinline__ long long fix_round(float w, float scale) { return (long long)rintf(w * scale); }

// (x - x0) / dx, the reference's XtoL/YtoL (domain.cpp:257-264), correctly
// rounded.  FP64: q = RN(a * r) with r = RN(1/dx), then two correction steps
// q = RN(q + RN(a - q*dx) * r) (fused multiply-adds).  The first brings q
// within one ulp of a/dx; the second is then exactly RN(a/dx) by Markstein's
// theorem (P. Markstein, "Computation of elementary functions on the IBM
// RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990: r within half
// an ulp of 1/b and q within one ulp of a/b give RN(q + (a - bq) r) =
// RN(a/b)).  The theorem needs the remainder and products to stay normal:
// |a| and dx in [2^-400, 2^400]; anything else (a == 0, infinities, NaN)
// takes the division instruction sequence.
template <typename P>
__device__ __forceinline__ P x_to_l(const PartArgs<P>& a, P x, P x0, P d, P rd) {
  return (x - x0) / d;
}
template <>
__device__ __forceinline__ double x_to_l<double>(const PartArgs<double>& a, double x, double x0, double d,
                                                 double rd) {
  const double n = x - x0;
  const double m = fabs(n);
  double q;
  if (a.fastdiv && m >= 0x1p-400 && m <= 0x1p400) {
    q = n * rd;
    double r = fma(-q, d, n);
    q = fma(r, rd, q);
    r = fma(-q, d, n);
    q = fma(r, rd, q);
  } else {
    q = n / d;
  }
  return q;
}

// One gather of (ex, ey) at logical position (lx, ly), following
// Domain::Interpolate (domain.cpp:231-255) operation for operation.
// Returns false when the reference's bounds-checked access would throw.
template <typename P>
__device__ __forceinline__ bool gather(const PartArgs<P>& a, P lx, P ly, int& ckey, P (&ce)[8], P& ex,
                                       P& ey) {
  if (a.shape == EPIC_SHAPE_NGP) {
    const int gi = (int)round(lx), gj = (int)round(ly);
    if (gi < 0 || gi >= a.nx || gj < 0 || gj >= a.ny) return false;
    const Vec2<P> e = a.E[(size_t)gi * a.ny + gj];
    ex = e.a;
    ey = e.b;
    return true;
  }
  const int i = (int)lx, j = (int)ly;
  if (i < 0 || i + 1 >= a.nx || j < 0 || j + 1 >= a.ny) return false;
  const P di = lx - (P)i, dj = ly - (P)j;
  const int key = i * a.ny + j;
  if (key != ckey) {
    const Vec2<P> e00 = a.E[key], e10 = a.E[key + a.ny], e01 = a.E[key + 1], e11 = a.E[key + a.ny + 1];
    ce[0] = e00.a; ce[1] = e10.a; ce[2] = e01.a; ce[3] = e11.a;
    ce[4] = e00.b; ce[5] = e10.b; ce[6] = e01.b; ce[7] = e11.b;
    ckey = key;
  }
  const P omdi = (P)1 - di, omdj = (P)1 - dj;
  ex = ce[0] * omdi * omdj + ce[1] * (di)*omdj + ce[2] * omdi * (dj) + ce[3] * (di) * (dj);
  ey = ce[4] * omdi * omdj + ce[5] * (di)*omdj + ce[6] * omdi * (dj) + ce[7] * (di) * (dj);
  return true;
}

// Cell of a position for the deposit: CIC base node (floor) or NGP node.
// Returns false where the reference's bounds-checked scatter would throw.
template <typename P>
__device__ __forceinline__ bool dep_cell(const PartArgs<P>& a, P x, P y, int& i, int& j, P& di, P& dj) {
  // FP64: the reference's XtoL division; FP32: the hot kernel's reciprocal
  const P lx = sizeof(P) == 8 ? x_to_l(a, x, a.x0, a.dx, a.idx) : (x - a.x0) * a.idx;
  const P ly = sizeof(P) == 8 ? x_to_l(a, y, a.y0, a.dy, a.idy) : (y - a.y0) * a.idy;
  if (a.shape == EPIC_SHAPE_NGP) {
    i = (int)round(lx);
    j = (int)round(ly);
    di = dj = 0;
    return i >= 0 && i < a.nx && j >= 0 && j < a.ny;
  }
  i = (int)lx;
  j = (int)ly;
  di = lx - (P)i;
  dj = ly - (P)j;
  return i >= 0 && i + 1 < a.nx && j >= 0 && j + 1 < a.ny;
}

// A chunk known to be full: the vector loads only.
template <typename P, int PPT>
__device__ __forceinline__ void ld_run_full(const P* __restrict__ a, int64_t wbase, int lane, P (&o)[PPT]) {
  constexpr int V = 16 / (int)sizeof(P);
#pragma unroll
  for (int g = 0; g < PPT / V; g++) {
    const v4u32 t = __builtin_nontemporal_load(
        reinterpret_cast<const v4u32*>(a + pslot<P>(wbase + (int64_t)g * 64 * V) + (int64_t)lane * V));
    __builtin_memcpy(&o[g * V], &t, 16);
  }
}

// Deposit cells of a thread's particles: base node (i, j), offsets (di, dj),
// -1 for dead or out-of-grid particles, and the thread's bounding box.
template <typename P, int PPT>
struct DepCells {
  int ci[PPT], cj[PPT];
  P di[PPT], dj[PPT];
  int imin, imax, jmin, jmax;
  int bad;  // live particles outside the grid (the reference's scatter throws)
};

template <typename P, int PPT>
__device__ __forceinline__ void dep_cells(const PartArgs<P>& a, const bool (&live)[PPT], const P (&x)[PPT],
                                          const P (&y)[PPT], DepCells<P, PPT>& d) {
  d.imin = 1 << 30; d.imax = -1; d.jmin = 1 << 30; d.jmax = -1;
  d.bad = 0;
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    d.ci[k] = -1;
    d.cj[k] = 0;
    d.di[k] = d.dj[k] = 0;
    if (!live[k]) continue;
    int i, j;
    P di, dj;
    if (!dep_cell(a, x[k], y[k], i, j, di, dj)) {
      d.bad++;
      continue;
    }
    d.ci[k] = i;
    d.cj[k] = j;
    d.di[k] = di;
    d.dj[k] = dj;
    d.imin = min(d.imin, i); d.imax = max(d.imax, i);
    d.jmin = min(d.jmin, j); d.jmax = max(d.jmax, j);
  }
}

// Deposit of one chunk's cells.  WHOLE: into the workgroup's whole-grid LDS
// histogram.  Otherwise the wavefront's LDS window is placed on the bounding
// box of its cells, filled, and flushed to the global accumulator.
template <typename P, int PPT, bool WHOLE>
__device__ __forceinline__ void deposit_cells(const PartArgs<P>& a, int lane, unsigned long long* wl, Window& win,
                                              const DepCells<P, PPT>& d) {
  const int ny = a.ny;
  const int ncorner = a.shape == EPIC_SHAPE_NGP ? 1 : 4;
  const P scale = (P)ldexp(1.0, a.fix_bits);
  const long long one = 1ll << a.fix_bits;
  bool any = true;
  if (!WHOLE) {
    const int gi0 = wave_min_dpp(d.imin), gi1 = wave_max_dpp(d.imax);
    any = gi1 >= 0;  // uniform
    if (any) {
      const int gj0 = wave_min_dpp(d.jmin), gj1 = wave_max_dpp(d.jmax);
      const int wi = gi1 - gi0 + 2, wj = gj1 - gj0 + 2;
      if (wi * wj <= WAVE_WIN) {
        win = Window{gi0, gj0, wi, wj};
      } else {
        // too wide (drifted outliers): a window on the wavefront's first
        // cells, the rest goes straight to global atomics
        const int wjc = min(wj, 32);
        win = Window{gi0, gj0, WAVE_WIN / wjc, wjc};
      }
      for (int k = lane; k < win.wi * win.wj; k += 64) wl[k] = 0ull;
      wave_fence();
    }
  }
  if (any) {
    // run-length accumulation over this thread's consecutive particles
    int rk_i = -1, rk_j = -1;
    unsigned long long acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      if (d.ci[k] < 0) continue;
      const int i = d.ci[k], j = d.cj[k];
      const P di = d.di[k], dj = d.dj[k];
      if (i != rk_i || j != rk_j) {
        if (rk_i >= 0) flush_block(rk_i, rk_j, ny, acc, ncorner, win, wl, a.Q);
        rk_i = i;
        rk_j = j;
        acc[0] = acc[1] = acc[2] = acc[3] = 0;
      }
      if (ncorner == 1) {
        acc[0] += (unsigned long long)one;
      } else if (sizeof(P) == 4) {
        // FP32 storage: 14-bit offsets, weights sum to 2^28 exactly
        const unsigned dI = (unsigned)min((int)rintf((float)di * 16384.0f), 16384);
        const unsigned dJ = (unsigned)min((int)rintf((float)dj * 16384.0f), 16384);
        const unsigned oI = 16384u - dI, oJ = 16384u - dJ;
        acc[0] += oI * oJ;
        acc[1] += dI * oJ;
        acc[2] += oI * dJ;
        acc[3] += dI * dJ;
      } else {
        const P omdi = (P)1 - di, omdj = (P)1 - dj;
        const long long q00 = fix_round(omdi * omdj, scale);
        const long long q10 = fix_round(di * omdj, scale);
        const long long q01 = fix_round(omdi * dj, scale);
        const long long q11 = one - q00 - q10 - q01;
        acc[0] += (unsigned long long)q00;
        acc[1] += (unsigned long long)q10;
        acc[2] += (unsigned long long)q01;
        acc[3] += (unsigned long long)q11;
      }
    }
    if (rk_i >= 0) flush_block(rk_i, rk_j, ny, acc, ncorner, win, wl, a.Q);

    if (!WHOLE) {
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
  }
  if (d.bad) atomicAdd(&a.errc[EPIC_ERRC_DEPOSIT_OOB], (unsigned long long)d.bad);
}

// WHOLE: the whole node grid is one LDS histogram per workgroup, flushed once
// at the end (small grids).  Otherwise each wavefront owns an LDS window of
// WAVE_WIN node accumulators placed on the bounding box of the cells its
// 64*PPT particles deposit to; deposits outside go to global atomics.  A
// wavefront never waits on the rest of its workgroup.
//
// Each particle's deposit cell and offsets are computed once, after its push,
// and kept for the deposit; the next chunk's particles are loaded before this
// chunk's deposit (software pipeline).  The arithmetic of every particle is
// the reference's, operation for operation, whatever the schedule.
template <typename P, int PPT, int MODE, bool WHOLE, bool DEP>
__global__ __launch_bounds__(PART_BLOCK) void particle_kernel(PartArgs<P> a) {
  constexpr int LDS_N = DEP ? (WHOLE ? DEP_WINDOW : (PART_BLOCK / 64) * WAVE_WIN) : 1;
  __shared__ unsigned long long lds[LDS_N];

  constexpr int CHW = 64 * PPT;          // particles per wavefront chunk
  constexpr int CH = PART_BLOCK * PPT;   // per workgroup chunk
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long* wl = WHOLE ? lds : lds + wid * WAVE_WIN;
  const int64_t n = a.n_fixed >= 0 ? a.n_fixed : *a.d_n;
  const int64_t nchunks = (n + CH - 1) / CH;
  Window win{0, 0, a.nx, a.ny};
  if (DEP && WHOLE) {
    for (int k = threadIdx.x; k < a.nx * a.ny; k += blockDim.x) lds[k] = 0ull;
    __syncthreads();
  }

  int64_t cb, ce_, cs;
  if (a.rr) {
    cb = blockIdx.x;
    ce_ = nchunks;
    cs = gridDim.x;
  } else {
    xcd_chunk_range(nchunks, &cb, &ce_, &cs);
  }
  P x[PPT], y[PPT], vx[PPT], vy[PPT];
  if (cb < ce_) {
    const int64_t wb0 = cb * CH + (int64_t)wid * CHW;
    ld_run<P, PPT>(a.x, wb0, lane, n, x);
    ld_run<P, PPT>(a.y, wb0, lane, n, y);
    if (MODE != MODE_DEPOSIT) {
      ld_run<P, PPT>(a.vx, wb0, lane, n, vx);
      ld_run<P, PPT>(a.vy, wb0, lane, n, vy);
    }
  }
  for (int64_t chunk = cb; chunk < ce_; chunk += cs) {
    const int64_t wb = chunk * CH + (int64_t)wid * CHW;
    bool live[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) live[k] = (part_index<P, PPT>(wb, lane, k) < n) && !isnan(x[k]);

    if (MODE != MODE_DEPOSIT) {
      int ckey = -1;
      P cev[8];
      int bad = 0;
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        if (!live[k]) continue;
        const P lx = x_to_l(a, x[k], a.x0, a.dx, a.idx);
        const P ly = x_to_l(a, y[k], a.y0, a.dy, a.idy);
        P ex = 0, ey = 0;
        if (!gather(a, lx, ly, ckey, cev, ex, ey)) bad++;
        if (MODE == MODE_REWIND) {
          vx[k] = vx[k] - (a.c1 * ex) * a.c2;
          vy[k] = vy[k] - (a.c1 * ey) * a.c2;
          continue;
        }
        vx[k] = vx[k] + (a.c1 * ex) * a.c2;
        vy[k] = vy[k] + (a.c1 * ey) * a.c2;
        P px = x[k] + (a.c3 * vx[k]) * a.c2;
        P py = y[k] + (a.c3 * vy[k]) * a.c2;
        if (isnan(px) || isnan(py)) {
          bad += NAN_UNIT;
          px = p_nan<P>();
          live[k] = false;
        } else if (a.bc == EPIC_BC_PBC) {
          if (px < a.x0) px = px + a.Lx;
          else if (px >= a.xhi) px = px - a.Lx;
          if (py < a.y0) py = py + a.Ly;
          else if (py >= a.yhi) py = py - a.Ly;
          if (sizeof(P) == 4) {
            // FP32: x0 - tiny + Lx can round up to xhi itself; that position
            // is the periodic image of x0
            if (px >= a.xhi) px = px - a.Lx;
            if (py >= a.yhi) py = py - a.Ly;
          }
        } else if (a.bc == EPIC_BC_OPEN) {
          if (px < a.x0 || px >= a.xhi || py < a.y0 || py >= a.yhi) {
            px = p_nan<P>();
            live[k] = false;
          }
        } else {
          if (px < a.x0) { px = a.x0 + (a.x0 - px); vx[k] = -vx[k]; }
          else if (px >= a.xhi) { px = a.xhi - (px - a.xhi); vx[k] = -vx[k]; }
          if (py < a.y0) { py = a.y0 + (a.y0 - py); vy[k] = -vy[k]; }
          else if (py >= a.yhi) { py = a.yhi - (py - a.yhi); vy[k] = -vy[k]; }
          if (sizeof(P) == 4) {
            // FP32: a position exactly on the upper wall mirrors onto
            // itself; keep live positions below it (the FP64 path keeps the
            // reference's arithmetic)
            if (px >= a.xhi) px = nextafter(a.xhi, a.x0);
            if (py >= a.yhi) py = nextafter(a.yhi, a.y0);
          }
        }
        // a conductor absorbs what the walls left (species.cpp:84-98,
        // 111-124, 152-165)
        if (a.ncond > 0 && live[k] && in_conductor(a.cond, a.ncond, (double)px, (double)py)) {
          px = p_nan<P>();
          live[k] = false;
        }
        x[k] = px;
        y[k] = py;
      }
      if (bad) flush_push_errors(a.errc, bad);
    }

    // deposit cells and offsets of the (new) positions, computed once
    DepCells<P, PPT> dc;
    if (DEP) dep_cells<P, PPT>(a, live, x, y, dc);
    // every gather fetch has been consumed: nothing outstanding but the
    // stores below and the next chunk's loads (this keeps the compiler from
    // draining those loads before the deposit)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (MODE != MODE_DEPOSIT) {
      st_run<P, PPT>(a.vx, wb, lane, n, vx);
      st_run<P, PPT>(a.vy, wb, lane, n, vy);
      if (MODE == MODE_PUSH) {
        st_run<P, PPT>(a.x, wb, lane, n, x);
        st_run<P, PPT>(a.y, wb, lane, n, y);
      }
    }

    // next chunk's particles (full chunks now, a partial one after the deposit)
    const int64_t nb = (chunk + cs) * CH + (int64_t)wid * CHW;
    const bool more = chunk + cs < ce_;
    const bool full = more && nb + CHW <= n;
    if (full) {
      ld_run_full<P, PPT>(a.x, nb, lane, x);
      ld_run_full<P, PPT>(a.y, nb, lane, y);
      if (MODE != MODE_DEPOSIT) {
        ld_run_full<P, PPT>(a.vx, nb, lane, vx);
        ld_run_full<P, PPT>(a.vy, nb, lane, vy);
      }
    }

    if (DEP) deposit_cells<P, PPT, WHOLE>(a, lane, wl, win, dc);

    if (more && !full) {
      ld_run<P, PPT>(a.x, nb, lane, n, x);
      ld_run<P, PPT>(a.y, nb, lane, n, y);
      if (MODE != MODE_DEPOSIT) {
        ld_run<P, PPT>(a.vx, nb, lane, n, vx);
        ld_run<P, PPT>(a.vy, nb, lane, n, vy);
      }
    }
  }

  if (DEP && WHOLE) {
    __syncthreads();
    const int W = a.nx * a.ny;
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
      const unsigned long long v = lds[k];
      if (v) atomicAdd(&a.Q[k], v);
    }
  }
}

// Buffer resource over one wavefront chunk of an FP64 particle array: the
// chw elements from wb, of which those below n are live.  Buffer loads past
// the range return zeros without a memory access, so every chunk (the last,
// partial one and one past the end included) is loaded by the same
// instructions.  wb and n are wavefront-uniform.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t chunk_rsrc(const double* p, int64_t wb, int64_t n, int chw) {
  const int64_t left = n - wb;
  const int recs = left <= 0 ? 0 : (int)((left < chw ? left : (int64_t)chw) * (int64_t)sizeof(double));
  return __builtin_amdgcn_make_buffer_rsrc((void*)(p + (left <= 0 ? 0 : wb)), (short)0, recs, 0x00020000);
}

template <int PPT>
__device__ __forceinline__ void ld_chunk_buf(__amdgpu_buffer_rsrc_t r, int lane, double (&o)[PPT]) {
#pragma unroll
  for (int g = 0; g < PPT / 2; g++) {
    // cache policy 2: non-temporal (streamed once)
    const v4u32 t = __builtin_amdgcn_raw_buffer_load_b128(r, (g * 128 + lane * 2) * 8, 0, 2);
    __builtin_memcpy(&o[2 * g], &t, 16);
  }
}

// FP64 CIC push + deposit over a windowed grid (the plain path's hot
// kernel).  Per chunk of 64*PPT particles a wavefront
//   1. computes the cells of the current positions (Domain::Interpolate's
//      logical coordinates, bounds and offsets),
//   2. issues the field loads of all its particles' cells at once,
//   3. issues the next chunk's particle loads (buffer loads, so the last,
//      partial chunk needs no separate path),
//   4. pushes (waiting only for the field loads: the next chunk stays in
//      flight), stores, and deposits the new positions.
// Every particle's arithmetic is that of particle_kernel<double> and of the
// reference, operation for operation.
template <int PPT, bool DEP, bool COND = false>
__global__ __launch_bounds__(PART_BLOCK) void push_f64_kernel(PartArgs<double> a) {
  using P = double;
  __shared__ unsigned long long lds[DEP ? (PART_BLOCK / 64) * WAVE_WIN : 1];
  constexpr int CHW = 64 * PPT;
  constexpr int CH = PART_BLOCK * PPT;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned long long* wl = lds + (DEP ? wid * WAVE_WIN : 0);
  const int64_t n = a.n_fixed >= 0 ? a.n_fixed : *a.d_n;
  const int64_t nchunks = (n + CH - 1) / CH;
  const int ny = a.ny;
  Window win{0, 0, a.nx, a.ny};
  // the (ex, ey) node pairs, 16 bytes each (the host keeps nx*ny*16 < 2^31)
  const __amdgpu_buffer_rsrc_t erc =
      __builtin_amdgcn_make_buffer_rsrc((void*)a.E, (short)0, a.nx * ny * 16, 0x00020000);

  int64_t cb, ce_, cs;
  if (a.rr) {
    cb = blockIdx.x;
    ce_ = nchunks;
    cs = gridDim.x;
  } else {
    xcd_chunk_range(nchunks, &cb, &ce_, &cs);
  }
  P x[PPT], y[PPT], vx[PPT], vy[PPT];
  {
    const int64_t wb0 = cb * CH + (int64_t)wid * CHW;
    ld_chunk_buf<PPT>(chunk_rsrc(a.x, wb0, n, CHW), lane, x);
    ld_chunk_buf<PPT>(chunk_rsrc(a.y, wb0, n, CHW), lane, y);
    ld_chunk_buf<PPT>(chunk_rsrc(a.vx, wb0, n, CHW), lane, vx);
    ld_chunk_buf<PPT>(chunk_rsrc(a.vy, wb0, n, CHW), lane, vy);
  }
  for (int64_t chunk = cb; chunk < ce_; chunk += cs) {
    const int64_t wb = chunk * CH + (int64_t)wid * CHW;
    bool live[PPT], ok[PPT];
    int key[PPT];
    P gdi[PPT], gdj[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      live[k] = (part_index<P, PPT>(wb, lane, k) < n) && !isnan(x[k]);
      const P lx = live[k] ? x_to_l(a, x[k], a.x0, a.dx, a.idx) : 0.0;
      const P ly = live[k] ? x_to_l(a, y[k], a.y0, a.dy, a.idy) : 0.0;
      const int i = (int)lx, j = (int)ly;
      ok[k] = live[k] && i >= 0 && i + 1 < a.nx && j >= 0 && j + 1 < ny;
      key[k] = ok[k] ? i * ny + j : 0;
      gdi[k] = lx - (P)i;
      gdj[k] = ly - (P)j;
    }
    Vec2<P> e00[PPT], e10[PPT], e01[PPT], e11[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int o = key[k] * 16;
      v4u32 t = __builtin_amdgcn_raw_buffer_load_b128(erc, o, 0, 0);
      __builtin_memcpy(&e00[k], &t, 16);
      t = __builtin_amdgcn_raw_buffer_load_b128(erc, o + ny * 16, 0, 0);
      __builtin_memcpy(&e10[k], &t, 16);
      t = __builtin_amdgcn_raw_buffer_load_b128(erc, o + 16, 0, 0);
      __builtin_memcpy(&e01[k], &t, 16);
      t = __builtin_amdgcn_raw_buffer_load_b128(erc, o + ny * 16 + 16, 0, 0);
      __builtin_memcpy(&e11[k], &t, 16);
    }
    P xn[PPT], yn[PPT], vxn[PPT], vyn[PPT];
    {
      const int64_t nb = (chunk + cs) * CH + (int64_t)wid * CHW;
      const int64_t nlim = chunk + cs < ce_ ? n : 0;
      ld_chunk_buf<PPT>(chunk_rsrc(a.x, nb, nlim, CHW), lane, xn);
      ld_chunk_buf<PPT>(chunk_rsrc(a.y, nb, nlim, CHW), lane, yn);
      ld_chunk_buf<PPT>(chunk_rsrc(a.vx, nb, nlim, CHW), lane, vxn);
      ld_chunk_buf<PPT>(chunk_rsrc(a.vy, nb, nlim, CHW), lane, vyn);
    }
    // Wait here for the field values (and only for them: the next chunk's
    // loads were issued after them and stay in flight), and keep the
    // compiler from sinking any field load into the branches below.
#pragma unroll
    for (int k = 0; k < PPT; k++)
      asm volatile("" ::"v"(e00[k].a), "v"(e00[k].b), "v"(e10[k].a), "v"(e10[k].b), "v"(e01[k].a), "v"(e01[k].b),
                   "v"(e11[k].a), "v"(e11[k].b));
    int bad = 0;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      if (!live[k]) continue;
      P ex = 0, ey = 0;
      if (ok[k]) {
        // Domain::Interpolate (domain.cpp:231-255), as gather()
        const P di = gdi[k], dj = gdj[k];
        const P omdi = (P)1 - di, omdj = (P)1 - dj;
        ex = e00[k].a * omdi * omdj + e10[k].a * (di)*omdj + e01[k].a * omdi * (dj) + e11[k].a * (di) * (dj);
        ey = e00[k].b * omdi * omdj + e10[k].b * (di)*omdj + e01[k].b * omdi * (dj) + e11[k].b * (di) * (dj);
      } else {
        bad++;
      }
      vx[k] = vx[k] + (a.c1 * ex) * a.c2;
      vy[k] = vy[k] + (a.c1 * ey) * a.c2;
      P px = x[k] + (a.c3 * vx[k]) * a.c2;
      P py = y[k] + (a.c3 * vy[k]) * a.c2;
      if (isnan(px) || isnan(py)) {
        bad += NAN_UNIT;
        px = p_nan<P>();
        live[k] = false;
      } else if (a.bc == EPIC_BC_PBC) {
        if (px < a.x0) px = px + a.Lx;
        else if (px >= a.xhi) px = px - a.Lx;
        if (py < a.y0) py = py + a.Ly;
        else if (py >= a.yhi) py = py - a.Ly;
      } else if (a.bc == EPIC_BC_OPEN) {
        if (px < a.x0 || px >= a.xhi || py < a.y0 || py >= a.yhi) {
          px = p_nan<P>();
          live[k] = false;
        }
      } else {
        if (px < a.x0) { px = a.x0 + (a.x0 - px); vx[k] = -vx[k]; }
        else if (px >= a.xhi) { px = a.xhi - (px - a.xhi); vx[k] = -vx[k]; }
        if (py < a.y0) { py = a.y0 + (a.y0 - py); vy[k] = -vy[k]; }
        else if (py >= a.yhi) { py = a.yhi - (py - a.yhi); vy[k] = -vy[k]; }
      }
      if (COND && live[k] && in_conductor(a.cond, a.ncond, px, py)) {
        px = p_nan<P>();
        live[k] = false;
      }
      x[k] = px;
      y[k] = py;
    }
    if (bad) flush_push_errors(a.errc, bad);
    DepCells<P, PPT> dc;
    if (DEP) dep_cells<P, PPT>(a, live, x, y, dc);
    st_run<P, PPT>(a.vx, wb, lane, n, vx);
    st_run<P, PPT>(a.vy, wb, lane, n, vy);
    st_run<P, PPT>(a.x, wb, lane, n, x);
    st_run<P, PPT>(a.y, wb, lane, n, y);
    if (DEP) deposit_cells<P, PPT, false>(a, lane, wl, win, dc);
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      x[k] = xn[k];
      y[k] = yn[k];
      vx[k] = vxn[k];
      vy[k] = vyn[k];
    }
  }
}

// rr = 0: each XCD walks its own contiguous range of chunks; rr = 1: all
// workgroups walk the chunks round-robin (the whole device streams through
// one region at a time).
template <bool DEP, int PPT, int WAVES, bool COND = false, int BC = BC_ANY>
__global__ __launch_bounds__(PART_BLOCK, WAVES) void push_f32_kernel(PushF32 a, int rr, const CondDev* cond,
                                                                     int ncond, const uint32_t* cmask) {
  __shared__ unsigned long long lds[DEP ? (PART_BLOCK / 64) * WAVE_WIN : 1];
  constexpr int CHW = 64 * PPT;
  constexpr int CH = PART_BLOCK * PPT;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long* wl = lds + (DEP ? wid * WAVE_WIN : 0);
  const int64_t n = *a.d_n;
  const int64_t nchunks = (n + CH - 1) / CH;
  int64_t cb, ce_, cs;
  if (rr) {
    cb = blockIdx.x;
    ce_ = nchunks;
    cs = gridDim.x;
  } else {
    xcd_chunk_range(nchunks, &cb, &ce_, &cs);
  }
  // software pipeline: the next chunk's loads go out (into the registers the
  // stores just freed) before this chunk's deposit
  // (full chunks only: a partial last chunk is loaded after the deposit)
  ChunkRegs<PPT> r;
  if (cb < ce_) load_chunk<PPT>(a, cb * CH + (int64_t)wid * CHW, n, lane, r);
  for (int64_t chunk = cb; chunk < ce_; chunk += cs) {
    ChunkDep<PPT> dp;
    push_chunk_regs<DEP, PPT, COND, BC>(a, chunk * CH + (int64_t)wid * CHW, n, lane, r, dp, cond, ncond, cmask);
    const int64_t nb = (chunk + cs) * CH + (int64_t)wid * CHW;
    const bool more = chunk + cs < ce_;
    const bool full = more && nb + 64 * PPT <= n;
    if (full) load_chunk_full<PPT>(a, nb, lane, r);
    if (DEP) deposit_chunk<PPT, BC == EPIC_BC_PBC>(a, lane, wl, dp);
    if (more && !full) load_chunk<PPT>(a, nb, n, lane, r);
  }
}

// All pushed FP32 species of a context in one launch: the species' chunk
// ranges are concatenated (species s owns global chunks [cbeg_s, cend_s)),
// and the workgroups walk the global chunk index round-robin with the
// software pipeline of push_f32_kernel, the next chunk possibly of the next
// species.  Per-particle arithmetic is push_f32_kernel's.
constexpr int PUSH_MAXSP = 8;

// A wavefront-uniform 64-bit value moved into scalar registers.
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(uint64_t)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
struct PushF32Set {
  PushF32 sp[PUSH_MAXSP];
  int nsp;
  const CondDev* cond;  // absorbing conductors (the first ncond are active)
  int ncond;
  const uint32_t* cmask;  // cells where a conductor can be (one bit per cell)
};

template <int PPT, bool COND, int WAVES, int BC = BC_ANY>
__global__ __launch_bounds__(PART_BLOCK, WAVES) void push_f32_multi_kernel(PushF32Set S) {
  __shared__ unsigned long long lds[(PART_BLOCK / 64) * WAVE_WIN];
  constexpr int CHW = 64 * PPT;
  constexpr int CH = PART_BLOCK * PPT;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long* wl = lds + wid * WAVE_WIN;
  // slot counts per species (LDS: indexed by a runtime species number)
  __shared__ int64_t ns[PUSH_MAXSP];
  if (threadIdx.x < PUSH_MAXSP) ns[threadIdx.x] = (int)threadIdx.x < S.nsp ? *S.sp[threadIdx.x].d_n : 0;
  __syncthreads();
  // the chunk ranges are the same for every lane: read into scalar
  // registers as 32-bit chunk counts (a species holds fewer than 2^31 * CH
  // slots), so that locating a chunk's species is scalar work
  int cend[PUSH_MAXSP];
  int tot = 0;
#pragma unroll
  for (int k = 0; k < PUSH_MAXSP; k++) {
    tot += (int)((uniform_i64(ns[k]) + CH - 1) / CH);
    cend[k] = tot;
  }
  // species and first slot of global chunk g (wavefront-uniform)
  auto locate = [&](int g, int& sp, int64_t& wb, int64_t& n) {
    int s = 0;
    int beg = 0;
#pragma unroll
    for (int k = 0; k + 1 < PUSH_MAXSP; k++)
      if (g >= cend[k]) {
        s = k + 1;
        beg = cend[k];
      }
    sp = s;
    n = uniform_i64(ns[s]);
    wb = (int64_t)(g - beg) * CH + (int64_t)wid * CHW;
  };
  int g = blockIdx.x;
  if (g >= tot) return;
  int s;
  int64_t wb, n;
  locate(g, s, wb, n);
  PushF32 a = S.sp[s];
  ChunkRegs<PPT> r;
  if (wb < n) load_chunk<PPT>(a, wb, n, lane, r);
  for (; g < tot; g += gridDim.x) {
    ChunkDep<PPT> dp;
    const bool active = wb < n;
    if (active) push_chunk_regs<true, PPT, COND, BC>(a, wb, n, lane, r, dp, S.cond, S.ncond, S.cmask);
    const int gn = g + (int)gridDim.x;
    const bool more = gn < tot;
    int sn = s;
    int64_t wbn = 0, nn = 0;
    if (more) locate(gn, sn, wbn, nn);
    const PushF32 an = S.sp[sn];
    const bool full = more && wbn + CHW <= nn;
    if (full) load_chunk_full<PPT>(an, wbn, lane, r);
    if (active) deposit_chunk<PPT, BC == EPIC_BC_PBC>(a, lane, wl, dp);
    if (more && !full && wbn < nn) load_chunk<PPT>(an, wbn, nn, lane, r);
    a = an;
    s = sn;
    n = nn;
    wb = wbn;
  }
}

// ---------------------------------------------------------------------------
template <typename P>
static PartArgs<P> make_args(Ctx* c, int s) {
  SpeciesDev& S = c->sp[s];
  const epic_domain_desc& d = c->dom;
  PartArgs<P> a{};
  a.x = (P*)S.x; a.y = (P*)S.y; a.vx = (P*)S.vx; a.vy = (P*)S.vy;
  a.d_n = S.d_n;
  a.n_fixed = -1;
  a.E = (const Vec2<P>*)c->E2;
  const double Lx = (d.nx - 1) * d.dx, Ly = (d.ny - 1) * d.dy;
  a.x0 = (P)d.x0; a.y0 = (P)d.y0; a.dx = (P)d.dx; a.dy = (P)d.dy;
  a.idx = (P)(1.0 / d.dx); a.idy = (P)(1.0 / d.dy);
  a.xhi = (P)(d.x0 + Lx); a.yhi = (P)(d.y0 + Ly);
  a.Lx = (P)Lx; a.Ly = (P)Ly;
  a.nx = d.nx; a.ny = d.ny;
  a.bc = d.bc; a.shape = d.shape;
  a.Q = S.Q;
  a.fix_bits = c->fix_bits;
  a.errc = c->errc;
  a.rr = c->part_map;
  a.fastdiv = sizeof(P) == 8 && d.dx >= 0x1p-400 && d.dx <= 0x1p400 && d.dy >= 0x1p-400 && d.dy <= 0x1p400;
  a.cond = c->cond.table;
  a.ncond = c->cond.active();
  return a;
}

// Push constants of species.cpp:44-51 evaluated in the reference's order:
// qm*((density*QE*L)/(EPS_0*W*vel_norm)) and vel_norm/(L*W).
static void push_consts(const Ctx* c, int s, int sub_cycle, bool rewind, double* c1, double* c2, double* c3) {
  const epic_domain_desc& d = c->dom;
  const epic_species_desc& sp = c->sp[s].desc;
  const double qm = sp.charge / sp.mass;
  const double K = (d.density * QE * d.L) / (EPS_0 * d.W * d.vel_norm);
  if (rewind) {
    *c1 = 0.5 * qm * K;  // species.cpp:242-243
    *c2 = d.DT;
  } else {
    *c1 = qm * K;
    *c2 = d.DT * sub_cycle;
  }
  *c3 = (d.vel_norm) / (d.L * d.W);
}

// Persistent grid: as many workgroups as can be resident (a multiple of the
// 8 XCDs, so the XCD-contiguous chunk walk applies).
template <typename K>
static int particle_grid(Ctx* c, K kernel) {
  const void* key = (const void*)kernel;
  auto it = c->occupancy.find(key);
  if (it != c->occupancy.end()) return c->num_cu * it->second;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, PART_BLOCK, 0) != hipSuccess || per_cu < 1)
    per_cu = 2;
  c->occupancy[key] = per_cu;
  return c->num_cu * per_cu;
}

// Persistent grid of the FP32 push: push_grid_per_cu workgroups per CU when
// set, else as many as are co-resident.
template <typename K>
static int push_grid(Ctx* c, K kernel) {
  if (c->push_grid_per_cu > 0) return c->num_cu * c->push_grid_per_cu;
  return particle_grid(c, kernel);
}

template <typename P, int PPT>
static hipError_t launch_particles_t(Ctx* c, int s, int mode, bool dep, double c1, double c2, double c3) {
  PartArgs<P> a = make_args<P>(c, s);
  a.c1 = (P)c1; a.c2 = (P)c2; a.c3 = (P)c3;
  const bool whole = (size_t)c->dom.nx * c->dom.ny <= (size_t)DEP_WINDOW;
  const dim3 b(PART_BLOCK);
  hipStream_t st = c->stream;
#define EPIC_LAUNCH(M, W, D)                                                    \
  do {                                                                         \
    auto k = particle_kernel<P, PPT, M, W, D>;                                 \
    hipLaunchKernelGGL(k, dim3(particle_grid(c, k)), b, 0, st, a);             \
  } while (0)
  if (mode == MODE_PUSH) {
    if (dep) {
      if (whole) EPIC_LAUNCH(MODE_PUSH, true, true);
      else EPIC_LAUNCH(MODE_PUSH, false, true);
    } else {
      EPIC_LAUNCH(MODE_PUSH, false, false);
    }
  } else if (mode == MODE_REWIND) {
    EPIC_LAUNCH(MODE_REWIND, false, false);
  } else {
    if (whole) EPIC_LAUNCH(MODE_DEPOSIT, true, true);
    else EPIC_LAUNCH(MODE_DEPOSIT, false, true);
  }
#undef EPIC_LAUNCH
  return hipGetLastError();
}

static hipError_t launch_particles(Ctx* c, int s, int mode, bool dep, double c1, double c2, double c3) {
  if (c->dom.precision == EPIC_FP64) {
    // particles per thread: EPIC_PART_PPT at context creation (2, 4 or 8)
    if (c->part_ppt == 4) return launch_particles_t<double, 4>(c, s, mode, dep, c1, c2, c3);
    if (c->part_ppt == 8) return launch_particles_t<double, 8>(c, s, mode, dep, c1, c2, c3);
    return launch_particles_t<double, 2>(c, s, mode, dep, c1, c2, c3);
  }
  return launch_particles_t<float, 8>(c, s, mode, dep, c1, c2, c3);
}

// Arguments of the FP32 push of species s: kv = c1 c2 multiplies E into a
// velocity increment, kx = c3 c2 a velocity into a displacement.
PushF32 make_push_f32(Ctx* c, int s, double kv, double kx, bool out_of_place) {
  const epic_domain_desc& d = c->dom;
  SpeciesDev& S = c->sp[s];
  const double Lx = (d.nx - 1) * d.dx, Ly = (d.ny - 1) * d.dy;
  PushF32 a{};
  a.x = (const float*)S.x;
  a.xo = (float*)(out_of_place ? S.x2 : S.x);
  a.d_n = S.d_n;
  a.Ec = (const float4*)c->Ec;
  a.x0 = (float)d.x0; a.y0 = (float)d.y0;
  a.idx = (float)(1.0 / d.dx); a.idy = (float)(1.0 / d.dy);
  a.xhi = (float)(d.x0 + Lx); a.yhi = (float)(d.y0 + Ly);
  a.Lx = (float)Lx; a.Ly = (float)Ly;
  a.kv = (float)kv;
  a.kx = (float)kx;
  a.nx = d.nx; a.ny = d.ny; a.bc = d.bc;
  a.Q = S.Q;
  a.errc = c->errc;
  return a;
}

// Push constants of species s for one step (sub_cycle 1), as used by the
// FP32 push: kv and kx of make_push_f32.
void push_f32_consts(Ctx* c, int s, double* kv, double* kx) {
  double c1, c2, c3;
  push_consts(c, s, 1, false, &c1, &c2, &c3);
  *kv = c1 * c2;
  *kx = c3 * c2;
}

hipError_t launch_push(Ctx* c, int s, int sub_cycle, bool deposit, double rewind_factor) {
  double c1, c2, c3;
  if (rewind_factor != 0.0) {
    push_consts(c, s, 1, true, &c1, &c2, &c3);
    return launch_particles(c, s, MODE_REWIND, false, c1, c2, c3);
  }
  push_consts(c, s, sub_cycle, false, &c1, &c2, &c3);
  const epic_domain_desc& d = c->dom;
  const bool whole = (size_t)d.nx * d.ny <= (size_t)DEP_WINDOW;
  const bool cond = c->cond.any();
  if (d.precision == EPIC_FP32 && d.shape == EPIC_SHAPE_CIC && !whole && d.nx <= 65535 && d.ny <= 65535) {
    const PushF32 a = make_push_f32(c, s, c1 * c2, c3 * c2, true);
    const CondDev* ct = c->cond.table;
    const int nct = c->cond.active();
    // the default shape, with the wall kind fixed at compile time
#define EPIC_D(DEP, COND, BC)                                                                     \
  do {                                                                                            \
    auto k = push_f32_kernel<DEP, 4, 1, COND, BC>;                                                \
    hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, a, c->push_map, ct, nct, \
                       c->cond.cmask);                                                            \
  } while (0)
#define EPIC_DB(DEP, COND)                                                                        \
  do {                                                                                            \
    if (d.bc == EPIC_BC_OPEN) EPIC_D(DEP, COND, EPIC_BC_OPEN);                                    \
    else if (d.bc == EPIC_BC_PBC) EPIC_D(DEP, COND, EPIC_BC_PBC);                                 \
    else EPIC_D(DEP, COND, EPIC_BC_REFLECTIVE);                                                   \
  } while (0)
    if (cond || c->push_variant < 1 || c->push_variant > 4) {
      // the default shape (with conductors always: its conductor test)
      if (cond && deposit) EPIC_DB(true, true);
      else if (cond) EPIC_DB(false, true);
      else if (deposit) EPIC_DB(true, false);
      else EPIC_DB(false, false);
      swap_xyv(c->sp[s]);
      return hipGetLastError();
    }
#undef EPIC_DB
#undef EPIC_D
    // kernel shape: particles per thread x occupancy target (waves per SIMD);
    // EPIC_PUSH_VARIANT at context creation
    switch (c->push_variant) {
#define EPIC_V(ID, PPT, WV)                                                                        \
  case ID:                                                                                         \
    if (deposit) {                                                                                 \
      auto k = push_f32_kernel<true, PPT, WV>;                                                     \
      hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, a, c->push_map, ct, nct, c->cond.cmask); \
    } else {                                                                                       \
      auto k = push_f32_kernel<false, PPT, WV>;                                                    \
      hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, a, c->push_map, ct, nct, c->cond.cmask); \
    }                                                                                              \
    break;
      EPIC_V(1, 8, 5)
      EPIC_V(2, 8, 1)
      EPIC_V(3, 4, 8)
      EPIC_V(4, 16, 1)
#undef EPIC_V
    }
    swap_xyv(c->sp[s]);
    return hipGetLastError();
  }
  if (d.precision == EPIC_FP64 && d.shape == EPIC_SHAPE_CIC && !whole && c->push64) {
    PartArgs<double> a = make_args<double>(c, s);
    a.c1 = c1; a.c2 = c2; a.c3 = c3;
#define EPIC_L(PPT)                                                                              \
    if (deposit) {                                                                               \
      auto k = cond ? push_f64_kernel<PPT, true, true> : push_f64_kernel<PPT, true, false>;      \
      hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, a);          \
    } else {                                                                                     \
      auto k = cond ? push_f64_kernel<PPT, false, true> : push_f64_kernel<PPT, false, false>;    \
      hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, a);          \
    }
    if (c->part_ppt == 4) {
      EPIC_L(4)
    } else {
      EPIC_L(2)
    }
#undef EPIC_L
    return hipGetLastError();
  }
  return launch_particles(c, s, MODE_PUSH, deposit, c1, c2, c3);
}

hipError_t launch_deposit(Ctx* c, int s) { return launch_particles(c, s, MODE_DEPOSIT, true, 0, 0, 0); }

// Whether one launch can push (with deposit) every species whose
// desc.pushed is set: the FP32 hot kernel applies to all of them (default
// shape) and there are 1 to PUSH_MAXSP of them.
bool push_all_ok(const Ctx* c) {
  const epic_domain_desc& d = c->dom;
  const bool whole = (size_t)d.nx * d.ny <= (size_t)DEP_WINDOW;
  if (!c->push_multi || d.precision != EPIC_FP32 || d.shape != EPIC_SHAPE_CIC || whole || d.nx > 65535 ||
      d.ny > 65535 || c->push_variant != 0)
    return false;
  int np = 0;
  for (int s = 0; s < c->nsp; s++) np += c->sp[s].desc.pushed ? 1 : 0;
  return np >= 1 && np <= PUSH_MAXSP;
}

hipError_t launch_push_all(Ctx* c) {
  PushF32Set S{};
  for (int s = 0; s < c->nsp; s++) {
    if (!c->sp[s].desc.pushed) continue;
    double c1, c2, c3;
    push_consts(c, s, 1, false, &c1, &c2, &c3);
    S.sp[S.nsp++] = make_push_f32(c, s, c1 * c2, c3 * c2, true);
  }
  S.cond = c->cond.table;
  S.ncond = c->cond.active();
  S.cmask = c->cond.cmask;
  // WAVES = 8 caps the registers at 64 (eight wavefronts per SIMD); the
  // wall kind is fixed at compile time (the compiler's choice of registers,
  // WAVES = 1, keeps the per-chunk wall dispatch)
  const bool cd = c->cond.any();
  const int bc = c->dom.bc;
  decltype(&push_f32_multi_kernel<4, false, 8>) k;
  if (c->push_multi_waves != 8)
    k = cd ? push_f32_multi_kernel<4, true, 1> : push_f32_multi_kernel<4, false, 1>;
  else if (bc == EPIC_BC_OPEN)
    k = cd ? push_f32_multi_kernel<4, true, 8, EPIC_BC_OPEN> : push_f32_multi_kernel<4, false, 8, EPIC_BC_OPEN>;
  else if (bc == EPIC_BC_PBC)
    k = cd ? push_f32_multi_kernel<4, true, 8, EPIC_BC_PBC> : push_f32_multi_kernel<4, false, 8, EPIC_BC_PBC>;
  else
    k = cd ? push_f32_multi_kernel<4, true, 8, EPIC_BC_REFLECTIVE>
           : push_f32_multi_kernel<4, false, 8, EPIC_BC_REFLECTIVE>;
  hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, S);
  for (int s = 0; s < c->nsp; s++)
    if (c->sp[s].desc.pushed) swap_xyv(c->sp[s]);
  return hipGetLastError();
}

// The pushed coordinates now live in the second buffer (vz and id stay).
void swap_xyv(SpeciesDev& S) {
  std::swap(S.x, S.x2);
  std::swap(S.y, S.y2);
  std::swap(S.vx, S.vx2);
  std::swap(S.vy, S.vy2);
}

}  // namespace epic
