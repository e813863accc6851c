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
#include "device_common.hpp"
#include "epic_internal.hpp"

namespace epic {

enum { MODE_PUSH = 0, MODE_REWIND = 1, MODE_DEPOSIT = 2 };

typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));

template <typename P>
struct alignas(2 * sizeof(P)) Vec2 {
  P a, b;
};

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
  int nx, ny;
  int bc, shape;
  P c1, c2, c3;        // dv = (c1*E)*c2 ; dpos = (c3*v)*c2
  unsigned long long* Q;
  int fix_bits;
  unsigned long long* errc;
};

template <typename P, int PPT>
__device__ __forceinline__ void ld_run(const P* __restrict__ a, int64_t i0, int64_t n, P (&o)[PPT]) {
  constexpr int NV = PPT * (int)sizeof(P) / 16;
  if (i0 + PPT <= n) {
    const v4u32* p = reinterpret_cast<const v4u32*>(a + i0);
    v4u32 t[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) t[k] = __builtin_nontemporal_load(p + k);
    __builtin_memcpy(&o[0], &t[0], sizeof(t));
  } else {
#pragma unroll
    for (int k = 0; k < PPT; k++) o[k] = (i0 + k < n) ? a[i0 + k] : P(0);
  }
}

template <typename P, int PPT>
__device__ __forceinline__ void st_run(P* __restrict__ a, int64_t i0, int64_t n, const P (&o)[PPT]) {
  constexpr int NV = PPT * (int)sizeof(P) / 16;
  if (i0 + PPT <= n) {
    v4u32* p = reinterpret_cast<v4u32*>(a + i0);
    v4u32 t[NV];
    __builtin_memcpy(&t[0], &o[0], sizeof(t));
#pragma unroll
    for (int k = 0; k < NV; k++) __builtin_nontemporal_store(t[k], p + k);
  } else {
#pragma unroll
    for (int k = 0; k < PPT; k++)
      if (i0 + k < n) a[i0 + k] = o[k];
  }
}

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

__device__ __forceinline__ long long fix_round(double w, double scale) { return (long long)rint(w * scale); }
__device__ __forceinline__ long long fix_round(float w, float scale) { return (long long)rintf(w * scale); }

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
  const P lx = (x - a.x0) / a.dx;
  const P ly = (y - a.y0) / a.dy;
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

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// WHOLE: the whole node grid is one LDS histogram per workgroup, flushed once
// at the end (small grids).  Otherwise each wavefront owns an LDS window of
// WAVE_WIN node accumulators placed on the bounding box of the cells its
// 64*PPT particles deposit to; deposits outside go to global atomics.  A
// wavefront never waits on the rest of its workgroup.
constexpr int WAVE_WIN = 512;

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
  const int ny = a.ny;
  const int ncorner = a.shape == EPIC_SHAPE_NGP ? 1 : 4;
  const P scale = (P)ldexp(1.0, a.fix_bits);
  const long long one = 1ll << a.fix_bits;

  Window win{0, 0, a.nx, a.ny};
  if (DEP && WHOLE) {
    for (int k = threadIdx.x; k < a.nx * a.ny; k += blockDim.x) lds[k] = 0ull;
    __syncthreads();
  }

  int64_t cb, ce_, cs;
  xcd_chunk_range(nchunks, &cb, &ce_, &cs);
  for (int64_t chunk = cb; chunk < ce_; chunk += cs) {
    const int64_t i0 = chunk * CH + (int64_t)wid * CHW + (int64_t)lane * PPT;
    P x[PPT], y[PPT], vx[PPT], vy[PPT];
    ld_run<P, PPT>(a.x, i0, n, x);
    ld_run<P, PPT>(a.y, i0, n, y);
    if (MODE != MODE_DEPOSIT) {
      ld_run<P, PPT>(a.vx, i0, n, vx);
      ld_run<P, PPT>(a.vy, i0, n, vy);
    }
    bool live[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) live[k] = (i0 + k < n) && !isnan(x[k]);

    if (MODE != MODE_DEPOSIT) {
      int ckey = -1;
      P cev[8];
      int bad = 0;
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        if (!live[k]) continue;
        const P lx = (x[k] - a.x0) / a.dx;
        const P ly = (y[k] - a.y0) / a.dy;
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
        if (a.bc == EPIC_BC_PBC) {
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
        x[k] = px;
        y[k] = py;
      }
      if (bad) atomicAdd(&a.errc[EPIC_ERRC_GATHER_OOB], (unsigned long long)bad);
      st_run<P, PPT>(a.vx, i0, n, vx);
      st_run<P, PPT>(a.vy, i0, n, vy);
      if (MODE == MODE_PUSH) {
        st_run<P, PPT>(a.x, i0, n, x);
        st_run<P, PPT>(a.y, i0, n, y);
      }
    }

    if (!DEP) continue;

    // ---- deposit the (new) positions ------------------------------------
    int bad = 0;
    if (!WHOLE) {
      int imin = 1 << 30, imax = -1, jmin = 1 << 30, jmax = -1;
#pragma unroll
      for (int k = 0; k < PPT; k++) {
        if (!live[k]) continue;
        int i, j;
        P di, dj;
        if (!dep_cell(a, x[k], y[k], i, j, di, dj)) continue;
        imin = min(imin, i); imax = max(imax, i);
        jmin = min(jmin, j); jmax = max(jmax, j);
      }
      const int gi0 = wave_min(imin), gi1 = wave_max(imax);
      if (gi1 < 0) continue;  // nothing to deposit in this wavefront chunk (uniform)
      const int gj0 = wave_min(jmin), gj1 = wave_max(jmax);
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

    // run-length accumulation over this thread's consecutive particles
    int rk_i = -1, rk_j = -1;
    unsigned long long acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      if (!live[k]) continue;
      int i, j;
      P di, dj;
      if (!dep_cell(a, x[k], y[k], i, j, di, dj)) {
        bad++;
        continue;
      }
      if (i != rk_i || j != rk_j) {
        if (rk_i >= 0) flush_block(rk_i, rk_j, ny, acc, ncorner, win, wl, a.Q);
        rk_i = i;
        rk_j = j;
        acc[0] = acc[1] = acc[2] = acc[3] = 0;
      }
      if (ncorner == 1) {
        acc[0] += (unsigned long long)one;
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
    if (bad) atomicAdd(&a.errc[EPIC_ERRC_DEPOSIT_OOB], (unsigned long long)bad);

    if (!WHOLE) {
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
  a.xhi = (P)(d.x0 + Lx); a.yhi = (P)(d.y0 + Ly);
  a.Lx = (P)Lx; a.Ly = (P)Ly;
  a.nx = d.nx; a.ny = d.ny;
  a.bc = d.bc; a.shape = d.shape;
  a.Q = S.Q;
  a.fix_bits = c->fix_bits;
  a.errc = c->errc;
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
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, PART_BLOCK, 0) != hipSuccess || per_cu < 1)
    per_cu = 2;
  return c->num_cu * per_cu;
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
  if (c->dom.precision == EPIC_FP64) return launch_particles_t<double, 4>(c, s, mode, dep, c1, c2, c3);
  return launch_particles_t<float, 8>(c, s, mode, dep, c1, c2, c3);
}

hipError_t launch_push(Ctx* c, int s, int sub_cycle, bool deposit, double rewind_factor) {
  double c1, c2, c3;
  if (rewind_factor != 0.0) {
    push_consts(c, s, 1, true, &c1, &c2, &c3);
    return launch_particles(c, s, MODE_REWIND, false, c1, c2, c3);
  }
  push_consts(c, s, sub_cycle, false, &c1, &c2, &c3);
  return launch_particles(c, s, MODE_PUSH, deposit, c1, c2, c3);
}

hipError_t launch_deposit(Ctx* c, int s) { return launch_particles(c, s, MODE_DEPOSIT, true, 0, 0, 0); }

}  // namespace epic
