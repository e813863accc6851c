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

__device__ __forceinline__ long long fix_round(double w, double scale) { return (long long)rint(w * scale); }
__device__ __forceinline__ long long fix_round(float w, float scale) { return (long long)rintf(w * scale); }

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
  const P lx = sizeof(P) == 8 ? (x - a.x0) / a.dx : (x - a.x0) * a.idx;
  const P ly = sizeof(P) == 8 ? (y - a.y0) / a.dy : (y - a.y0) * a.idy;
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

// WHOLE: the whole node grid is one LDS histogram per workgroup, flushed once
// at the end (small grids).  Otherwise each wavefront owns an LDS window of
// WAVE_WIN node accumulators placed on the bounding box of the cells its
// 64*PPT particles deposit to; deposits outside go to global atomics.  A
// wavefront never waits on the rest of its workgroup.
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
    const int64_t wb = chunk * CH + (int64_t)wid * CHW;
    P x[PPT], y[PPT], vx[PPT], vy[PPT];
    ld_run<P, PPT>(a.x, wb, lane, n, x);
    ld_run<P, PPT>(a.y, wb, lane, n, y);
    if (MODE != MODE_DEPOSIT) {
      ld_run<P, PPT>(a.vx, wb, lane, n, vx);
      ld_run<P, PPT>(a.vy, wb, lane, n, vy);
    }
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
        }
        x[k] = px;
        y[k] = py;
      }
      if (bad) atomicAdd(&a.errc[EPIC_ERRC_GATHER_OOB], (unsigned long long)bad);
      st_run<P, PPT>(a.vx, wb, lane, n, vx);
      st_run<P, PPT>(a.vy, wb, lane, n, vy);
      if (MODE == MODE_PUSH) {
        st_run<P, PPT>(a.x, wb, lane, n, x);
        st_run<P, PPT>(a.y, wb, lane, n, y);
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

// rr = 0: each XCD walks its own contiguous range of chunks; rr = 1: all
// workgroups walk the chunks round-robin (the whole device streams through
// one region at a time).
template <bool DEP, int PPT, int WAVES>
__global__ __launch_bounds__(PART_BLOCK, WAVES) void push_f32_kernel(PushF32 a, int rr) {
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
    push_chunk_regs<DEP, PPT>(a, chunk * CH + (int64_t)wid * CHW, n, lane, r, dp);
    const int64_t nb = (chunk + cs) * CH + (int64_t)wid * CHW;
    const bool more = chunk + cs < ce_;
    const bool full = more && nb + 64 * PPT <= n;
    if (full) load_chunk_full<PPT>(a, nb, lane, r);
    if (DEP) deposit_chunk<PPT>(a, lane, wl, dp);
    if (more && !full) load_chunk<PPT>(a, nb, n, lane, r);
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
  if (c->dom.precision == EPIC_FP64) return launch_particles_t<double, 4>(c, s, mode, dep, c1, c2, c3);
  return launch_particles_t<float, 8>(c, s, mode, dep, c1, c2, c3);
}

// Arguments of the FP32 push of species s: kv = c1 c2 multiplies E into a
// velocity increment, kx = c3 c2 a velocity into a displacement.
PushF32 make_push_f32(Ctx* c, int s, double kv, double kx) {
  const epic_domain_desc& d = c->dom;
  SpeciesDev& S = c->sp[s];
  const double Lx = (d.nx - 1) * d.dx, Ly = (d.ny - 1) * d.dy;
  PushF32 a{};
  a.x = (float*)S.x; a.y = (float*)S.y; a.vx = (float*)S.vx; a.vy = (float*)S.vy;
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
  if (d.precision == EPIC_FP32 && d.shape == EPIC_SHAPE_CIC && !whole && d.nx <= 65535 && d.ny <= 65535) {
    const PushF32 a = make_push_f32(c, s, c1 * c2, c3 * c2);
    // kernel shape: particles per thread x occupancy target (waves per SIMD);
    // EPIC_PUSH_VARIANT at context creation
    switch (c->push_variant) {
#define EPIC_V(ID, PPT, WV)                                                                        \
  case ID:                                                                                         \
    if (deposit) {                                                                                 \
      auto k = push_f32_kernel<true, PPT, WV>;                                                     \
      hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, a, c->push_map); \
    } else {                                                                                       \
      auto k = push_f32_kernel<false, PPT, WV>;                                                    \
      hipLaunchKernelGGL(k, dim3(push_grid(c, k)), dim3(PART_BLOCK), 0, c->stream, a, c->push_map); \
    }                                                                                              \
    break;
      EPIC_V(1, 8, 5)
      EPIC_V(2, 8, 1)
      EPIC_V(3, 4, 8)
      EPIC_V(4, 16, 1)
      default:
      EPIC_V(0, 4, 1)
#undef EPIC_V
    }
    return hipGetLastError();
  }
  return launch_particles(c, s, MODE_PUSH, deposit, c1, c2, c3);
}

hipError_t launch_deposit(Ctx* c, int s) { return launch_particles(c, s, MODE_DEPOSIT, true, 0, 0, 0); }

}  // namespace epic
