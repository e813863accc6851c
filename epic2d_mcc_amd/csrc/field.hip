// Poisson solvers and E = -grad(phi) for gfx950.
//
// pcg_kernel: Jacobi-preconditioned CG with the reference recurrence of
// pcg() (solvers.h:72-100): r = b - A x, w = M^-1 r (M^-1 = -1/4 I, the
// diagonalPreconditioner of the 5-point A), v = w, alpha = w.r; per
// iteration u = A v, t = alpha/(v.u), x += v t, r -= u t, w = M^-1 r,
// beta = w.r, v = w + v (beta/alpha).  The stopping test ||r||_2 < tol is
// taken before every iteration, iterations capped at min(n^2, max_iteration).
//
// One persistent kernel runs the whole solve (a single workgroup on small
// grids; one co-resident workgroup per CU with a counter barrier otherwise).  Each thread owns
// a strided set of nodes and keeps x, r, w and v for them in registers; only
// the search direction crosses threads, through a double-buffered global
// array.  The next iteration's stencil rebuilds neighbour v from neighbour
// w and the previous v (bit-identical to what the owner computes), so one
// iteration needs two grid barriers, both fused with a dot-product
// reduction.  Dots: wavefront butterfly -> per-block partial -> every block
// sums the partials in the same fixed order, so all blocks hold identical
// scalars and runs are reproducible.  Dots accumulate in FP64 whatever the
// vector precision T.
#include <algorithm>
#include <cmath>

#include "device_common.hpp"
#include "epic_internal.hpp"
#include "node_ops.hpp"
#include "push_f32.hpp"

namespace epic {

constexpr int SOLVER_BLOCK = 1024;

template <typename T>
struct PcgArgs {
  int nx, ny;           // node grid (data layout i*ny + j)
  int op;               // EPIC_OP_REF_GHOST | EPIC_OP_DIRICHLET_WALL
  const double* b;      // right-hand side, full grid, FP64
  T* x_out;             // solution, full grid
  T* vb;                // 2*n: double-buffered search direction
  T* wg;                // n: current preconditioned residual
  double* part;         // 4 * gridDim partial sums
  double* scalars;      // out: [0] ||r||, [1] error, [2] alpha
  int* iters_out;
  long long cap;        // min(n^2, max_iteration)
  double tol;           // stop when ||r|| < tol (absolute)
  int relative;         // tol relative to ||b||
  int compute_error;    // ||A x - b|| / n (field.cpp:258-259)
  unsigned* bar;        // grid barrier counter
  unsigned long long* noconv;  // +1 when the cap ends the loop above tol (nullptr: not counted)
};

template <typename T>
__device__ __forceinline__ T rd(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Grid barrier for a co-resident (cooperative) launch: a monotonic counter
// in device memory, zeroed by the host before the launch.  Thread 0 of each
// workgroup arrives with an agent-scope release and waits with agent-scope
// acquires; the workgroup barriers order the other threads.
struct GridBar {
  unsigned* ctr;
  unsigned epoch;
  __device__ __forceinline__ void sync() {
    __syncthreads();
    if (threadIdx.x == 0) {
      epoch += gridDim.x;
      __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();
  }
};

// Reduce NV doubles over the whole grid.  `slot` selects a partial-sum region.
template <int NV, bool COOP>
__device__ __forceinline__ void grid_sum(double (&v)[NV], double* part, int slot, double* sh, GridBar& grid) {
  block_sum<NV>(v, sh);
  if (!COOP) return;
  const int G = gridDim.x;
  double* p = part + (size_t)slot * 4 * G;  // 4 values per slot, fixed offsets
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) p[k * G + blockIdx.x] = v[k];
  }
  grid.sync();
  if (threadIdx.x < 64) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0;
      for (int b = threadIdx.x; b < G; b += 64) s += __builtin_nontemporal_load(&p[k * G + b]);
      s = wave_sum(s);
      if (threadIdx.x == 0) sh[k] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = sh[k];
  __syncthreads();
}

template <bool COOP>
__device__ __forceinline__ void gsync(GridBar& grid) {
  if (COOP) {
    grid.sync();
  } else {
    __syncthreads();
  }
}

// LDSV (single workgroup): the search direction lives in LDS (n values of
// T, dynamic shared memory) instead of the global double buffer, so the
// stencil reads its neighbours from LDS.
template <typename T, int OP, int NPT, bool COOP, bool LDSV, int BS = SOLVER_BLOCK>
__global__ __launch_bounds__(BS) void pcg_kernel(PcgArgs<T> a) {
  __shared__ double sh[64];
  extern __shared__ double vs_raw[];
  T* const vs = reinterpret_cast<T*>(vs_raw);
  GridBar grid{a.bar, 0u};
  const int n = a.nx * a.ny;
  const int TT = gridDim.x * blockDim.x;
  const int g0 = blockIdx.x * blockDim.x + threadIdx.x;
  const Stencil<OP> S{a.nx, a.ny};

  T x[NPT], r[NPT], w[NPT], vold[NPT];
  bool own[NPT];
  unsigned nbm[NPT];  // the stencil's neighbour mask per owned node
  double d2[2] = {0, 0};
  double bb = 0;
#pragma unroll
  for (int m = 0; m < NPT; m++) {
    const int k = g0 + m * TT;
    own[m] = k < n && S.unknown(k);
    nbm[m] = own[m] ? S.mask(k) : 0u;
    const T bk = own[m] ? (T)a.b[k] : (T)0;
    x[m] = 0;
    // r = b - A*0 (solvers.h:75): A*0 = 0 exactly
    r[m] = bk - (T)0;
    w[m] = (T)(1.0 / -4.0) * r[m];
    vold[m] = 0;
    if (k < n) {
      a.wg[k] = w[m];
      a.vb[k] = 0;
      a.vb[n + k] = 0;
    }
    d2[0] += (double)w[m] * (double)r[m];
    d2[1] += (double)r[m] * (double)r[m];
    bb += (double)bk * (double)bk;
  }
  double d3[3] = {d2[0], d2[1], bb};
  grid_sum<3, COOP>(d3, a.part, 0, sh, grid);
  double alpha = d3[0], rr = d3[1];
  const double tol = a.relative ? a.tol * sqrt(d3[2]) : a.tol;
  double s = 0;
  int cur = 0;
  long long it = 0;
  bool conv = false;
  for (; it < a.cap; it++) {
    if (sqrt(rr) < tol) {
      conv = true;
      break;
    }
    // v = w + v_old * s ; u = A v
    T v[NPT], u[NPT];
    const T* vo = a.vb + (size_t)cur * n;
    const T sT = (T)s;
    auto vec = [&](int j) -> double { return (double)(a.wg[j] + vo[j] * sT); };
    double dvu = 0;
    if (LDSV) {
      // every thread has finished the previous stencil (the reductions
      // since then end in workgroup barriers)
#pragma unroll
      for (int m = 0; m < NPT; m++) {
        const int k = g0 + m * TT;
        v[m] = own[m] ? w[m] + vold[m] * sT : (T)0;
        if (k < n) vs[k] = v[m];
      }
      __syncthreads();
      auto vl = [&](int j) -> double { return (double)vs[j]; };
#pragma unroll
      for (int m = 0; m < NPT; m++) {
        const int k = g0 + m * TT;
        u[m] = own[m] ? (T)S.apply_mask(k, nbm[m], (double)v[m], vl) : (T)0;
        dvu += (double)v[m] * (double)u[m];
      }
    } else {
#pragma unroll
      for (int m = 0; m < NPT; m++) {
        const int k = g0 + m * TT;
        v[m] = w[m] + vold[m] * sT;
        if (own[m]) {
          u[m] = (T)S.apply_mask(k, nbm[m], (double)v[m], vec);
        } else {
          u[m] = 0;
          v[m] = 0;
        }
        dvu += (double)v[m] * (double)u[m];
      }
#pragma unroll
      for (int m = 0; m < NPT; m++) {
        const int k = g0 + m * TT;
        if (k < n) a.vb[(size_t)(cur ^ 1) * n + k] = v[m];
      }
    }
    double d1[1] = {dvu};
    if (LDSV)
      block_sum_open<1, BS / 64>(d1, sh);  // region 0..15; the next sum uses 32..63
    else
      grid_sum<1, COOP>(d1, a.part, 1, sh, grid);
    const double t = alpha / d1[0];
    const T tT = (T)t;
    double db[2] = {0, 0};
#pragma unroll
    for (int m = 0; m < NPT; m++) {
      const int k = g0 + m * TT;
      x[m] = x[m] + v[m] * tT;
      r[m] = r[m] - u[m] * tT;
      w[m] = (T)(1.0 / -4.0) * r[m];
      if (!LDSV && k < n) a.wg[k] = w[m];
      db[0] += (double)w[m] * (double)r[m];
      db[1] += (double)r[m] * (double)r[m];
      vold[m] = v[m];
    }
    if (LDSV)
      block_sum_open<2, BS / 64>(db, sh + 32);
    else
      grid_sum<2, COOP>(db, a.part, 2, sh, grid);
    const double beta = db[0];
    s = beta / alpha;
    alpha = beta;
    rr = db[1];
    cur ^= 1;
    if (!COOP && !LDSV) __syncthreads();
  }
  // write x
#pragma unroll
  for (int m = 0; m < NPT; m++) {
    const int k = g0 + m * TT;
    if (k < n) a.x_out[k] = x[m];
  }
  if (a.compute_error) {
    gsync<COOP>(grid);
    double e = 0;
    auto xv = [&](int j) -> double { return (double)a.x_out[j]; };
#pragma unroll
    for (int m = 0; m < NPT; m++) {
      const int k = g0 + m * TT;
      if (k < n && S.unknown(k)) {
        const double ax = S.apply(k, (double)x[m], xv);
        const double d = ax - a.b[k];
        e += d * d;
      } else if (k < n && OP == EPIC_OP_REF_GHOST) {
        const double d = -a.b[k];
        e += d * d;
      }
    }
    double de[1] = {e};
    grid_sum<1, COOP>(de, a.part, 3, sh, grid);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.scalars[1] = sqrt(de[0]) / n;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.scalars[0] = sqrt(rr);
    a.scalars[2] = alpha;
    *a.iters_out = (int)it;
    // the reference stops silently at the cap (solvers.h:81-97); the final
    // residual is still tested, as the next iteration would test it
    if (!conv && !(sqrt(rr) < tol) && a.noconv) atomicAdd(a.noconv, 1ull);
  }
}

// ---------------------------------------------------------------------------
// Red-black SOR on gssolver's problem (field.cpp:70-139): phi = 0 start,
// boundary phi = 0, interior update
//   phi_new = (1/coeff)(rho s + (phi_{i+1}+phi_{i-1})/dx^2 + (phi_{j+1}+phi_{j-1})/dy^2)
//   phi += omega (phi_new - phi)
// with the RMS residual sqrt(sum R^2/(nx ny)) checked every 25 iterations.
// Lexicographic and red-black SOR converge to the same discrete solution;
// their iterates differ.  BOLTZ: gsfluidsolver's problem (field.cpp:7-68),
// rho s replaced by (rho - exp(phi)) s in the update and the residual, and
// the residual checked every 50 iterations.
struct SorArgs {
  int nx, ny;
  const double* rho;
  double* phi;
  double dx2, dy2, coeff_inv, s, omega, tol;
  int maxit;
  double* part;
  double* scalars;
  int* iters_out;
  unsigned* bar;
  unsigned long long* noconv;  // +1 when maxit ends the loop above tol
};

template <bool COOP, bool BOLTZ>
__global__ __launch_bounds__(SOLVER_BLOCK) void sor_kernel(SorArgs a) {
  constexpr int CHECK = BOLTZ ? 50 : 25;
  __shared__ double sh[64];
  GridBar grid{a.bar, 0u};
  const int n = a.nx * a.ny;
  const int TT = gridDim.x * blockDim.x;
  const int g0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = g0; k < n; k += TT) a.phi[k] = 0.0;
  gsync<COOP>(grid);
  int it = 0;
  double L2 = 0;
  bool conv = false;
  for (; it < a.maxit; it++) {
    for (int color = 0; color < 2; color++) {
      for (int k = g0; k < n; k += TT) {
        const int i = k / a.ny, j = k % a.ny;
        if (i < 1 || i > a.nx - 2 || j < 1 || j > a.ny - 2 || ((i + j) & 1) != color) continue;
        const double* p = a.phi;
        const double src = BOLTZ ? (a.rho[k] - exp(p[k])) * a.s : a.rho[k] * a.s;
        const double phi_new = a.coeff_inv * (src + ((p[k + a.ny] + p[k - a.ny]) / a.dx2) +
                                              ((p[k + 1] + p[k - 1]) / a.dy2));
        a.phi[k] = p[k] + a.omega * (phi_new - p[k]);
      }
      gsync<COOP>(grid);
    }
    if (it % CHECK == 0) {
      double sum = 0;
      for (int k = g0; k < n; k += TT) {
        const int i = k / a.ny, j = k % a.ny;
        if (i < 1 || i > a.nx - 2 || j < 1 || j > a.ny - 2) continue;
        const double* p = a.phi;
        const double src = BOLTZ ? (a.rho[k] - exp(p[k])) * a.s : (a.rho[k]) * a.s;
        const double R = ((p[k + a.ny] - 2 * p[k] + p[k - a.ny]) / a.dx2) +
                         ((p[k + 1] - 2 * p[k] + p[k - 1]) / a.dy2) + src;
        sum += R * R;
      }
      double d[1] = {sum};
      grid_sum<1, COOP>(d, a.part, (it / CHECK) & 1, sh, grid);
      L2 = sqrt(d[0] / (a.nx * a.ny));
      if (L2 < a.tol) {
        conv = true;
        it++;
        break;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.scalars[1] = L2;
    *a.iters_out = it;
    // gssolver reports this on stderr (field.cpp:137)
    if (!conv && a.noconv) atomicAdd(a.noconv, 1ull);
  }
}

// ---------------------------------------------------------------------------
// RHS of pcgsolver (field.cpp:234-251): b = -rho*dx^2*((L*L)/(LDe*LDe)),
// zero on every boundary node.
__global__ void rhs_kernel(const double* rho, double* b, int nx, int ny, double dx2, double s) {
  const int n = nx * ny;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    b[k] = rhs_at(rho[k], k / ny, k % ny, nx, ny, dx2, s);
}

// r = b - A phi (FP64 true residual) and its squared norm partials.
template <int OP>
__global__ void residual_kernel(const double* b, const double* phi, double* r, int nx, int ny, double* part) {
  __shared__ double sh[64];
  const Stencil<OP> S{nx, ny};
  const int n = nx * ny;
  double acc = 0;
  auto pv = [&](int j) -> double { return phi[j]; };
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double rk = 0;
    if (S.unknown(k)) rk = b[k] - S.apply(k, phi[k], pv);
    r[k] = rk;
    acc += rk * rk;
  }
  double v[1] = {acc};
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void sum_partials_kernel(const double* part, int n, double* out) {
  double s = 0;
  for (int b = threadIdx.x; b < n; b += 64) s += part[b];
  s = wave_sum(s);
  if (threadIdx.x == 0) *out = s;
}

template <typename T>
__global__ void axpy_phi_kernel(double* phi, const T* c, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    phi[k] += (double)c[k];
}

__global__ void fill_kernel(double* p, double v, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = v;
}

// ---------------------------------------------------------------------------
// FieldSolve::CalculateEfield (field.cpp:276-331).  Also writes the
// interleaved (ex, ey) node array the particle gather reads, in particle
// precision.
template <typename P>
__global__ void efield_kernel(const double* phi, double* efx, double* efy, P* E2, int nx, int ny, double dx,
                              double dy, int bc) {
  const int n = nx * ny;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double ex, ey;
    efield_at(phi, k / ny, k % ny, nx, ny, dx, dy, bc, &ex, &ey);
    efx[k] = ex;
    efy[k] = ey;
    E2[2 * (size_t)k] = (P)ex;
    E2[2 * (size_t)k + 1] = (P)ey;
  }
}

// FP32 gather table: bilinear coefficients of every cell from E2.
__global__ void cell_table_kernel(const float2* E2, float4* Ec, int nx, int ny) {
  const int nc = (nx - 1) * (ny - 1);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
    cell_coeffs(E2, c / (ny - 1), c % (ny - 1), ny, Ec + 2 * (size_t)c);
}

static hipError_t launch_cell_table(Ctx* c) {
  const epic_domain_desc& d = c->dom;
  if (d.precision != EPIC_FP32 || !c->Ec) return hipSuccess;
  const int nc = (d.nx - 1) * (d.ny - 1);
  const int grid = std::min((nc + 255) / 256, 4096);
  hipLaunchKernelGGL(cell_table_kernel, dim3(grid), dim3(256), 0, c->stream, (const float2*)c->E2, (float4*)c->Ec,
                     d.nx, d.ny);
  return hipGetLastError();
}

static __global__ void norms_kernel(NormsArgs a) { norms_wave(a); }

// FP32 contexts: the E-field of efield_kernel and the gather table of
// cell_table_kernel in one launch.  The thread of node (i, j) writes the
// node's field and, for i < nx - 1 and j < ny - 1, the table of cell (i, j),
// evaluating the field at the cell's three other corners itself (the same
// arithmetic, so the same values as through E2).  Optionally one more
// workgroup forms the convergence norms of a mixed solve (NormsArgs.n > 0),
// which do not depend on the field.
__global__ void efield_cells_kernel(const double* phi, double* efx, double* efy, float2* E2, float4* Ec, int nx,
                                    int ny, double dx, double dy, int bc, NormsArgs na) {
  if (na.n > 0 && blockIdx.x == gridDim.x - 1) {
    if (threadIdx.x < 64) norms_wave(na);
    return;
  }
  const int n = nx * ny;
  const int nthr = (gridDim.x - (na.n > 0 ? 1 : 0)) * blockDim.x;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += nthr) {
    const int i = k / ny, j = k % ny;
    double ex, ey;
    efield_at(phi, i, j, nx, ny, dx, dy, bc, &ex, &ey);
    efx[k] = ex;
    efy[k] = ey;
    const float2 e00 = make_float2((float)ex, (float)ey);
    E2[k] = e00;
    if (i < nx - 1 && j < ny - 1) {
      double ax, ay, bx, by, cx, cy;
      efield_at(phi, i + 1, j, nx, ny, dx, dy, bc, &ax, &ay);
      efield_at(phi, i, j + 1, nx, ny, dx, dy, bc, &bx, &by);
      efield_at(phi, i + 1, j + 1, nx, ny, dx, dy, bc, &cx, &cy);
      float4* out = Ec + 2 * (size_t)(i * (ny - 1) + j);
      out[0] = bilinear_coeffs(e00.x, (float)ax, (float)bx, (float)cx);
      out[1] = bilinear_coeffs(e00.y, (float)ay, (float)by, (float)cy);
    }
  }
}

hipError_t launch_efield_norms(Ctx* c, const NormsArgs& na) {
  const epic_domain_desc& d = c->dom;
  const int n = d.nx * d.ny;
  const int grid = std::min((n + 255) / 256, 4096);
  if (d.precision == EPIC_FP32 && c->Ec) {
    hipLaunchKernelGGL(efield_cells_kernel, dim3(grid + (na.n > 0 ? 1 : 0)), dim3(256), 0, c->stream, c->phi,
                       c->efx, c->efy, (float2*)c->E2, (float4*)c->Ec, d.nx, d.ny, d.dx, d.dy, d.bc, na);
    return hipGetLastError();
  }
  if (na.n > 0) {
    hipLaunchKernelGGL(norms_kernel, dim3(1), dim3(64), 0, c->stream, na);
    EPIC_HIP_TRY(hipGetLastError());
  }
  if (d.precision == EPIC_FP64)
    hipLaunchKernelGGL(efield_kernel<double>, dim3(grid), dim3(256), 0, c->stream, c->phi, c->efx, c->efy,
                       (double*)c->E2, d.nx, d.ny, d.dx, d.dy, d.bc);
  else
    hipLaunchKernelGGL(efield_kernel<float>, dim3(grid), dim3(256), 0, c->stream, c->phi, c->efx, c->efy,
                       (float*)c->E2, d.nx, d.ny, d.dx, d.dy, d.bc);
  EPIC_HIP_TRY(hipGetLastError());
  return launch_cell_table(c);
}

hipError_t launch_efield(Ctx* c) { return launch_efield_norms(c, NormsArgs{}); }

template <typename P>
__global__ void pack_E2_kernel(const double* efx, const double* efy, P* E2, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    E2[2 * (size_t)k] = (P)efx[k];
    E2[2 * (size_t)k + 1] = (P)efy[k];
  }
}

hipError_t launch_pack_E2(Ctx* c) {
  const int n = c->dom.nx * c->dom.ny;
  const int grid = std::min((n + 255) / 256, 4096);
  if (c->dom.precision == EPIC_FP64)
    hipLaunchKernelGGL(pack_E2_kernel<double>, dim3(grid), dim3(256), 0, c->stream, c->efx, c->efy, (double*)c->E2, n);
  else
    hipLaunchKernelGGL(pack_E2_kernel<float>, dim3(grid), dim3(256), 0, c->stream, c->efx, c->efy, (float*)c->E2, n);
  EPIC_HIP_TRY(hipGetLastError());
  return launch_cell_table(c);
}

// ---------------------------------------------------------------------------
// launch helpers
static int solver_blocks(Ctx* c, int n, int* npt) {
  // one workgroup while 8 nodes per thread suffice (exactly ceil(n/1024) per
  // thread); otherwise one workgroup per CU (co-resident for the cooperative
  // launch), 1, 2, 4 or 8 nodes per thread
  if (n <= SOLVER_BLOCK * 8) {
    *npt = std::max(1, (n + SOLVER_BLOCK - 1) / SOLVER_BLOCK);
    return 1;
  }
  int G = std::min(c->max_solver_blocks, (n + SOLVER_BLOCK - 1) / SOLVER_BLOCK);
  const int p = (n + G * SOLVER_BLOCK - 1) / (G * SOLVER_BLOCK);
  *npt = p <= 1 ? 1 : p <= 2 ? 2 : p <= 4 ? 4 : 8;
  return G;
}

__global__ void zero_bar_kernel(unsigned* b) { *b = 0u; }

// LDS for the single-workgroup solve's search direction: up to 60 KiB
constexpr size_t PCG_LDS_MAX = 60 * 1024;

template <typename T, int OP, int NPT>
static hipError_t run_pcg_t(Ctx* c, int G, PcgArgs<T>& a) {
  if (G == 1) {
    hipLaunchKernelGGL((pcg_kernel<T, OP, NPT, false, false>), dim3(1), dim3(SOLVER_BLOCK), 0, c->stream, a);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(zero_bar_kernel, dim3(1), dim3(1), 0, c->stream, a.bar);
  void* args[] = {&a};
  return hipLaunchCooperativeKernel((const void*)pcg_kernel<T, OP, NPT, true, false>, dim3(G), dim3(SOLVER_BLOCK),
                                    args, 0, c->stream);
}

// Single workgroup, NPT nodes per thread: the search direction in LDS when
// it fits, else the global double buffer (NPT rounded up to 4 or 8).
template <typename T, int OP, int NPT>
static hipError_t run_pcg_single(Ctx* c, PcgArgs<T>& a) {
  const size_t lds = (size_t)a.nx * a.ny * sizeof(T);
  if (lds <= PCG_LDS_MAX) {
    hipLaunchKernelGGL((pcg_kernel<T, OP, NPT, false, true>), dim3(1), dim3(SOLVER_BLOCK), lds, c->stream, a);
    return hipGetLastError();
  }
  return run_pcg_t<T, OP, (NPT <= 4 ? 4 : 8)>(c, 1, a);
}

// Single workgroup of BS < 1024 threads (EPIC_PCG_BLOCK), search direction
// in LDS: NPT >= n / BS nodes per thread.
template <typename T, int OP, int BS, int NPT>
static hipError_t run_pcg_small(Ctx* c, PcgArgs<T>& a) {
  const size_t lds = (size_t)a.nx * a.ny * sizeof(T);
  hipLaunchKernelGGL((pcg_kernel<T, OP, NPT, false, true, BS>), dim3(1), dim3(BS), lds, c->stream, a);
  return hipGetLastError();
}

template <typename T, int OP>
static hipError_t run_pcg_op(Ctx* c, int G, int npt, PcgArgs<T>& a) {
  const int n = a.nx * a.ny;
  // 512 threads (measured 2.5 % faster per iteration at 65^2) once each of
  // 1024 threads would own more than 4 nodes
  if (G == 1 && (size_t)n * sizeof(T) <= PCG_LDS_MAX &&
      (c->pcg_block == 256 || (c->pcg_block == 512 && n > 4096))) {
    hipLaunchKernelGGL(zero_bar_kernel, dim3(1), dim3(1), 0, c->stream, a.bar);
    const int p = (n + c->pcg_block - 1) / c->pcg_block;
    if (c->pcg_block == 512) {
      if (p <= 9) return run_pcg_small<T, OP, 512, 9>(c, a);
      if (p <= 12) return run_pcg_small<T, OP, 512, 12>(c, a);
      return run_pcg_small<T, OP, 512, 16>(c, a);
    }
    if (p <= 17) return run_pcg_small<T, OP, 256, 17>(c, a);
    if (p <= 24) return run_pcg_small<T, OP, 256, 24>(c, a);
    return run_pcg_small<T, OP, 256, 32>(c, a);
  }
  if (G == 1) {
    hipLaunchKernelGGL(zero_bar_kernel, dim3(1), dim3(1), 0, c->stream, a.bar);
    switch (npt) {
      case 1: return run_pcg_single<T, OP, 1>(c, a);
      case 2: return run_pcg_single<T, OP, 2>(c, a);
      case 3: return run_pcg_single<T, OP, 3>(c, a);
      case 4: return run_pcg_single<T, OP, 4>(c, a);
      case 5: return run_pcg_single<T, OP, 5>(c, a);
      case 6: return run_pcg_single<T, OP, 6>(c, a);
      case 7: return run_pcg_single<T, OP, 7>(c, a);
      default: return run_pcg_single<T, OP, 8>(c, a);
    }
  }
  switch (npt) {
    case 1: return run_pcg_t<T, OP, 1>(c, G, a);
    case 2: return run_pcg_t<T, OP, 2>(c, G, a);
    case 4: return run_pcg_t<T, OP, 4>(c, G, a);
    default: return run_pcg_t<T, OP, 8>(c, G, a);
  }
}

// count_noconv: a cap reached above tolerance adds to the context's
// EPIC_ERRC_NOCONV counter (the outer solves; not the mixed scheme's inner
// FP32 solves, whose tolerance is only a stopping rule).
template <typename T>
static hipError_t run_pcg(Ctx* c, const double* b, T* x_out, long long cap, double tol, int relative,
                          int compute_error, bool count_noconv = false) {
  const epic_domain_desc& d = c->dom;
  const int n = d.nx * d.ny;
  int npt = 1;
  const int G = solver_blocks(c, n, &npt);
  if ((long long)G * SOLVER_BLOCK * npt < n) return hipErrorInvalidValue;
  PcgArgs<T> a{};
  a.nx = d.nx; a.ny = d.ny; a.op = d.poisson_op;
  a.b = b; a.x_out = x_out;
  a.vb = (T*)c->ws_v; a.wg = (T*)c->ws_w;
  a.part = c->ws_part; a.scalars = c->ws_scalar; a.iters_out = c->ws_iter;
  a.cap = cap; a.tol = tol; a.relative = relative; a.compute_error = compute_error;
  a.bar = (unsigned*)(c->ws_iter + 1);
  a.noconv = count_noconv ? c->errc + EPIC_ERRC_NOCONV : nullptr;
  if (d.poisson_op == EPIC_OP_REF_GHOST) return run_pcg_op<T, EPIC_OP_REF_GHOST>(c, G, npt, a);
  return run_pcg_op<T, EPIC_OP_DIRICHLET_WALL>(c, G, npt, a);
}

static hipError_t read_scalars(Ctx* c, int* iters, double* scal3) {
  EPIC_HIP_TRY(hipMemcpyAsync(scal3, c->ws_scalar, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  EPIC_HIP_TRY(hipMemcpyAsync(iters, c->ws_iter, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  return hipStreamSynchronize(c->stream);
}

hipError_t launch_spectral_solve(Ctx* c, double* b, double* phi, bool fp32_inner, double tol, int max_ref,
                                 int* solves, double* relres, const double* rho, double dx2, double s);

// rho := voltage on the conductors' nodes (the grids' initialize() at the
// start of gssolver, field.cpp:100-103); the host resolved repeated nodes to
// the last write, so the entries are distinct.
__global__ void conductor_rho_kernel(double* rho, const int* node, const double* value, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) rho[node[k]] = value[k];
}

hipError_t launch_conductor_rho(Ctx* c) {
  Conductors& q = c->cond;
  if (q.nnode > 0) {
    hipLaunchKernelGGL(conductor_rho_kernel, dim3(std::min((q.nnode + 255) / 256, 1024)), dim3(256), 0, c->stream,
                       c->rho, q.node, q.value, q.nnode);
    EPIC_HIP_TRY(hipGetLastError());
  }
  // CircularGrid::initialize (grid.cpp:40-50) creates the electrodes that
  // IsGrid tests from now on
  q.elec_live = true;
  return hipSuccess;
}

// nrpcgsolver (field.cpp:141-210) in one workgroup: ten Newton steps, each
// cg() (solvers.h:32-69) from zero on J = A - diag(exp(x)) with F = A x -
// rhovec, ||r|| < 1e-4, at most min(n, 10000) iterations, then x -= y.  As
// written in the reference, rhovec accumulates exp(x) L^2/LDe^2 over the
// steps while J's diagonal takes exp(x) alone.  A is pcgsolver's matrix
// (flat index k = row*nx + col, neighbours k +- 1, k +- nx, summed in column
// order); dot products are workgroup sums.
struct NrArgs {
  int nx, ny;
  const double* rho;
  double* x;  // phi
  double dx2, s;
  double* ws;  // rhovec, ex, y, r, p, Ap: 6n
  int* iters_out;
  double* scalars;
};

__device__ __forceinline__ double nr_row(const double* v, const double* ex, int k, int nx, int ny, bool jac) {
  const int row = k / nx, col = k % nx;
  double sum = 0;
  if (row > 0) sum += 1.0 * v[k - nx];
  if (col > 0) sum += 1.0 * v[k - 1];
  sum += (jac ? (-4.0 - ex[k]) : -4.0) * v[k];
  if (col < nx - 1) sum += 1.0 * v[k + 1];
  if (row < ny - 1) sum += 1.0 * v[k + nx];
  return sum;
}

__global__ __launch_bounds__(SOLVER_BLOCK) void newton_cg_kernel(NrArgs a) {
  __shared__ double sh[SOLVER_BLOCK / 64];
  const int nx = a.nx, ny = a.ny, n = nx * ny, t = threadIdx.x, T = blockDim.x;
  double* rv = a.ws;
  double* ex = a.ws + n;
  double* y = a.ws + 2 * (size_t)n;
  double* r = a.ws + 3 * (size_t)n;
  double* p = a.ws + 4 * (size_t)n;
  double* Ap = a.ws + 5 * (size_t)n;
  auto dot = [&](const double* u, const double* v) {
    double d[1] = {0.0};
    for (int k = t; k < n; k += T) d[0] += u[k] * v[k];
    block_sum<1>(d, sh);
    return d[0];
  };
  for (int k = t; k < n; k += T) {
    rv[k] = rhs_at(a.rho[k], k / ny, k % ny, nx, ny, a.dx2, a.s);
    a.x[k] = 0.0;
  }
  __syncthreads();
  int total = 0;
  const int cap = min(n, 10000);
  for (int newton = 0; newton < 10; newton++) {
    for (int k = t; k < n; k += T) {
      const double e = exp(a.x[k]);
      rv[k] = rv[k] + e * a.s;
      ex[k] = e;
    }
    __syncthreads();
    // F = A x - rhovec; r0 = F - J 0 = F; p0 = r0; y = 0
    for (int k = t; k < n; k += T) {
      const double F = nr_row(a.x, ex, k, nx, ny, false) - rv[k];
      r[k] = F;
      p[k] = F;
      y[k] = 0.0;
    }
    __syncthreads();
    double rr = dot(r, r);
    for (int it = 0; it < cap; it++) {
      if (sqrt(rr) < 1e-4) break;
      for (int k = t; k < n; k += T) Ap[k] = nr_row(p, ex, k, nx, ny, true);
      __syncthreads();
      const double pAp = dot(p, Ap);
      const double alpha = rr / pAp;
      for (int k = t; k < n; k += T) {
        y[k] = y[k] + p[k] * alpha;
        r[k] = r[k] - Ap[k] * alpha;
      }
      __syncthreads();
      const double r12 = dot(r, r);
      const double beta = r12 / rr;
      for (int k = t; k < n; k += T) p[k] = r[k] + p[k] * beta;
      __syncthreads();
      rr = r12;
      total++;
    }
    for (int k = t; k < n; k += T) a.x[k] = a.x[k] - y[k];
    __syncthreads();
  }
  if (t == 0) {
    *a.iters_out = total;
    a.scalars[1] = 0.0;  // nrpcgsolver sets error = 0
  }
}

// gsfluidsolver's boundary rho (field.cpp:23-33): 0 on the bottom, left and
// right edges, top_v on the top edge (j = ny-1) except its two corners, which
// the second loop resets to 0.
__global__ void boltz_boundary_rho_kernel(double* rho, int nx, int ny, double top_v) {
  const int n = nx * ny;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = k / ny, j = k % ny;
    if (i == 0 || i == nx - 1 || j == 0) rho[k] = 0.0;
    else if (j == ny - 1) rho[k] = top_v;
  }
}

hipError_t launch_solve(Ctx* c, int* iters, double* err) {
  const epic_domain_desc& d = c->dom;
  const int n = d.nx * d.ny;
  const double dx2 = d.dx * d.dx;
  const double s = (d.L * d.L) / (d.LDe * d.LDe);
  const int grid = std::min((n + 255) / 256, 4096);
  const long long cap = std::min<long long>((long long)n * n, (long long)d.max_iteration);
  if (d.solver == EPIC_SOLVER_NEWTON_CG) {
    if (!c->nr_ws) EPIC_HIP_TRY(hipMalloc(&c->nr_ws, 6 * (size_t)n * sizeof(double)));
    NrArgs a{};
    a.nx = d.nx; a.ny = d.ny; a.rho = c->rho; a.x = c->phi;
    a.dx2 = dx2; a.s = s;
    a.ws = c->nr_ws; a.iters_out = c->ws_iter; a.scalars = c->ws_scalar;
    hipLaunchKernelGGL(newton_cg_kernel, dim3(1), dim3(SOLVER_BLOCK), 0, c->stream, a);
    EPIC_HIP_TRY(hipGetLastError());
    double sc[3];
    EPIC_HIP_TRY(read_scalars(c, iters, sc));
    *err = sc[1];
    return hipSuccess;
  }
  if (d.solver == EPIC_SOLVER_GS || d.solver == EPIC_SOLVER_GS_BOLTZMANN) {
    const bool boltz = d.solver == EPIC_SOLVER_GS_BOLTZMANN;
    // gssolver runs the conductors' initialize(); gsfluidsolver does not
    if (!boltz && c->cond.any()) EPIC_HIP_TRY(launch_conductor_rho(c));
    if (boltz) {
      // 10000 * (eV / (K_b * EV_to_K)) with the reference's constants (domain.h:204-215)
      const double QE_ = 1.602176565e-19, K_b = 1.380648e-23, EV_to_K = QE_ / K_b, eV_ = 1.602176565e-19;
      const double top_v = 10000 * (eV_ / (K_b * EV_to_K));
      hipLaunchKernelGGL(boltz_boundary_rho_kernel, dim3(grid), dim3(256), 0, c->stream, c->rho, d.nx, d.ny, top_v);
      EPIC_HIP_TRY(hipGetLastError());
    }
    SorArgs a{};
    a.nx = d.nx; a.ny = d.ny; a.rho = c->rho; a.phi = c->phi;
    a.dx2 = d.dx * d.dx; a.dy2 = d.dy * d.dy;
    a.coeff_inv = 1 / ((2 / a.dx2) + (2 / a.dy2));
    a.s = s; a.omega = d.sor_omega; a.tol = d.tolerance; a.maxit = d.max_iteration;
    a.part = c->ws_part; a.scalars = c->ws_scalar; a.iters_out = c->ws_iter;
    a.bar = (unsigned*)(c->ws_iter + 1);
    a.noconv = c->errc + EPIC_ERRC_NOCONV;
    int npt;
    const int G = solver_blocks(c, n, &npt);
    hipLaunchKernelGGL(zero_bar_kernel, dim3(1), dim3(1), 0, c->stream, a.bar);
    if (G == 1) {
      auto k1 = boltz ? sor_kernel<false, true> : sor_kernel<false, false>;
      hipLaunchKernelGGL(k1, dim3(1), dim3(SOLVER_BLOCK), 0, c->stream, a);
      EPIC_HIP_TRY(hipGetLastError());
    } else {
      void* args[] = {&a};
      const void* k = boltz ? (const void*)sor_kernel<true, true> : (const void*)sor_kernel<true, false>;
      EPIC_HIP_TRY(hipLaunchCooperativeKernel(k, dim3(G), dim3(SOLVER_BLOCK), args, 0, c->stream));
    }
    double sc[3];
    EPIC_HIP_TRY(read_scalars(c, iters, sc));
    *err = sc[1];
    return hipSuccess;
  }

  if (d.solver == EPIC_SOLVER_SPECTRAL || (d.solver == EPIC_SOLVER_PCG_MIXED && d.precond == EPIC_PRECOND_SPECTRAL)) {
    // b is built from rho by the solve's first kernel
    const bool fp32 = d.solver == EPIC_SOLVER_PCG_MIXED;
    const double tol = d.tol_relative ? d.tolerance : 1e-12;
    return launch_spectral_solve(c, c->b64, c->phi, fp32, tol, fp32 ? 6 : 3, iters, err, c->rho, dx2, s);
  }

  hipLaunchKernelGGL(rhs_kernel, dim3(grid), dim3(256), 0, c->stream, c->rho, c->b64, d.nx, d.ny, dx2, s);
  EPIC_HIP_TRY(hipGetLastError());

  if (d.solver == EPIC_SOLVER_PCG) {
    EPIC_HIP_TRY(run_pcg<double>(c, c->b64, c->phi, cap, d.tolerance, d.tol_relative, 1, true));
    double sc[3];
    EPIC_HIP_TRY(read_scalars(c, iters, sc));
    *err = sc[1];
    return hipSuccess;
  }

  // Mixed precision: FP64 true residual + refinement around FP32 Jacobi-PCG
  // correction solves.  Stops when ||b - A phi||_2 <= tol ||b||_2.
  const int rgrid = std::min((n + 255) / 256, 1024);
  double bnorm = -1, rnorm = 0;
  int total = 0;
  hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, c->stream, c->phi, 0.0, n);
  for (int outer = 0; outer < 8; outer++) {
    if (d.poisson_op == EPIC_OP_REF_GHOST)
      hipLaunchKernelGGL(residual_kernel<EPIC_OP_REF_GHOST>, dim3(rgrid), dim3(256), 0, c->stream, c->b64, c->phi,
                         c->r64, d.nx, d.ny, c->ws_part);
    else
      hipLaunchKernelGGL(residual_kernel<EPIC_OP_DIRICHLET_WALL>, dim3(rgrid), dim3(256), 0, c->stream, c->b64,
                         c->phi, c->r64, d.nx, d.ny, c->ws_part);
    EPIC_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, c->stream, c->ws_part, rgrid, c->ws_scalar + 3);
    double r2;
    EPIC_HIP_TRY(hipMemcpyAsync(&r2, c->ws_scalar + 3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
    rnorm = sqrt(r2);
    if (bnorm < 0) bnorm = rnorm;  // phi = 0: r = b
    if (rnorm <= d.tolerance * bnorm || rnorm == 0) break;
    // inner FP32 solve of A c = r to 1e-5 relative
    EPIC_HIP_TRY(run_pcg<float>(c, c->r64, (float*)c->ws_x, cap, 1e-5, 1, 0));
    int it;
    double sc[3];
    EPIC_HIP_TRY(read_scalars(c, &it, sc));
    total += it;
    hipLaunchKernelGGL(axpy_phi_kernel<float>, dim3(grid), dim3(256), 0, c->stream, c->phi,
                       (const float*)c->ws_x, n);
    EPIC_HIP_TRY(hipGetLastError());
  }
  *iters = total;
  *err = bnorm > 0 ? rnorm / bnorm : 0;
  if (!(rnorm <= d.tolerance * bnorm || rnorm == 0)) c->errc_host[EPIC_ERRC_NOCONV]++;
  return hipSuccess;
}

}  // namespace epic

