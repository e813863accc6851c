// Exact solve of the 5-point operator by sine transforms, on the matrix cores.
//
// Both Poisson operators are the 5-point Laplacian with -4 on the diagonal
// and zero values outside a p x q block of unknowns: the Dirichlet-wall
// problem (interior nodes, gssolver's field.cpp:89-98) and pcgsolver's
// zero-ghost operator (every node, field.cpp:223-232).  Its eigenvectors are
// sine modes, S[a][k] = sin(pi (a+1)(k+1)/(p+1)), with eigenvalues
// 2cos(pi(k+1)/(p+1)) - 2 per direction, and S^-1 = 2/(p+1) S.  So
//
//   U = S_p ( (S_p X S_q) o G ) S_q,   G[k][l] = c / (lam_k + mu_l),
//   c = (2/(p+1)) (2/(q+1))
//
// is four dense matrix products on the matrix cores (gemm_tile.hpp: FP32 or
// FP64 MFMA, hand-written) with the scaling by G in the second product's
// epilogue and the write into phi in the fourth's.  In FP32 it is the inner solve of the mixed-precision
// scheme: FP64 true residual r = b - A phi, c = A^-1 r in FP32, phi += c,
// repeated until ||r|| <= tol ||b|| (FP64 dot products); in FP64 it is a
// direct solve followed by the same refinement.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "epic_internal.hpp"
#include "gemm_tile.hpp"
#include "node_ops.hpp"

namespace epic {

// Transform dimensions padded to a multiple of 32 with zero rows and
// columns, the shapes the tile products take.
static int spec_pad(int p) { return spec_pad32(p); }

static SpecLayout spec_layout(const Ctx* c) {
  const epic_domain_desc& d = c->dom;
  if (d.poisson_op == EPIC_OP_DIRICHLET_WALL)
    return SpecLayout{d.ny - 2, d.nx - 2, d.ny, d.ny + 1, spec_pad(d.ny - 2), spec_pad(d.nx - 2)};
  // pcgsolver's A: flat index row*nx + col, all nodes unknown
  return SpecLayout{d.nx, d.ny, d.nx, 0, spec_pad(d.nx), spec_pad(d.ny)};
}

template <typename T>
__global__ void pack_kernel(const double* src, T* X, SpecLayout L) {
  const int n = L.p * L.q;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int a = k % L.p, b = k / L.p;
    X[a + (size_t)b * L.pp] = (T)src[L.off + a + (size_t)b * L.ld];
  }
}

// residual r = b - A phi over the unknowns (FP64) and ||r||^2 partials;
// X (when given): r packed in FP32 as the transform's operand
// The solve's first residual can also build b from rho: pcgsolver's
// right-hand side (field.cpp:234-251), when rho is given.
struct RhsSrc {
  const double* rho;
  double* b;
  double dx2, s;
};

// FIN: rho is first finalised here from the species' accumulators (the work
// of rho_finalize_kernel, node by node, in the same launch).
template <int OP, bool FIN>
__global__ void spec_residual_kernel(const double* b, const double* phi, double* r, int nx, int ny, double* part,
                                     float* X, SpecLayout L, double* zero_phi, RhsSrc src, FinArgs fin) {
  __shared__ double sh[64];
  const int n = nx * ny;
  double acc = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    // zero_phi: the solve's start, phi = 0 (written here, not read)
    double rk;
    if (zero_phi) {
      if (src.rho) {
        const double rho = FIN ? fin_node(fin, k) : src.rho[k];
        src.b[k] = rhs_at(rho, k / ny, k % ny, nx, ny, src.dx2, src.s);
      }
      rk = residual_at_zero(OP, b, k, nx, ny);
      zero_phi[k] = 0.0;
    } else {
      rk = residual_at(OP, b, phi, k, nx, ny);
    }
    r[k] = rk;
    acc += rk * rk;
    if (X) {
      const int x = spec_index(L, k);
      if (x >= 0) X[x] = (float)rk;
    }
  }
  double v[1] = {acc};
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

static __global__ void norms_kernel(NormsArgs a) { norms_wave(a); }

struct SpecState {
  SpecLayout L{};
  // FP64 and FP32 copies of S_p, S_q, G
  double *Sp64 = nullptr, *Sq64 = nullptr, *G64 = nullptr, *X64 = nullptr, *Y64 = nullptr;
  float *Sp32 = nullptr, *Sq32 = nullptr, *G32 = nullptr, *X32 = nullptr, *Y32 = nullptr;
};

static hipError_t spec_init(Ctx* c) {
  if (c->blas) return hipSuccess;
  SpecState* S = new SpecState();
  S->L = spec_layout(c);
  const int p = S->L.p, q = S->L.q, P = S->L.pp, Q = S->L.qq;
  std::vector<double> sp((size_t)P * P, 0.0), sq((size_t)Q * Q, 0.0), g((size_t)P * Q, 0.0);
  for (int a = 0; a < p; a++)
    for (int k = 0; k < p; k++) sp[a + (size_t)k * P] = sin(M_PI * (double)(a + 1) * (k + 1) / (p + 1));
  for (int a = 0; a < q; a++)
    for (int k = 0; k < q; k++) sq[a + (size_t)k * Q] = sin(M_PI * (double)(a + 1) * (k + 1) / (q + 1));
  const double cc = (2.0 / (p + 1)) * (2.0 / (q + 1));
  for (int k = 0; k < p; k++)
    for (int l = 0; l < q; l++) {
      const double lam = 2 * cos(M_PI * (k + 1) / (p + 1)) - 2 + 2 * cos(M_PI * (l + 1) / (q + 1)) - 2;
      g[k + (size_t)l * P] = cc / lam;
    }
  std::vector<float> sp32(sp.begin(), sp.end()), sq32(sq.begin(), sq.end()), g32(g.begin(), g.end());
  EPIC_HIP_TRY(hipMalloc(&S->Sp64, sp.size() * 8));
  EPIC_HIP_TRY(hipMalloc(&S->Sq64, sq.size() * 8));
  EPIC_HIP_TRY(hipMalloc(&S->G64, g.size() * 8));
  EPIC_HIP_TRY(hipMalloc(&S->X64, g.size() * 8));
  EPIC_HIP_TRY(hipMalloc(&S->Y64, g.size() * 8));
  EPIC_HIP_TRY(hipMalloc(&S->Sp32, sp.size() * 4));
  EPIC_HIP_TRY(hipMalloc(&S->Sq32, sq.size() * 4));
  EPIC_HIP_TRY(hipMalloc(&S->G32, g.size() * 4));
  EPIC_HIP_TRY(hipMalloc(&S->X32, g.size() * 4));
  EPIC_HIP_TRY(hipMalloc(&S->Y32, g.size() * 4));
  EPIC_HIP_TRY(hipMemcpy(S->Sp64, sp.data(), sp.size() * 8, hipMemcpyHostToDevice));
  EPIC_HIP_TRY(hipMemcpy(S->Sq64, sq.data(), sq.size() * 8, hipMemcpyHostToDevice));
  EPIC_HIP_TRY(hipMemcpy(S->G64, g.data(), g.size() * 8, hipMemcpyHostToDevice));
  EPIC_HIP_TRY(hipMemcpy(S->Sp32, sp32.data(), sp32.size() * 4, hipMemcpyHostToDevice));
  EPIC_HIP_TRY(hipMemcpy(S->Sq32, sq32.data(), sq32.size() * 4, hipMemcpyHostToDevice));
  EPIC_HIP_TRY(hipMemcpy(S->G32, g32.data(), g32.size() * 4, hipMemcpyHostToDevice));
  // the work matrices' padding must start (and so stays) zero
  EPIC_HIP_TRY(hipMemset(S->X64, 0, g.size() * 8));
  EPIC_HIP_TRY(hipMemset(S->Y64, 0, g.size() * 8));
  EPIC_HIP_TRY(hipMemset(S->X32, 0, g.size() * 4));
  EPIC_HIP_TRY(hipMemset(S->Y32, 0, g.size() * 4));
  c->blas = S;
  return hipSuccess;
}

void spec_destroy(Ctx* c) {
  SpecState* S = (SpecState*)c->blas;
  if (!S) return;
  void* ps[] = {S->Sp64, S->Sq64, S->G64, S->X64, S->Y64, S->Sp32, S->Sq32, S->G32, S->X32, S->Y32};
  for (void* p : ps)
    if (p) hipFree(p);
  delete S;
  c->blas = nullptr;
}

// phi (+)= A^-1 src over the unknowns, in precision T: four tile products,
// the scaling by G in the second's epilogue and the write into phi in the
// fourth's.  FP32: X holds the packed residual (written by the residual
// kernel); FP64: src is packed here.
template <typename T>
static hipError_t spec_apply(Ctx* c, SpecState* S, const double* src, double* phi, bool overwrite, const T* Sp,
                             const T* Sq, const T* G, T* X, T* Y) {
  const SpecLayout L = S->L;
  const int P = L.pp, Q = L.qq;
  if (sizeof(T) == 8) {
    const int n = L.p * L.q;
    const int grid = std::min((n + 255) / 256, 4096);
    hipLaunchKernelGGL(pack_kernel<T>, dim3(grid), dim3(256), 0, c->stream, src, X, L);
    EPIC_HIP_TRY(hipGetLastError());
  }
  const TileGemmT<T> t1{Sp, P, X, P, Y, P, nullptr, nullptr, L, 0, 0, 0, 0};
  const TileGemmT<T> t2{Y, P, Sq, Q, X, P, G, nullptr, L, 0, 0, 0, 0};
  const TileGemmT<T> t3{Sp, P, X, P, Y, P, nullptr, nullptr, L, 0, 0, 0, 0};
  const TileGemmT<T> t4{Y, P, Sq, Q, X, P, nullptr, phi, L, overwrite ? 1 : 0, 0, 0, 0};
  EPIC_HIP_TRY(gemm_tile32(c->stream, P, Q, P, t1));
  EPIC_HIP_TRY(gemm_tile32(c->stream, P, Q, Q, t2));
  EPIC_HIP_TRY(gemm_tile32(c->stream, P, Q, P, t3));
  EPIC_HIP_TRY(gemm_tile32(c->stream, P, Q, Q, t4));
  return hipSuccess;
}

// Partial sums of the residual kernels: the latest residual's at
// ws_part[0, g), the first one's (r = b) at ws_part[SPEC_PART_B, +g).
constexpr int SPEC_PART_B = 1024;

static int residual_blocks(const Ctx* c) { return std::min((c->dom.nx * c->dom.ny + 255) / 256, SPEC_PART_B); }

// r = b - A phi and its partial sums.  First residual of a solve (zero_phi
// given): phi := 0 in the same kernel and r = b (phi itself is not read),
// partials to the second region; b is built from rho there when src.rho is
// given.
static hipError_t residual(Ctx* c, const SpecLayout& L, const double* b, const double* phi, double* r, float* X,
                           double* zero_phi = nullptr, RhsSrc src = RhsSrc{nullptr, nullptr, 0, 0},
                           const FinArgs* fin = nullptr) {
  const epic_domain_desc& d = c->dom;
  const int g = residual_blocks(c);
  double* part = c->ws_part + (zero_phi ? SPEC_PART_B : 0);
  const FinArgs fa = fin ? *fin : FinArgs{};
#define EPIC_R(OP, FIN)                                                                                      \
  hipLaunchKernelGGL((spec_residual_kernel<OP, FIN>), dim3(g), dim3(256), 0, c->stream, b, phi, r, d.nx, d.ny, \
                     part, X, L, zero_phi, src, fa)
  if (d.poisson_op == EPIC_OP_REF_GHOST) {
    if (fin) EPIC_R(EPIC_OP_REF_GHOST, true);
    else EPIC_R(EPIC_OP_REF_GHOST, false);
  } else {
    if (fin) EPIC_R(EPIC_OP_DIRICHLET_WALL, true);
    else EPIC_R(EPIC_OP_DIRICHLET_WALL, false);
  }
#undef EPIC_R
  return hipGetLastError();
}

// ||b||^2 and the latest residual's ||r||^2 into out (first_only: no
// correction ran, r = b); seq > 0: also to the pinned host words.
static NormsArgs norms_args(Ctx* c, double* out, bool first_only, unsigned long long seq) {
  const double* pb = c->ws_part + SPEC_PART_B;
  return NormsArgs{first_only ? pb : c->ws_part, pb, residual_blocks(c), out, seq ? c->hscal : nullptr, seq};
}
static hipError_t norms(Ctx* c, double* out, bool first_only, unsigned long long seq) {
  hipLaunchKernelGGL(norms_kernel, dim3(1), dim3(64), 0, c->stream, norms_args(c, out, first_only, seq));
  return hipGetLastError();
}

// Waits until the pinned host words carry seq and reads them.  A stream
// synchronisation spins only briefly and then sleeps on an interrupt, whose
// wake-up costs tens of microseconds once per solve; polling the words
// costs one PCIe read per poll.  After 50 ms the wait falls back to a stream
// synchronisation (which also reports a failed launch).
static hipError_t wait_host_scalars(Ctx* c, unsigned long long seq, double* bb, double* rr) {
  HostScal* h = c->hscal;
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned spin = 1; __atomic_load_n(&h->seq, __ATOMIC_ACQUIRE) != seq; spin++) {
    if ((spin & 255u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
      if (__atomic_load_n(&h->seq, __ATOMIC_ACQUIRE) != seq) return hipErrorUnknown;
      break;
    }
  }
  const unsigned long long b = __atomic_load_n(&h->bb, __ATOMIC_RELAXED);
  const unsigned long long r = __atomic_load_n(&h->rr, __ATOMIC_RELAXED);
  std::memcpy(bb, &b, sizeof(double));
  std::memcpy(rr, &r, sizeof(double));
  return hipSuccess;
}

// Solve A phi = b to ||b - A phi|| <= tol ||b|| on the workspace S.
// fp32_inner: the corrections are FP32 sine-transform solves (mixed
// precision); otherwise FP64.  rho (when given): b is first built from it,
// b = -rho dx^2 s on the unknowns.  poll: the convergence check's norms come
// through the pinned host words (c->hscal); otherwise by a plain copy.
// Returns the number of transform solves in *solves and the final relative
// residual in *relres.
static hipError_t spectral_solve_on(Ctx* c, SpecState* S, double* b, double* phi, bool fp32_inner, double tol,
                                    int max_ref, int* solves, double* relres, const double* rho, double dx2,
                                    double s, bool poll) {
  double* dscal = c->ws_scalar + 4;  // [4] ||b||^2, [5] ||r||^2
  float* Xpack = fp32_inner ? S->X32 : nullptr;
  const SpecLayout& L = S->L;
  // b from rho, phi := 0 and r = b in one kernel (the corrections write
  // phi's interior)
  // (fused step: rho finalised from the deposit in the same launch)
  if (c->rho_in_solve && rho) {
    const FinArgs fa = fin_args(c, (1u << c->nsp) - 1u, 1);
    EPIC_HIP_TRY(residual(c, L, b, nullptr, c->r64, Xpack, phi, RhsSrc{rho, b, dx2, s}, &fa));
    EPIC_HIP_TRY(zero_q_after_fin(c, fa));
  } else {
    EPIC_HIP_TRY(residual(c, L, b, nullptr, c->r64, Xpack, phi, RhsSrc{rho, b, dx2, s}));
  }
  c->rho_in_solve = false;
  int k = 0;
  bool checked = false;
  double host[2] = {0, 0};
  for (; k < max_ref; k++) {
    c->efield_done = false;
    if (fp32_inner)
      EPIC_HIP_TRY(spec_apply<float>(c, S, c->r64, phi, k == 0, S->Sp32, S->Sq32, S->G32, S->X32, S->Y32));
    else
      EPIC_HIP_TRY(spec_apply<double>(c, S, c->r64, phi, k == 0, S->Sp64, S->Sq64, S->G64, S->X64, S->Y64));
    EPIC_HIP_TRY(residual(c, L, b, phi, c->r64, Xpack));
    // the FP32 correction gains ~6 digits: two applications reach 1e-10 and
    // more; check the true residual after the second (the norms are formed
    // only for a check)
    if (k + 1 >= (fp32_inner ? 2 : 1)) {
      if (poll) {
        const unsigned long long seq = ++c->hscal_seq;
        // with E of this phi, in case it is the final one (a further
        // correction launches it again)
        if (c->efield_in_solve) EPIC_HIP_TRY(launch_efield_norms(c, norms_args(c, dscal, false, seq)));
        else EPIC_HIP_TRY(norms(c, dscal, false, seq));
        c->efield_done = c->efield_in_solve;
        EPIC_HIP_TRY(wait_host_scalars(c, seq, &host[0], &host[1]));
      } else {
        EPIC_HIP_TRY(norms(c, dscal, false, 0));
        EPIC_HIP_TRY(hipMemcpyAsync(host, dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
      }
      checked = true;
      if (sqrt(host[1]) <= tol * sqrt(host[0]) || host[1] == 0) {
        k++;
        break;
      }
    }
  }
  if (!checked) {
    EPIC_HIP_TRY(norms(c, dscal, k == 0, 0));
    EPIC_HIP_TRY(hipMemcpyAsync(host, dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EPIC_HIP_TRY(hipStreamSynchronize(c->stream));
  }
  *solves = k;
  *relres = host[0] > 0 ? sqrt(host[1] / host[0]) : 0.0;
  // max_ref corrections left the residual above tolerance (or not finite)
  if (!(*relres <= tol)) c->errc_host[EPIC_ERRC_NOCONV]++;
  return hipSuccess;
}

hipError_t launch_spectral_solve(Ctx* c, double* b, double* phi, bool fp32_inner, double tol, int max_ref,
                                 int* solves, double* relres, const double* rho, double dx2, double s) {
  EPIC_HIP_TRY(spec_init(c));
  if (!c->hscal) {
    EPIC_HIP_TRY(hipHostMalloc((void**)&c->hscal, sizeof(HostScal), hipHostMallocCoherent));
    std::memset(c->hscal, 0, sizeof(HostScal));
  }
  return spectral_solve_on(c, (SpecState*)c->blas, b, phi, fp32_inner, tol, max_ref, solves, relres, rho, dx2, s,
                           true);
}

// ---- the same FP32 solve without allocation ---------------------------------
// The transform matrices and work arrays of launch_spectral_solve come from a
// file-scope device array instead of hipMalloc on the context's first solve,
// and the convergence check reads the norms by a copy instead of pinned host
// words.  Dimensions are padded as the allocating solve pads them; the
// array holds transform dimensions up to 512, i.e. node grids up to
// 514 x 514 (Dirichlet walls) or 512 x 512 (pcgsolver's operator).
constexpr int NOALLOC_MAX_DIM = 512;
constexpr size_t NOALLOC_FLOATS = (size_t)5 * NOALLOC_MAX_DIM * NOALLOC_MAX_DIM;  // Sp, Sq, G, X, Y: 5 MiB
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static float* noalloc_ptr[NOALLOC_MAX_DEVICES];       // scratch's device address, per device
static SpecLayout noalloc_built[NOALLOC_MAX_DEVICES];  // layout of the matrices scratch holds
static bool noalloc_valid[NOALLOC_MAX_DEVICES];


// S_p, S_q, G of the layout into scratch, X and Y zeroed (their padding
// must be zero), in the values spec_init computes (FP64 arithmetic rounded
// to FP32).
__global__ void noalloc_build_kernel(float* ws, SpecLayout L) {
  const int P = L.pp, Q = L.qq, p = L.p, q = L.q;
  float* Sp = ws;
  float* Sq = Sp + (size_t)P * P;
  float* G = Sq + (size_t)Q * Q;
  float* X = G + (size_t)P * Q;
  float* Y = X + (size_t)P * Q;
  const double cc = (2.0 / (p + 1)) * (2.0 / (q + 1));
  const size_t tot = (size_t)P * P + (size_t)Q * Q + (size_t)3 * P * Q;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    if (t < (size_t)P * P) {
      const int a = (int)(t % P), k = (int)(t / P);
      Sp[t] = (a < p && k < p) ? (float)sin(M_PI * (double)(a + 1) * (k + 1) / (p + 1)) : 0.f;
    } else if (t < (size_t)P * P + (size_t)Q * Q) {
      const size_t u = t - (size_t)P * P;
      const int a = (int)(u % Q), k = (int)(u / Q);
      Sq[u] = (a < q && k < q) ? (float)sin(M_PI * (double)(a + 1) * (k + 1) / (q + 1)) : 0.f;
    } else {
      const size_t u = t - (size_t)P * P - (size_t)Q * Q;
      const size_t m = u % ((size_t)P * Q);
      const int which = (int)(u / ((size_t)P * Q));
      if (which == 0) {
        const int k = (int)(m % P), l = (int)(m / P);
        float v = 0.f;
        if (k < p && l < q) {
          const double lam = 2 * cos(M_PI * (k + 1) / (p + 1)) - 2 + 2 * cos(M_PI * (l + 1) / (q + 1)) - 2;
          v = (float)(cc / lam);
        }
        G[m] = v;
      } else {
        (which == 1 ? X : Y)[m] = 0.f;
      }
    }
  }
}

hipError_t launch_spectral_solve_noalloc(Ctx* c, int* solves, double* relres, int* status, const char** why) {
  const epic_domain_desc& d = c->dom;
  SpecLayout L = spec_layout(c);
  L.pp = spec_pad32(L.p);
  L.qq = spec_pad32(L.q);
  *status = 0;
  if (L.pp > NOALLOC_MAX_DIM || L.qq > NOALLOC_MAX_DIM) {
    *status = EPIC_ERR_CAPACITY;
    *why = "grid beyond the no-allocation solve's maximum (transform dimensions up to 512: nodes up to 514 x 514 "
           "with Dirichlet walls, 512 x 512 with pcgsolver's operator)";
    return hipSuccess;
  }
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES) {
    *status = EPIC_ERR_ARG;
    *why = "device index beyond the no-allocation solve's table";
    return hipSuccess;
  }
  float* ws = nullptr;
  bool build = false;
  {
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[c->device]) {
      // the array's device address on this device, looked up once with the
      // device current (the host-side symbol is not a device address)
      int cur = 0;
      EPIC_HIP_TRY(hipGetDevice(&cur));
      if (cur != c->device) EPIC_HIP_TRY(hipSetDevice(c->device));
      void* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (cur != c->device) EPIC_HIP_TRY(hipSetDevice(cur));
      EPIC_HIP_TRY(e);
      noalloc_ptr[c->device] = (float*)ptr;
    }
    ws = noalloc_ptr[c->device];
    const SpecLayout& B = noalloc_built[c->device];
    build = !noalloc_valid[c->device] || B.p != L.p || B.q != L.q || B.pp != L.pp || B.qq != L.qq;
    if (build) {
      noalloc_built[c->device] = L;
      noalloc_valid[c->device] = true;
    }
  }
  if (build) {
    hipLaunchKernelGGL(noalloc_build_kernel, dim3(1024), dim3(256), 0, c->stream, ws, L);
    EPIC_HIP_TRY(hipGetLastError());
  }
  SpecState S{};
  S.L = L;
  S.Sp32 = ws;
  S.Sq32 = S.Sp32 + (size_t)L.pp * L.pp;
  S.G32 = S.Sq32 + (size_t)L.qq * L.qq;
  S.X32 = S.G32 + (size_t)L.pp * L.qq;
  S.Y32 = S.X32 + (size_t)L.pp * L.qq;
  const double tol = d.tol_relative ? d.tolerance : 1e-12;
  const double s = (d.L * d.L) / (d.LDe * d.LDe);
  const bool eis = c->efield_in_solve;
  c->efield_in_solve = false;
  const hipError_t e =
      spectral_solve_on(c, &S, c->b64, c->phi, true, tol, 6, solves, relres, c->rho, d.dx * d.dx, s, false);
  c->efield_in_solve = eis;
  return e;
}

}  // namespace epic
