// Dense products of the sine-transform Poisson solves on the matrix cores:
// C = A B (times G elementwise when G is given), column-major, FP32 on
// v_mfma_f32_16x16x4f32 and FP64 on v_mfma_f64_16x16x4f64.  M, N, K
// multiples of 32 (the transform matrices are zero-padded to multiples of
// 32).  Shared by the single-simulation solves (spectral.hip) and the
// batched ensemble (batch.hip).
//
// One workgroup per 32 x 32 tile of C (blockIdx.z: the product of a batch,
// operands offset by the strides); its four wavefronts take disjoint parts
// of K with a 2 x 2 block of 16x16 accumulators each, and the four partial
// tiles are summed in LDS in wavefront order, so the result does not depend
// on scheduling (nor on how many products share the launch).  Operand layout
// of the 16x16x4 MFMA: lane l holds A[l%16][k] and B[k][l%16] for its
// k = l/16, and D[4(l/16)+r][l%16] in register r.  The contraction index is
// free to permute, so lane group g takes the four consecutive
// k = kb + 4g + e over the four MFMAs e of a 16-deep step: the B operand
// (contiguous along k) is then one vector load per lane.
#pragma once

#include "device_common.hpp"
#include "node_ops.hpp"

namespace epic {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// Operands of one tile GEMM (or of nbatch of them).  dst: the product goes to
// dst through the layout L (overwritten or added) instead of to C.  G: the
// product is scaled by G elementwise (G has C's layout; shared by a batch).
template <typename T>
struct TileGemmT {
  const T* A;
  int lda;
  const T* B;
  int ldb;
  T* C;
  int ldc;
  const T* G;
  double* dst;
  SpecLayout L;
  int overwrite;
  size_t sA, sB, sC;  // batch strides in elements (blockIdx.z)
};
using TileGemm = TileGemmT<float>;

// Epilogue: the NW wavefronts' partial tiles summed in wavefront order.
template <typename T, int NW = 4>
__device__ __forceinline__ void tile_store(const TileGemmT<T>& t, T (*part)[32 * 32], int tm, int tn, T* C) {
  for (int q = threadIdx.x; q < 32 * 32; q += blockDim.x) {
    T v = part[0][q] + part[1][q];
#pragma unroll
    for (int w = 2; w < NW; w++) v += part[w][q];
    const int i = tm + (q & 31), j = tn + (q >> 5);
    const size_t o = (size_t)i + (size_t)j * t.ldc;
    if (t.G) v *= t.G[o];
    if (t.dst) {
      if (i < t.L.p && j < t.L.q) {
        double* d = t.dst + t.L.off + i + (size_t)j * t.L.ld;
        *d = t.overwrite ? (double)v : *d + (double)v;
      }
    } else {
      C[o] = v;
    }
  }
}

// FP32: each wavefront a contiguous quarter of K = 64 NSTEP, every operand
// load of that slice issued before the first product (one or two
// wavefronts per SIMD: the register budget holds the whole slice).
template <int NSTEP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void sgemm_tile32_kernel(TileGemm t) {
  __shared__ float part[4][32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tm = blockIdx.x * 32, tn = blockIdx.y * 32;
  const float* A = t.A + blockIdx.z * t.sA;
  const float* B = t.B + blockIdx.z * t.sB;
  float* C = t.C + blockIdx.z * t.sC;
  const int r16 = lane & 15, g = lane >> 4;
  const int k0 = w * NSTEP * 16;  // this wavefront's quarter of K = 64 NSTEP
  v4f b[NSTEP][2];
  float a[NSTEP][2][4];
#pragma unroll
  for (int st = 0; st < NSTEP; st++) {
    const int kb = k0 + st * 16;
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
      const int col = tn + cb * 16 + r16;
      b[st][cb] = *reinterpret_cast<const v4f*>(B + (size_t)col * t.ldb + kb + 4 * g);
    }
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
      for (int e = 0; e < 4; e++)
        a[st][rb][e] = A[(size_t)(tm + rb * 16 + r16) + (size_t)(kb + 4 * g + e) * t.lda];
  }
  // the scheduler may not move anything across: every load above is in
  // flight before the first product waits for its own operands
  __builtin_amdgcn_sched_barrier(0);
  v4f acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) acc[rb][cb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int st = 0; st < NSTEP; st++)
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[st][rb][e], b[st][cb][e], acc[rb][cb], 0, 0, 0);
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) part[w][(rb * 16 + 4 * g + r) + (cb * 16 + r16) * 32] = acc[rb][cb][r];
  __syncthreads();
  tile_store<float>(t, part, tm, tn, C);
}

// FP32 for K = 32 (2j + 1) (K / 16 blocks not a multiple of 4): wavefront w
// takes the contiguous blocks [w nb / 4, (w + 1) nb / 4) of the nb = K / 16,
// at most NSTEP of them, every operand load issued before the first product
// as in sgemm_tile32_kernel.  A wavefront's unused steps load an in-range
// block and multiply zeros (selected, not branched around, so that no load
// waits for another).
template <int NSTEP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void sgemm_tile32_part_kernel(
    TileGemm t, int K) {
  __shared__ float part[4][32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tm = blockIdx.x * 32, tn = blockIdx.y * 32;
  const float* A = t.A + blockIdx.z * t.sA;
  const float* B = t.B + blockIdx.z * t.sB;
  float* C = t.C + blockIdx.z * t.sC;
  const int r16 = lane & 15, g = lane >> 4;
  const int nb = K / 16, b0 = (w * nb) / 4, cw = ((w + 1) * nb) / 4 - b0;
  v4f b[NSTEP][2];
  float a[NSTEP][2][4];
#pragma unroll
  for (int st = 0; st < NSTEP; st++) {
    const int kb = 16 * min(b0 + st, nb - 1);
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
      const int col = tn + cb * 16 + r16;
      b[st][cb] = *reinterpret_cast<const v4f*>(B + (size_t)col * t.ldb + kb + 4 * g);
    }
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
      for (int e = 0; e < 4; e++)
        a[st][rb][e] = A[(size_t)(tm + rb * 16 + r16) + (size_t)(kb + 4 * g + e) * t.lda];
  }
  __builtin_amdgcn_sched_barrier(0);
  v4f acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) acc[rb][cb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int st = 0; st < NSTEP; st++) {
    const bool used = st < cw;
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(used ? a[st][rb][e] : 0.f, b[st][cb][e], acc[rb][cb],
                                                              0, 0, 0);
  }
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) part[w][(rb * 16 + 4 * g + r) + (cb * 16 + r16) * 32] = acc[rb][cb][r];
  __syncthreads();
  tile_store<float>(t, part, tm, tn, C);
}

// FP64: wavefront w takes the 16-deep k blocks w, w + 4, w + 8, ... (K/16
// blocks in all), U of them per round (all loads of a round issued before
// its products).
template <int U>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void dgemm_tile32_kernel(
    TileGemmT<double> t, int K) {
  __shared__ double part[4][32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tm = blockIdx.x * 32, tn = blockIdx.y * 32;
  const double* A = t.A + blockIdx.z * t.sA;
  const double* B = t.B + blockIdx.z * t.sB;
  double* C = t.C + blockIdx.z * t.sC;
  const int r16 = lane & 15, g = lane >> 4;
  const int nbt = K / 16;  // 16-deep blocks of K
  v4d acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) acc[rb][cb] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int i0 = w; i0 < nbt; i0 += 4 * U) {
    v4d b[U][2];
    double a[U][2][4];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + 4 * u >= nbt) break;
      const int kb = 16 * (i0 + 4 * u);
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        const int col = tn + cb * 16 + r16;
        b[u][cb] = *reinterpret_cast<const v4d*>(B + (size_t)col * t.ldb + kb + 4 * g);
      }
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int e = 0; e < 4; e++) a[u][rb][e] = A[(size_t)(tm + rb * 16 + r16) + (size_t)(kb + 4 * g + e) * t.lda];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + 4 * u >= nbt) break;
#pragma unroll
      for (int e = 0; e < 4; e++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
          for (int cb = 0; cb < 2; cb++)
            acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][rb][e], b[u][cb][e], acc[rb][cb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) part[w][(rb * 16 + 4 * g + r) + (cb * 16 + r16) * 32] = acc[rb][cb][r];
  __syncthreads();
  tile_store<double>(t, part, tm, tn, C);
}

// FP64, K = 128 NB: eight wavefronts, each a contiguous eighth of K (NB
// 16-deep blocks), every operand load of that slice issued before the first
// product (160 registers at NB = 4: two wavefronts per SIMD), so a
// wavefront waits for memory once; the eight partial tiles summed in LDS in
// wavefront order.
template <int NB>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(1, 2))) void dgemm_tile32_w8_kernel(
    TileGemmT<double> t) {
  __shared__ double part[8][32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tm = blockIdx.x * 32, tn = blockIdx.y * 32;
  const double* A = t.A + blockIdx.z * t.sA;
  const double* B = t.B + blockIdx.z * t.sB;
  double* C = t.C + blockIdx.z * t.sC;
  const int r16 = lane & 15, g = lane >> 4;
  const int k0 = w * NB * 16;
  v4d b[NB][2];
  double a[NB][2][4];
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int kb = k0 + 16 * u;
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
      const int col = tn + cb * 16 + r16;
      b[u][cb] = *reinterpret_cast<const v4d*>(B + (size_t)col * t.ldb + kb + 4 * g);
    }
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
      for (int e = 0; e < 4; e++) a[u][rb][e] = A[(size_t)(tm + rb * 16 + r16) + (size_t)(kb + 4 * g + e) * t.lda];
  }
  __builtin_amdgcn_sched_barrier(0);
  v4d acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) acc[rb][cb] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < NB; u++)
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][rb][e], b[u][cb][e], acc[rb][cb], 0, 0, 0);
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) part[w][(rb * 16 + 4 * g + r) + (cb * 16 + r16) * 32] = acc[rb][cb][r];
  __syncthreads();
  tile_store<double, 8>(t, part, tm, tn, C);
}

// FP32 for K beyond the register-resident slice (K > TILE32_MAX_K): the
// FP64 kernel's round structure, U 16-deep blocks per round.
template <int U>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void sgemm_tile32_loop_kernel(
    TileGemm t, int K) {
  __shared__ float part[4][32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tm = blockIdx.x * 32, tn = blockIdx.y * 32;
  const float* A = t.A + blockIdx.z * t.sA;
  const float* B = t.B + blockIdx.z * t.sB;
  float* C = t.C + blockIdx.z * t.sC;
  const int r16 = lane & 15, g = lane >> 4;
  const int nbt = K / 16;
  v4f acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) acc[rb][cb] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int i0 = w; i0 < nbt; i0 += 4 * U) {
    v4f b[U][2];
    float a[U][2][4];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + 4 * u >= nbt) break;
      const int kb = 16 * (i0 + 4 * u);
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        const int col = tn + cb * 16 + r16;
        b[u][cb] = *reinterpret_cast<const v4f*>(B + (size_t)col * t.ldb + kb + 4 * g);
      }
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int e = 0; e < 4; e++) a[u][rb][e] = A[(size_t)(tm + rb * 16 + r16) + (size_t)(kb + 4 * g + e) * t.lda];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + 4 * u >= nbt) break;
#pragma unroll
      for (int e = 0; e < 4; e++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
          for (int cb = 0; cb < 2; cb++)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][rb][e], b[u][cb][e], acc[rb][cb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) part[w][(rb * 16 + 4 * g + r) + (cb * 16 + r16) * 32] = acc[rb][cb][r];
  __syncthreads();
  tile_store<float>(t, part, tm, tn, C);
}

// Largest K of the register-resident FP32 kernels (K = 64 NSTEP, NSTEP <= 9,
// or K = 32 (2j + 1) up to it); larger K take the round-by-round kernel.
constexpr int TILE32_MAX_K = 576;

inline bool tile32_ok(int m, int n, int k) { return m % 32 == 0 && n % 32 == 0 && k % 32 == 0 && k >= 32; }

// C (m x n) = A (m x k) B (k x n) (times G), nbatch products (blockIdx.z).
// Returns hipErrorInvalidValue for shapes the kernels do not take.
inline hipError_t gemm_tile32(hipStream_t st, int m, int n, int k, const TileGemmT<float>& t, int nbatch = 1) {
  if (!tile32_ok(m, n, k) || nbatch < 1) return hipErrorInvalidValue;
  const dim3 grid(m / 32, n / 32, nbatch);
  if (k % 64 != 0 && k <= TILE32_MAX_K) {
    switch ((k / 16 + 3) / 4) {
#define EPIC_P(N) \
  case N: hipLaunchKernelGGL(sgemm_tile32_part_kernel<N>, grid, dim3(256), 0, st, t, k); break;
      EPIC_P(1) EPIC_P(2) EPIC_P(3) EPIC_P(4) EPIC_P(5) EPIC_P(6) EPIC_P(7) EPIC_P(8) EPIC_P(9)
#undef EPIC_P
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (k > TILE32_MAX_K || k % 64 != 0) {
    hipLaunchKernelGGL(sgemm_tile32_loop_kernel<4>, grid, dim3(256), 0, st, t, k);
    return hipGetLastError();
  }
  switch (k / 64) {
#define EPIC_G(N) \
  case N: hipLaunchKernelGGL(sgemm_tile32_kernel<N>, grid, dim3(256), 0, st, t); break;
    EPIC_G(1) EPIC_G(2) EPIC_G(3) EPIC_G(4) EPIC_G(5) EPIC_G(6) EPIC_G(7) EPIC_G(8) EPIC_G(9)
#undef EPIC_G
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

inline hipError_t gemm_tile32(hipStream_t st, int m, int n, int k, const TileGemmT<double>& t, int nbatch = 1) {
  if (!tile32_ok(m, n, k) || nbatch < 1) return hipErrorInvalidValue;
  const dim3 grid(m / 32, n / 32, nbatch);
  // K = 128 NB up to 512: the eight-wavefront kernel with the whole slice in
  // registers; other K the round-by-round kernel
  switch (k % 128 == 0 ? k / 128 : 0) {
    case 1: hipLaunchKernelGGL(dgemm_tile32_w8_kernel<1>, grid, dim3(512), 0, st, t); break;
    case 2: hipLaunchKernelGGL(dgemm_tile32_w8_kernel<2>, grid, dim3(512), 0, st, t); break;
    case 3: hipLaunchKernelGGL(dgemm_tile32_w8_kernel<3>, grid, dim3(512), 0, st, t); break;
    case 4: hipLaunchKernelGGL(dgemm_tile32_w8_kernel<4>, grid, dim3(512), 0, st, t); break;
    default: hipLaunchKernelGGL(dgemm_tile32_kernel<2>, grid, dim3(256), 0, st, t, k); break;
  }
  return hipGetLastError();
}

// Transform dimensions padded to a multiple of 32 with zero rows and
// columns: S, G and the work matrices are zero there, so the padding stays
// zero through every product.
inline int spec_pad32(int p) { return (p + 31) / 32 * 32; }

}  // namespace epic
