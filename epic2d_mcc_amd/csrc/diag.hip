// Energy / momentum diagnostics: FP64 reductions over particles and nodes.
//   Species::Compute_KE        species.cpp:248-271   sum of v_c^2
//   Species::Compute_Momentum  species.cpp:273-297   sum of v_c
//   Domain::Compute_PE         domain.cpp:266-286    sum of efx^2 + efy^2
// Two passes: per-block partials in fixed lane/wave order, then one block
// sums the partials in order: reproducible FP64 results.
#include "device_common.hpp"
#include "epic_internal.hpp"

namespace epic {

constexpr int DIAG_BLOCK = 256;

template <typename P>
__global__ __launch_bounds__(DIAG_BLOCK) void moments_kernel(const P* x, const P* vx, const P* vy, const P* vz,
                                                             const int64_t* d_n, double* part) {
  __shared__ double sh[7 * (DIAG_BLOCK / 64)];
  const int64_t n = *d_n;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = pslot<P>(k);
    if (isnan(x[q])) continue;
    const double a = vx[q], b = vy[q], c = vz[k];
    acc[0] += a * a; acc[1] += b * b; acc[2] += c * c;
    acc[3] += a; acc[4] += b; acc[5] += c;
    acc[6] += 1.0;
  }
  block_sum<7>(acc, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < 7; q++) part[q * gridDim.x + blockIdx.x] = acc[q];
}

__global__ void field_energy_kernel(const double* efx, const double* efy, int64_t n, double* part) {
  __shared__ double sh[DIAG_BLOCK / 64];
  double acc[1] = {0};
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const double ef2 = efx[k] * efx[k] + efy[k] * efy[k];
    acc[0] += ef2;
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void final_sum_kernel(const double* part, int nblk, int nq, double* out) {
  // one wavefront per quantity
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (q >= nq) return;
  double s = 0;
  for (int b = lane; b < nblk; b += 64) s += part[q * nblk + b];
  s = wave_sum(s);
  if (lane == 0) out[q] = s;
}

// out7: sum vx^2, vy^2, vz^2, vx, vy, vz, live count
hipError_t launch_particle_moments(Ctx* c, int s, double out6[7]) {
  SpeciesDev& S = c->sp[s];
  const int nblk = c->red_blocks;
  if (c->dom.precision == EPIC_FP64)
    hipLaunchKernelGGL(moments_kernel<double>, dim3(nblk), dim3(DIAG_BLOCK), 0, c->stream, (const double*)S.x,
                       (const double*)S.vx, (const double*)S.vy, (const double*)S.vz, S.d_n, c->red_part);
  else
    hipLaunchKernelGGL(moments_kernel<float>, dim3(nblk), dim3(DIAG_BLOCK), 0, c->stream, (const float*)S.x,
                       (const float*)S.vx, (const float*)S.vy, (const float*)S.vz, S.d_n, c->red_part);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(7 * 64), 0, c->stream, c->red_part, nblk, 7, c->red_out);
  EPIC_HIP_TRY(hipGetLastError());
  EPIC_HIP_TRY(hipMemcpyAsync(out6, c->red_out, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return hipStreamSynchronize(c->stream);
}

hipError_t launch_field_energy(Ctx* c, double* sum_e2) {
  const int nblk = c->red_blocks;
  hipLaunchKernelGGL(field_energy_kernel, dim3(nblk), dim3(DIAG_BLOCK), 0, c->stream, c->efx, c->efy,
                     (int64_t)c->nnodes, c->red_part);
  EPIC_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(64), 0, c->stream, c->red_part, nblk, 1, c->red_out);
  EPIC_HIP_TRY(hipGetLastError());
  EPIC_HIP_TRY(hipMemcpyAsync(sum_e2, c->red_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return hipStreamSynchronize(c->stream);
}

}  // namespace epic
