// Fixed-point deposit finalisation: Q (int64, 2^B per particle) -> species
// density den_s and charge density rho.
//
//   den_s = Q * 2^-B * spwt / (dx*dy*L^2*density)        species.cpp:188-190
//   pbc: each node takes the sum over its periodic images  species.cpp:193-205
//        (done on the integers, so the fold is exact)
//   open / reflective: edge nodes x 1/2, corners x 1/4     species.cpp:206-225
//   rho = 0; rho += den_s * (charge_s * (1/QE))             domain.cpp:194-202
#include "device_common.hpp"
#include "epic_internal.hpp"
#include "node_ops.hpp"

namespace epic {

__global__ void rho_finalize_kernel(FinArgs a) {
  const int64_t nn = (int64_t)a.nx * a.ny;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nn; k += (int64_t)gridDim.x * blockDim.x)
    fin_node(a, k);
}

// Clears the accumulators of every species in mask in one launch (under
// periodic walls a node's fold reads its images, which other threads read
// too, so rho_finalize_kernel cannot clear them as it reads).
struct ZeroQArgs {
  unsigned long long* Q[FIN_MAXSP];
  int nsp;
  unsigned mask;
  int64_t nn;
};

__global__ void zero_q_kernel(ZeroQArgs a) {
  for (int s = 0; s < a.nsp; s++) {
    if (!(a.mask & (1u << s))) continue;
    unsigned long long* q = a.Q[s];
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < a.nn; k += (int64_t)gridDim.x * blockDim.x)
      q[k] = 0ull;
  }
}

static hipError_t zero_q(Ctx* c, unsigned mask) {
  if (!mask) return hipSuccess;
  ZeroQArgs a{};
  a.nsp = c->nsp;
  a.mask = mask;
  a.nn = (int64_t)c->dom.nx * c->dom.ny;
  for (int s = 0; s < c->nsp; s++) a.Q[s] = c->sp[s].Q;
  const int grid = (int)std::min<int64_t>((a.nn + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_q_kernel, dim3(grid), dim3(256), 0, c->stream, a);
  return hipGetLastError();
}

FinArgs fin_args(Ctx* c, unsigned mask, int do_rho) {
  const epic_domain_desc& d = c->dom;
  FinArgs a{};
  a.nsp = c->nsp;
  a.mask = mask;
  a.do_rho = do_rho;
  for (int s = 0; s < c->nsp; s++) {
    a.Q[s] = c->sp[s].Q;
    a.den[s] = c->sp[s].den;
    a.scale[s] = ldexp(1.0, -c->fix_bits) * c->sp[s].desc.spwt;
    a.factor[s] = c->sp[s].desc.charge * (1 / QE);
  }
  const double cell_area = d.dx * d.dy;
  const double L2 = d.L * d.L;
  a.norm = (cell_area * L2 * d.density);
  a.nx = d.nx; a.ny = d.ny; a.bc = d.bc;
  a.rho = c->rho;
  a.errc = c->errc;
  a.clear_q = d.bc != EPIC_BC_PBC;
  return a;
}

hipError_t zero_q_after_fin(Ctx* c, const FinArgs& a) { return a.clear_q ? hipSuccess : zero_q(c, a.mask); }

static hipError_t finalize(Ctx* c, unsigned mask, int do_rho) {
  const epic_domain_desc& d = c->dom;
  const FinArgs a = fin_args(c, mask, do_rho);
  const int64_t nn = (int64_t)d.nx * d.ny;
  int grid = (int)std::min<int64_t>((nn + 255) / 256, 4096);
  hipLaunchKernelGGL(rho_finalize_kernel, dim3(grid), dim3(256), 0, c->stream, a);
  EPIC_HIP_TRY(hipGetLastError());
  return a.clear_q ? hipSuccess : zero_q(c, mask);
}

// All species' Q -> den and rho (the fused-step path).
hipError_t launch_rho_finalize(Ctx* c) {
  unsigned mask = 0;
  for (int s = 0; s < c->nsp; s++) mask |= 1u << s;
  return finalize(c, mask, 1);
}

// One species' Q -> den (ScatterSpecies_serial) when called with s >= 0;
// rho from the stored den of all species (ComputeRho) when s < 0.
hipError_t launch_den_only(Ctx* c, int s) {
  if (s >= 0) return finalize(c, 1u << s, 0);
  return finalize(c, 0u, 1);
}

hipError_t launch_zero_Q(Ctx* c) {
  unsigned mask = 0;
  for (int s = 0; s < c->nsp; s++) mask |= 1u << s;
  return zero_q(c, mask);
}

}  // namespace epic
