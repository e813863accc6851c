// Per-node arithmetic of the field stages, shared by the single-simulation
// kernels (deposit.hip, field.hip, spectral.hip) and the batched ensemble
// kernels (batch.hip), so both compute every node value with the same
// operations in the same order.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/epic2d.h"
#include "device_common.hpp"
#include "epic_internal.hpp"

namespace epic {

// Convergence norms of a mixed solve from the residual kernels' per-block
// partial sums (pr: the latest residual's, pb: the first residual's, whose
// r is b): ||b||^2 and ||r||^2 to out[0], out[1]; hs (when given): both, then
// seq with a system-scope release, to pinned host memory for the host's
// polling wait.  One wavefront; n = 0: nothing to form.
struct NormsArgs {
  const double* pr;
  const double* pb;
  int n;
  double* out;
  HostScal* hs;
  unsigned long long seq;
};

__device__ __forceinline__ void norms_wave(const NormsArgs& a) {
  double sr = 0, sb = 0;
  for (int b = threadIdx.x & 63; b < a.n; b += 64) {
    sr += a.pr[b];
    sb += a.pb[b];
  }
  sr = wave_sum(sr);
  sb = wave_sum(sb);
  if ((threadIdx.x & 63) == 0) {
    a.out[0] = sb;
    a.out[1] = sr;
    if (a.hs) {
      __hip_atomic_store(&a.hs->bb, (unsigned long long)__double_as_longlong(sb), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&a.hs->rr, (unsigned long long)__double_as_longlong(sr), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&a.hs->seq, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// Fixed-point density of node (i, j) from one species' accumulator Q
// (species.cpp:188-225): pbc takes the sum over the node's periodic images on
// the integers (exact); Returns the raw sum; the caller scales and applies
// the open/reflective edge factors with den_from_q.
__device__ __forceinline__ unsigned long long fold_q(const unsigned long long* Q, int i, int j, int nx, int ny,
                                                     int bc) {
  if (bc != EPIC_BC_PBC) return Q[(size_t)i * ny + j];
  const bool iedge = (i == 0 || i == nx - 1), jedge = (j == 0 || j == ny - 1);
  const int ia0 = iedge ? 0 : i, ia1 = iedge ? nx - 1 : i;
  const int ja0 = jedge ? 0 : j, ja1 = jedge ? ny - 1 : j;
  unsigned long long q = Q[(size_t)ia0 * ny + ja0];
  if (ja1 != ja0) q += Q[(size_t)ia0 * ny + ja1];
  if (ia1 != ia0) {
    q += Q[(size_t)ia1 * ny + ja0];
    if (ja1 != ja0) q += Q[(size_t)ia1 * ny + ja1];
  }
  return q;
}

// den = q * 2^-B * spwt / (dx dy L^2 n0), edges x1/2 and corners x1/4 for
// open / reflective walls (species.cpp:188-190, 206-225).
__device__ __forceinline__ double den_from_q(long long q, double scale, double norm, int i, int j, int nx, int ny,
                                             int bc) {
  double den = ((double)q * scale) / norm;
  if (bc != EPIC_BC_PBC) {
    if (i == 0 || i == nx - 1) den *= 0.5;
    if (j == 0 || j == ny - 1) den *= 0.5;
  }
  return den;
}

// Q -> den and rho of one node: the arguments of the finalisation.
constexpr int FIN_MAXSP = 16;
struct FinArgs {
  unsigned long long* Q[FIN_MAXSP];
  double* den[FIN_MAXSP];
  double scale[FIN_MAXSP];   // 2^-B * spwt
  double factor[FIN_MAXSP];  // charge * (1/QE)
  int nsp;
  unsigned mask;         // species whose den is (re)computed from Q
  int do_rho;            // 1: rho from den of all species
  double norm;           // dx*dy*L^2*density
  int nx, ny, bc;
  double* rho;
  unsigned long long* errc;
  int clear_q;           // clear each read accumulator (walls without images)
};

// den_s of node k from Q_s (species in mask) and rho = sum den_s q_s/e
// (domain.cpp:194-202); returns rho (0 when do_rho is 0).
__device__ __forceinline__ double fin_node(const FinArgs& a, int64_t k) {
  const int i = (int)(k / a.ny), j = (int)(k % a.ny);
  double rho = 0;
  for (int s = 0; s < a.nsp; s++) {
    double den;
    if (a.mask & (1u << s)) {
      const long long qs = (long long)fold_q(a.Q[s], i, j, a.nx, a.ny, a.bc);
      // without periodic images a node reads only its own accumulator
      if (a.clear_q) a.Q[s][k] = 0ull;
      if (qs < 0) atomicAdd(&a.errc[EPIC_ERRC_FIXED_OVF], 1ull);
      den = den_from_q(qs, a.scale[s], a.norm, i, j, a.nx, a.ny, a.bc);
      a.den[s][k] = den;
    } else {
      den = a.den[s][k];
    }
    if (a.do_rho) rho += den * a.factor[s];
  }
  if (a.do_rho) a.rho[k] = rho;
  return rho;
}

// Right-hand side of pcgsolver (field.cpp:234-251): -rho dx^2 (L/LDe)^2,
// zero on every boundary node.
__device__ __forceinline__ double rhs_at(double rho, int i, int j, int nx, int ny, double dx2, double s) {
  const bool bnd = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
  return bnd ? 0.0 : -rho * dx2 * s;
}

// The 5-point operator of the Poisson solvers.
//   EPIC_OP_REF_GHOST: pcgsolver's dense A (field.cpp:223-232): flat index
//     k = row*nx + col over all nodes, -4 on the diagonal, neighbours outside
//     the block are zero (ghost ring);
//   EPIC_OP_DIRICHLET_WALL: interior nodes (i*ny + j) with phi = 0 on the
//     boundary (gssolver's problem, field.cpp:89-98).
template <int OP>
struct Stencil {
  int nx, ny;
  __device__ __forceinline__ bool unknown(int k) const {
    if (OP == EPIC_OP_REF_GHOST) return true;
    const int i = k / ny, j = k % ny;
    return i > 0 && i < nx - 1 && j > 0 && j < ny - 1;
  }
  // Row k of A applied to vec(), non-zero entries in increasing column order.
  template <typename F>
  __device__ __forceinline__ double apply(int k, double self, F vec) const {
    if (OP == EPIC_OP_REF_GHOST) {
      const int row = k / nx, col = k % nx;
      double sum = 0;
      if (row > 0) sum += vec(k - nx);
      if (col > 0) sum += vec(k - 1);
      sum += -4.0 * self;
      if (col < nx - 1) sum += vec(k + 1);
      if (row < ny - 1) sum += vec(k + nx);
      return sum;
    } else {
      double sum = 0;
      sum += vec(k - ny);
      sum += vec(k - 1);
      sum += -4.0 * self;
      sum += vec(k + 1);
      sum += vec(k + ny);
      return sum;
    }
  }
  // Which of row k's off-diagonal entries exist (bit 0: k - stride, 1: k - 1,
  // 2: k + 1, 3: k + stride), for apply_mask: the index arithmetic of apply
  // done once instead of per product.
  __device__ __forceinline__ unsigned mask(int k) const {
    if (OP != EPIC_OP_REF_GHOST) return 0xfu;
    const int row = k / nx, col = k % nx;
    return (row > 0 ? 1u : 0u) | (col > 0 ? 2u : 0u) | (col < nx - 1 ? 4u : 0u) | (row < ny - 1 ? 8u : 0u);
  }
  // apply() with the entries given by mask(k): the same terms in the same order.
  template <typename F>
  __device__ __forceinline__ double apply_mask(int k, unsigned m, double self, F vec) const {
    const int st = OP == EPIC_OP_REF_GHOST ? nx : ny;
    double sum = 0;
    if (m & 1u) sum += vec(k - st);
    if (m & 2u) sum += vec(k - 1);
    sum += -4.0 * self;
    if (m & 4u) sum += vec(k + 1);
    if (m & 8u) sum += vec(k + st);
    return sum;
  }
};

// r_k = b_k - (A phi)_k over the unknowns, 0 elsewhere; pv(m) reads phi_m.
template <class PV>
__device__ __forceinline__ double residual_with(int op, const double* b, PV pv, int k, int nx, int ny) {
  if (op == EPIC_OP_REF_GHOST) {
    const Stencil<EPIC_OP_REF_GHOST> S{nx, ny};
    return b[k] - S.apply(k, pv(k), pv);
  }
  const Stencil<EPIC_OP_DIRICHLET_WALL> S{nx, ny};
  if (!S.unknown(k)) return 0.0;
  return b[k] - S.apply(k, pv(k), pv);
}

__device__ __forceinline__ double residual_at(int op, const double* b, const double* phi, int k, int nx, int ny) {
  return residual_with(op, b, [&](int m) -> double { return phi[m]; }, k, nx, ny);
}

// The residual of phi = 0, in the same arithmetic (so the same bits as
// residual_at over a zeroed phi).
__device__ __forceinline__ double residual_at_zero(int op, const double* b, int k, int nx, int ny) {
  return residual_with(op, b, [](int) -> double { return 0.0; }, k, nx, ny);
}

// FieldSolve::CalculateEfield (field.cpp:276-331) at node (i, j): central
// differences inside; pbc edges (phi_1 - phi_{n-2}) / 2dx; open/reflective
// edges one-sided.
__device__ __forceinline__ void efield_at(const double* phi, int i, int j, int nx, int ny, double dx, double dy,
                                          int bc, double* ex_out, double* ey_out) {
  auto ph = [&](int ii, int jj) { return phi[(size_t)ii * ny + jj]; };
  double ex, ey;
  if (i > 0 && i < nx - 1) {
    ex = -(ph(i + 1, j) - ph(i - 1, j)) / (2 * dx);
  } else if (bc == EPIC_BC_PBC) {
    ex = -(ph(1, j) - ph(nx - 2, j)) / (2 * dx);
  } else if (i == 0) {
    ex = -(ph(1, j) - ph(0, j)) / dx;
  } else {
    ex = -(ph(nx - 1, j) - ph(nx - 2, j)) / dx;
  }
  if (j > 0 && j < ny - 1) {
    ey = -(ph(i, j + 1) - ph(i, j - 1)) / (2 * dy);
  } else if (bc == EPIC_BC_PBC) {
    ey = -(ph(i, 1) - ph(i, ny - 2)) / (2 * dy);
  } else if (j == 0) {
    ey = -(ph(i, 1) - ph(i, 0)) / dy;
  } else {
    ey = -(ph(i, ny - 1) - ph(i, ny - 2)) / dy;
  }
  *ex_out = ex;
  *ey_out = ey;
}

// Column-major p x q block of unknowns inside the node array (sine-transform
// solver): element (a, b) is node off + a + b*ld.
// The p x q block of unknowns: node (a, b) of the block is grid node
// off + a + b*ld; in the transform matrices it is entry a + b*pp (pp >= p:
// rows padded with zeros so the GEMM dimensions are multiples of the tile).
struct SpecLayout {
  int p, q, ld, off;
  int pp, qq;  // padded transform dimensions (>= p, >= q)
};

// Index of node k inside the block, or -1 when k is not an unknown.
__device__ __forceinline__ int spec_index(const SpecLayout& L, int k) {
  const int t = k - L.off;
  if (t < 0) return -1;
  const int a = t % L.ld, b = t / L.ld;
  if (a >= L.p || b >= L.q) return -1;
  return a + b * L.pp;
}

}  // namespace epic
