// Device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "epic2d.h"

namespace epic {

// ---- Philox4x32-10 ---------------------------------------------------------
// Counter-based generator of Salmon, Moraes, Dror and Shaw (SC'11): 10 rounds
// of multiply-hi/lo mixing with a Weyl-sequence key schedule.  Stateless: a
// draw is a pure function of (counter, key), so every particle gets its own
// reproducible stream keyed by (particle id, step).
struct u32x4 {
  uint32_t a, b, c, d;
};

// The 32 x 32 -> 64-bit products are written as one 64-bit multiply, which
// gfx950 does in one v_mad_u64_u32 (instead of a multiply-hi and a
// multiply-lo).
__device__ __forceinline__ u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * ctr.a;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr.c;
    u32x4 n;
    n.a = (uint32_t)(p1 >> 32) ^ ctr.b ^ k0;
    n.b = (uint32_t)p1;
    n.c = (uint32_t)(p0 >> 32) ^ ctr.d ^ k1;
    n.d = (uint32_t)p0;
    ctr = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return ctr;
}

// Philox2x32-10 of the same family (one multiply-hi/lo per round, 32-bit key
// with a Weyl increment): 64 random bits for half the multiplies of
// Philox4x32-10, for draws that need no more.
__device__ __forceinline__ uint2 philox2x32_10(uint32_t c0, uint32_t c1, uint32_t k) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t p = (uint64_t)0xD256D193u * c0;
    c0 = (uint32_t)(p >> 32) ^ k ^ c1;
    c1 = (uint32_t)p;
    k += 0x9E3779B9u;
  }
  return make_uint2(c0, c1);
}

// 53 random bits -> [0, 1)
__device__ __forceinline__ double u01_53(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  return (double)m * (1.0 / 9007199254740992.0);
}
// 24 random bits -> [0, 1)
__device__ __forceinline__ float u01_24(uint32_t a) { return (float)(a >> 8) * (1.0f / 16777216.0f); }

// Stream purposes: the third counter word separates independent uses.
enum : uint32_t {
  RNG_LOAD_POS = 1,
  RNG_LOAD_VEL = 2,
  RNG_LOAD_VEL2 = 3,
  RNG_MCC = 4,
  RNG_MCC2 = 5,
  RNG_MCC3 = 6,
  RNG_INJECT = 7,
  RNG_MCC_KEY = 8,  // host-side derivation of the candidate draws' Philox2x32 key
  RNG_MCC_ION_SCATTER = 12,
  RNG_MCC_NEWION = 13,
  RNG_MCC_NEWION2 = 14,
  RNG_INJECT2 = 16,  // 16 + 2*target, 17 + 2*target
  RNG_INJECT3 = 20,  // 20 + target
};

// ---- reductions ------------------------------------------------------------
// Fixed butterfly over the 64-lane wavefront: the same lanes are combined in
// the same order every run, so FP sums are reproducible.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// Wave-uniform min / max of an int over the 64 lanes with DPP moves: an
// inclusive scan inside each 16-lane row (row_shr 1, 2, 4, 8), the row
// results carried across rows (row_bcast 15, 31), lane 63 read back.  Lanes
// whose DPP source is outside the row or masked off contribute the identity.
template <bool MIN>
__device__ __forceinline__ int wave_ext_dpp(int v) {
  const int id = MIN ? 0x7fffffff : (int)0x80000000;
  auto op = [](int a, int b) { return MIN ? min(a, b) : max(a, b); };
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min_dpp(int v) { return wave_ext_dpp<true>(v); }
__device__ __forceinline__ int wave_max_dpp(int v) { return wave_ext_dpp<false>(v); }

// One DPP move of a double (two 32-bit halves); lanes without a source in
// the pattern receive +0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Wave sum of a double with DPP moves (the scan of wave_ext_dpp: row_shr 1,
// 2, 4, 8 inside each 16-lane row, row_bcast 15 and 31 across rows), read
// back from lane 63: uniform, fixed order, no LDS traffic.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_f64<0x111, 0xf>(v);  // row_shr:1
  v += dpp_f64<0x112, 0xf>(v);  // row_shr:2
  v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
  v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)u, 63);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Block sum of NV doubles; result valid in every thread.  `sh` holds
// NV * (blockDim/64) doubles.  Deterministic (fixed order).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum_dpp(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k * nw + wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0;
    for (int w = 0; w < nw; w++) s += sh[k * nw + w];
    v[k] = s;
  }
  __syncthreads();
}

// Block sum without the closing barrier: `sh` (NV * (blockDim/64) doubles)
// may be written again only after one more workgroup barrier, so callers
// alternate two regions between consecutive sums.
template <int NV, int NW>
__device__ __forceinline__ void block_sum_open(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int nw = NW;  // wavefronts per workgroup
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum_dpp(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k * nw + wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < nw; w++) s += sh[k * nw + w];
    v[k] = s;
  }
}

__device__ __forceinline__ void atomic_add_u64(unsigned long long* p, unsigned long long v) {
  atomicAdd(p, v);
}

// XCD-aware chunk walk: MI355X dispatches consecutive workgroups round-robin
// over its 8 XCDs.  Giving XCD x a contiguous run of chunks keeps a chunk
// range's grid-cell working set in that XCD's own L2.
__device__ __forceinline__ void xcd_chunk_range(int64_t nchunks, int64_t* begin, int64_t* end,
                                                int64_t* step) {
  const int nxcd = 8;
  const int g = gridDim.x;
  if (g % nxcd != 0 || g < nxcd) {
    *begin = blockIdx.x;
    *end = nchunks;
    *step = g;
    return;
  }
  const int xcd = blockIdx.x % nxcd;
  const int local = blockIdx.x / nxcd;
  const int per = g / nxcd;
  const int64_t span = (nchunks + nxcd - 1) / nxcd;
  const int64_t lo = (int64_t)xcd * span;
  int64_t hi = lo + span;
  if (hi > nchunks) hi = nchunks;
  *begin = lo + local;
  *end = hi;
  *step = per;
}

// Push error tally of one thread: gathers outside the grid in the low half,
// particles whose new position is not finite (a non-finite velocity or
// field) in the high half; such a particle leaves the list.
constexpr int NAN_UNIT = 1 << 16;
__device__ __forceinline__ void flush_push_errors(unsigned long long* errc, int bad) {
  if (bad & (NAN_UNIT - 1)) atomicAdd(&errc[EPIC_ERRC_GATHER_OOB], (unsigned long long)(bad & (NAN_UNIT - 1)));
  if (bad >= NAN_UNIT) atomicAdd(&errc[EPIC_ERRC_NAN], (unsigned long long)(bad / NAN_UNIT));
}

}  // namespace epic
