// Null-collision Monte-Carlo collisions against tabulated cross sections.
//
// Replaces CollisionHandler::handle_collisions (collision.cpp:190-207), which
// draws P = 1 - exp(-nu(E) dt) per electron.  Here every particle of the
// colliding species draws from its own counter-based streams keyed by
// (particle id, step): R0 from Philox2x32-10 under a per-(seed, species) key
// makes it a candidate with probability P_null = 1 - exp(-nu_max dt), and a
// candidate collides for real with probability (1 - exp(-nu(E) dt)) / P_null
// (R1 and the later draws from Philox4x32-10; else a null collision), so each
// particle's collision probability is exactly the reference's; only the
// candidates (a fraction ~nu_max dt) evaluate cross sections.
//
// Electron-neutral kinematics follow CollisionHandler::collision_electron
// (collision.cpp:60-188): centre-of-mass split F1 = m_e/(m_e+M), F2 = M/(m_e+M),
// Euler angles of g, isotropic chi = acos(1-2R), excitation and ionisation
// thresholds, ejected-electron energy 10 tan(R atan(E/20)) eV.  Ionisation
// births are written to an event list, ranked by parent slot and appended in
// that order, so slots and ids of new particles do not depend on scheduling.
//
// Ion-neutral (Ar+ - Ar): isotropic elastic scattering in the centre-of-mass
// frame and resonant charge exchange (the ion takes the neutral's velocity),
// against Maxwellian neutrals at gas_temp_K.  No reference counterpart.
#include <algorithm>
#include <cmath>
#include <vector>

#include "device_common.hpp"
#include "epic_internal.hpp"

namespace epic {

constexpr int MCC_BLOCK = 256;
constexpr int EVW = 8;  // doubles per ionisation event: slot, x, y, vex, vey, vez, (ion from RNG)

struct MccTab {
  const double* sigma;  // NPROC * n
  int n;
  double loge0, inv_dloge;
  // where energy e falls on the log-spaced grid, found once for all the
  // processes read at e: entry k and the fraction f towards k + 1 (f < 0:
  // outside the grid, entry k as it is)
  struct Pos {
    int k;
    double f;
  };
  __device__ __forceinline__ Pos pos(double e) const {
    if (!(e > 0)) return {0, -1.0};
    const double t = (log(e) - loge0) * inv_dloge;
    if (t <= 0) return {0, -1.0};
    if (t >= n - 1) return {n - 1, -1.0};
    const int k = (int)t;
    return {k, t - k};
  }
  __device__ __forceinline__ double at(int p, Pos q) const {
    if (q.f < 0) return sigma[p * n + q.k];
    return sigma[p * n + q.k] * (1 - q.f) + sigma[p * n + q.k + 1] * q.f;
  }
};

template <typename P>
struct MccArgs {
  P* x;
  P* y;
  P* vx;
  P* vy;
  P* vz;
  const uint32_t* id;
  const int64_t* d_n;
  MccTab tab;
  int kind;             // 0 electron, 1 ion
  unsigned enable;      // bit per process
  int ref_selection;
  double nu_max, p_null, dt;
  double vel_norm, mass, gas_mass, gas_density, gas_vth;  // gas_vth normalised sqrt(kT/M)
  double e_exc, e_ion;  // eV
  double ion_vth;       // normalised thermal spread of newborn ions
  uint32_t step, sp, k0, k1;
  uint32_t key_r0;  // Philox2x32 key of the candidate draw R0 (per seed and species)
  uint64_t r0_below;  // R0 < p_null exactly when its 53-bit integer is below this (ceil(p_null 2^53))
  unsigned long long* counts;
  double* events;
  int* event_count;
  int event_cap;
  int slot;  // batched runs: index of the simulation's step counter
  uint2* list;     // real collisions: (slot, process | (event slot + 1) << 3)
  int* list_n;     // entries in list (cleared by the finish kernel)
  int* chunk_ev;   // electrons: stored ionisation events per range
  int waves;       // pass 1: wavefronts of the grid, one range each (0: MCC_RANGE_BATCH-slot ranges)
  unsigned long long* errc;  // EPIC_ERRC_MCC_BOUND: collision probability above p_null
  // 1: the species holds no absorbed (NaN x) slot unless the push dropped a
  // non-finite particle (errc[EPIC_ERRC_NAN] > 0): walls that do not absorb,
  // no conductors, nothing uploaded or cleared since (SpeciesDev::maybe_dead)
  int no_dead;
};

__device__ __forceinline__ void rotate(double g, double chi, double eta, double st, double ct, double sp, double cp,
                                       double& gx, double& gy, double& gz) {
  const double sc = sin(chi), cc = cos(chi), se = sin(eta), ce = cos(eta);
  gx = g * (ct * cc - st * sc * ce);
  gy = g * (st * cp * cc + ct * cp * sc * ce - sp * sc * se);
  gz = g * (st * sp * cc + ct * sp * sc * ce + cp * sc * se);
}

// Box-Muller pair; the angle 2 pi u2 through sincospi(2 u2) (2 u2 in [0, 2):
// no large-argument reduction).
__device__ __forceinline__ double gauss(uint32_t a, uint32_t b, uint32_t c, uint32_t d, double* second) {
  const double u1 = 1.0 - u01_53(a, b), u2 = u01_53(c, d);
  const double m = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  *second = m * sn;
  return m * cs;
}

// The neutral partner of an ion collision: a Maxwellian at gas_temp_K from
// the particle's (id, step) streams 2 and 3.
template <typename P>
__device__ __forceinline__ void neutral_velocity(const MccArgs<P>& a, uint32_t pid, uint32_t step, double& vnx,
                                                 double& vny, double& vnz) {
  const u32x4 r2 = philox4x32_10({pid, step, RNG_MCC2, a.sp}, a.k0, a.k1);
  const u32x4 r3 = philox4x32_10({pid, step, RNG_MCC3, a.sp}, a.k0, a.k1);
  double g2nd, g3nd;
  vnx = a.gas_vth * gauss(r2.a, r2.b, r2.c, r2.d, &g2nd);
  vny = a.gas_vth * g2nd;
  vnz = a.gas_vth * gauss(r3.a, r3.b, r3.c, r3.d, &g3nd);
}

// One candidate of the null-collision method (R0 < P_null already drawn):
// the real/null decision and, for a real collision, its process (0-2
// electron-neutral elastic / excitation / ionisation, 3-4 ion-neutral
// elastic / charge exchange).  Returns -1 for a null collision or an
// absorbed slot.
// KIND: the species kind when fixed at compile time (0 electron, 1 ion), -1
// to read it from the arguments.
// DEAD: absorbed slots (NaN x) may exist and are skipped; without, x is not
// read (a quarter of the candidates' lines).
template <typename P, int KIND, bool DEAD>
__device__ __forceinline__ int mcc_null_test(const MccArgs<P>& a, uint32_t step, int64_t p) {
  // x, y, vx, vy at slot q (tiled FP32 storage), vz and id at p; every load
  // is issued before the first test (p is a live slot, so all are in bounds)
  const int64_t q = pslot<P>(p);
  const uint32_t pid = a.id[p];
  const P xp = DEAD ? a.x[q] : (P)0;
  const double vxe = a.vx[q], vye = a.vy[q], vze = a.vz[p];
  const u32x4 r = philox4x32_10({pid, step, RNG_MCC, a.sp}, a.k0, a.k1);
  if (DEAD && isnan(xp)) return -1;
  const double R1 = u01_53(r.c, r.d);
  if ((KIND >= 0 ? KIND : a.kind) == 0) {
    // electron - neutral (neutral at rest)
    const double v_sqr = (vxe * vxe + vye * vye + vze * vze) * a.vel_norm * a.vel_norm;
    const double velocity = sqrt(v_sqr);
    const double energy = (0.5 * ME * v_sqr) / eV;
    const MccTab::Pos at_e = a.tab.pos(energy);
    double s_el = (a.enable & 1) ? a.tab.at(0, at_e) : 0.0;
    double s_ex = (a.enable & 2) && energy > a.e_exc ? a.tab.at(1, at_e) : 0.0;
    double s_io = (a.enable & 4) && energy > a.e_ion ? a.tab.at(2, at_e) : 0.0;
    const double t0 = s_el, t1 = t0 + s_ex, t2 = t1 + s_io;
    const double nu = a.gas_density * t2 * velocity;
    const double pc = -expm1(-nu * a.dt);
    // a speed beyond the one nu_max covers: the particle collides with
    // p_null instead of pc (counted, as an undercount)
    if (pc > a.p_null) atomicAdd(&a.errc[EPIC_ERRC_MCC_BOUND], 1ull);
    if (R1 * a.p_null >= pc) return -1;  // null collision
    const u32x4 r2 = philox4x32_10({pid, step, RNG_MCC2, a.sp}, a.k0, a.k1);
    if (a.ref_selection) {
      // two independent draws (collision.cpp:117,125)
      if (u01_53(r2.a, r2.b) < (t0 / t2)) return 0;
      if (u01_53(r2.c, r2.d) < (t1 / t2)) return 1;
      return 2;
    }
    const double R = u01_53(r2.a, r2.b) * t2;
    return R < t0 ? 0 : (R < t1 ? 1 : 2);
  }
  // ion - neutral against a Maxwellian neutral
  double vnx, vny, vnz;
  neutral_velocity(a, pid, step, vnx, vny, vnz);
  const double gx = vxe - vnx, gy = vye - vny, gz = vze - vnz;
  const double g2 = gx * gx + gy * gy + gz * gz;
  const double g = sqrt(g2);
  const double energy = 0.5 * a.mass * g2 * a.vel_norm * a.vel_norm / eV;
  const MccTab::Pos at_e = a.tab.pos(energy);
  const double s_el = (a.enable & 8) ? a.tab.at(3, at_e) : 0.0;
  const double s_cx = (a.enable & 16) ? a.tab.at(4, at_e) : 0.0;
  const double nu_el = a.gas_density * s_el * g * a.vel_norm;
  const double nu_cx = a.gas_density * s_cx * g * a.vel_norm;
  const double nu = nu_el + nu_cx;
  const double pc = -expm1(-nu * a.dt);
  if (pc > a.p_null) atomicAdd(&a.errc[EPIC_ERRC_MCC_BOUND], 1ull);
  if (R1 * a.p_null >= pc) return -1;  // null collision
  // real collision: process by frequency share
  const double R = (R1 * a.p_null / pc) * nu;
  return R < nu_el ? 3 : 4;
}

// The kinematics of one real collision (slot p, process from pass 1).  An
// ionisation with ev >= 0 fills event ev (its rank in the chunk, ev[7], is
// pass 1's).
template <typename P>
__device__ __forceinline__ void mcc_collide(const MccArgs<P>& a, uint32_t step, int64_t p, int type, int evi,
                                            unsigned long long (&cnt)[EPIC_MCC_NPROC]) {
  const int64_t q = pslot<P>(p);
  const uint32_t pid = a.id[p];
  const double vxe = a.vx[q], vye = a.vy[q], vze = a.vz[p];
  if (a.kind == 0) {
    const u32x4 r3 = philox4x32_10({pid, step, RNG_MCC3, a.sp}, a.k0, a.k1);
    const double F1 = ME / (ME + a.gas_mass), F2 = a.gas_mass / (ME + a.gas_mass);
    double gx = vxe, gy = vye, gz = vze;
    double g = sqrt(gx * gx + gy * gy + gz * gz);
    const double wx = F1 * vxe, wy = F1 * vye, wz = F1 * vze;
    const double theta = (gx == 0) ? 0.5 * PI : atan2(sqrt(gy * gy + gz * gz), gx);
    const double phi = (gy == 0) ? (gz > 0 ? 0.5 * PI : -0.5 * PI) : atan2(gz, gy);
    const double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
    double chi, eta;
    const double Ra = u01_53(r3.a, r3.b), Rb = u01_53(r3.c, r3.d);
    if (type == 0) {
      chi = acos(1.0 - 2.0 * Ra);
      eta = 2 * PI * Rb;
    } else if (type == 1) {
      double en = 0.5 * ME * g * g * a.vel_norm * a.vel_norm;
      en = fabs(en - a.e_exc * eV);
      g = sqrt(2.0 * en / ME) / a.vel_norm;
      chi = acos(1.0 - 2.0 * Ra);
      eta = 2 * PI * Rb;
    } else {
      double en = 0.5 * ME * g * g * a.vel_norm * a.vel_norm;
      en = fabs(en - a.e_ion * eV);
      const double e_ej = 10.0 * tan(Ra * atan(en / eV / 20.0)) * eV;
      const double e_sc = fabs(en - e_ej);
      g = sqrt(2.0 * e_sc / ME) / a.vel_norm;
      const double g2 = sqrt(2.0 * e_ej / ME) / a.vel_norm;
      chi = acos(sqrt(e_sc / en));
      const double chi2 = acos(sqrt(e_ej / en));
      eta = 2 * PI * Rb;
      const double eta2 = eta + PI;
      double hx, hy, hz;
      rotate(g2, chi2, eta2, st, ct, sp, cp, hx, hy, hz);
      if (evi >= 0) {
        double* ev = a.events + (size_t)evi * EVW;
        ev[0] = (double)p;
        ev[1] = (double)a.x[q];
        ev[2] = (double)a.y[q];
        ev[3] = wx + F2 * hx;
        ev[4] = wy + F2 * hy;
        ev[5] = wz + F2 * hz;
        ev[6] = (double)pid;
      }
    }
    double ngx, ngy, ngz;
    rotate(g, chi, eta, st, ct, sp, cp, ngx, ngy, ngz);
    a.vx[q] = (P)(wx + F2 * ngx);
    a.vy[q] = (P)(wy + F2 * ngy);
    a.vz[p] = (P)(wz + F2 * ngz);
  } else {
    double vnx, vny, vnz;
    neutral_velocity(a, pid, step, vnx, vny, vnz);
    if (type == 3) {
      const double gx = vxe - vnx, gy = vye - vny, gz = vze - vnz;
      const double g = sqrt(gx * gx + gy * gy + gz * gz);
      const u32x4 r4 = philox4x32_10({pid, step, RNG_MCC_ION_SCATTER, a.sp}, a.k0, a.k1);
      const double M = a.gas_mass, m = a.mass;
      const double cx = (m * vxe + M * vnx) / (m + M), cy = (m * vye + M * vny) / (m + M),
                   cz = (m * vze + M * vnz) / (m + M);
      const double ct = 1.0 - 2.0 * u01_53(r4.a, r4.b), stt = sqrt(fmax(0.0, 1.0 - ct * ct));
      const double ph = 2.0 * M_PI * u01_53(r4.c, r4.d);
      const double f = M / (m + M);
      a.vx[q] = (P)(cx + f * g * stt * cos(ph));
      a.vy[q] = (P)(cy + f * g * stt * sin(ph));
      a.vz[p] = (P)(cz + f * g * ct);
    } else {
      a.vx[q] = (P)vnx;
      a.vy[q] = (P)vny;
      a.vz[p] = (P)vnz;
    }
  }
  cnt[type]++;
}

// Pass 1 walks a species in ranges of slots (a multiple of 64, at most
// MCC_RMAX), one wavefront per range at a time; a real collision's list
// entry packs its process (3 bits) and its event slot + 1 (29 bits).
static_assert(MCC_SUB == 512, "the selection loads 8 groups of 64 slots at a time");
static_assert(MCC_EVENT_MAX < (1 << 29) - 1, "event slot + 1 in 29 bits");
constexpr int MCC_RMAX = 3072;  // LDS candidate list of a wavefront (16-bit offsets)
static_assert(MCC_RMAX % 64 == 0 && MCC_RMAX <= 65536 && MCC_RMAX >= MCC_RANGE_BATCH, "range bounds");

__device__ __forceinline__ int lane_rank(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One null-test round: up to 64 candidates (cand[t0 + lane], t < todo; slot
// offsets from base), real collisions appended to the species' list with
// one atomic.  Electron ionisations reserve their event slots here, in slot
// order within the range, so each stored event gets its rank among the
// range's stored events (ev[7]; nstored counts them).
template <typename P, int KIND, bool DEAD>
__device__ __forceinline__ void mcc_null_round(const MccArgs<P>& a, uint32_t step, int64_t base, const uint16_t* cand,
                                               int t0, int todo, int& nstored) {
  const int lane = threadIdx.x & 63;
  const int kind = KIND >= 0 ? KIND : a.kind;
  const int t = t0 + lane;
  const int64_t p = base + (t < todo ? (int64_t)cand[t] : 0);
  const int type = t < todo ? mcc_null_test<P, KIND, DEAD>(a, step, p) : -1;
  const unsigned long long real = __ballot(type >= 0);
  if (!real) return;
  int evi = -1;
  if (kind == 0) {
    const unsigned long long ion = __ballot(type == 2);
    if (ion) {
      const int first = __builtin_ctzll(ion), k = (int)__popcll(ion);
      int e0 = 0;
      if (lane == first) e0 = atomicAdd(a.event_count, k);
      e0 = __shfl(e0, first, 64);
      const int i = lane_rank(ion);
      if (type == 2 && e0 + i < a.event_cap) {
        evi = e0 + i;
        a.events[(size_t)evi * EVW + 7] = (double)(nstored + i);
      }
      nstored += max(0, min(k, a.event_cap - e0));
    }
  }
  const int first = __builtin_ctzll(real);
  int b0 = 0;
  if (lane == first) b0 = atomicAdd(a.list_n, (int)__popcll(real));
  b0 = __shfl(b0, first, 64);
  if (type >= 0) a.list[b0 + lane_rank(real)] = make_uint2((uint32_t)p, (uint32_t)type | ((uint32_t)(evi + 1) << 3));
}

// Slots per selection range: n spread over `waves` ranges (rounded up to 64
// slots, at most MCC_RMAX), or MCC_RANGE_BATCH when waves is 0.  The append
// computes the same.
__device__ __forceinline__ int64_t mcc_range(int64_t n, int waves) {
  if (waves <= 0) return MCC_RANGE_BATCH;
  const int64_t per = (n + waves - 1) / waves;
  return min((int64_t)MCC_RMAX, max((int64_t)64, (per + 63) / 64 * 64));
}

// Pass 1 of a colliding species over range rg, in two phases.  Selection:
// every particle draws R0 from its (id, step) Philox2x32 stream, MCC_SUB
// slots at a time (the next MCC_SUB slots' ids load while these are drawn);
// those with R0 < P_null (a fraction ~nu_max dt) are candidates, listed in
// LDS (cand) in slot order.  Null test: the candidates in rounds of 64.
// chunk_ev[rg] gets the range's stored electron ionisation events: the birth
// append orders births by (range, rank), i.e. by parent slot, whatever order
// the list has.
template <typename P, int KIND, bool DEAD>
__device__ __forceinline__ void mcc_select_range(const MccArgs<P>& a, uint32_t step, int64_t rg, int64_t range,
                                                 int64_t n, uint16_t* cand) {
  const int lane = threadIdx.x & 63;
  const int kind = KIND >= 0 ? KIND : a.kind;
  const int64_t base = rg * range;
  const int64_t end = min(base + range, n);
  int cnt = 0;
  // loads and draws run on every lane (slots past the end read the range's
  // last slot and are not candidates): no branches, so the next slots' loads
  // stay in flight through the draws
  uint32_t nxt[8];
#pragma unroll
  for (int j = 0; j < 8; j++) nxt[j] = a.id[min(base + j * 64 + lane, end - 1)];
  for (int64_t sb = base; sb < end; sb += MCC_SUB) {
    uint32_t pid[8];
#pragma unroll
    for (int j = 0; j < 8; j++) pid[j] = nxt[j];
#pragma unroll
    for (int j = 0; j < 8; j++) nxt[j] = a.id[min(sb + MCC_SUB + j * 64 + lane, end - 1)];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int64_t p = sb + j * 64 + lane;
      // the key goes through an empty asm so that its ten round values are
      // formed by scalar adds here rather than kept live (and spilled to
      // VGPR lanes) across the loop
      uint32_t key = __builtin_amdgcn_readfirstlane(a.key_r0);
      asm volatile("" : "+s"(key));
      const uint2 r0 = philox2x32_10(pid[j], step, key);
      // u01_53(r0) < p_null, as an exact integer comparison
      const uint64_t m53 = ((uint64_t)r0.x << 21) ^ (uint64_t)(r0.y >> 11);
      const bool is_cand = p < end && m53 < a.r0_below;
      const unsigned long long m = __ballot(is_cand);
      if (is_cand) cand[cnt + lane_rank(m)] = (uint16_t)(p - base);
      cnt += (int)__popcll(m);
    }
  }
  wave_lds_sync();
  int nstored = 0;
  for (int t0 = 0; t0 < cnt; t0 += 64) mcc_null_round<P, KIND, DEAD>(a, step, base, cand, t0, cnt, nstored);
  if (kind == 0 && lane == 0) a.chunk_ev[rg] = nstored;
  wave_lds_sync();  // the next range overwrites the candidate list
}

template <typename P, int KIND, bool DEAD>
__device__ __forceinline__ void mcc_select_body(const MccArgs<P>& a, uint32_t step, int bid, int nblk,
                                                uint16_t* cand) {
  const int64_t n = a.d_n[0];
  const int64_t range = mcc_range(n, a.waves);
  const int wpb = blockDim.x >> 6;
  const int64_t nr = (n + range - 1) / range;
  for (int64_t rg = (int64_t)bid * wpb + (threadIdx.x >> 6); rg < nr; rg += (int64_t)nblk * wpb)
    mcc_select_range<P, KIND, DEAD>(a, step, rg, range, n, cand);
}

// Pass 2: the kinematics of the species' real collisions, one per lane over
// the list pass 1 built.  Each particle's draws are keyed by (id, step): the
// list's order changes nothing.
template <typename P>
__device__ __forceinline__ void mcc_body(const MccArgs<P>& a, uint32_t step, int bid, int nblk) {
  const int m = *a.list_n;
  unsigned long long cnt[EPIC_MCC_NPROC] = {0, 0, 0, 0, 0};
  for (int k = bid * (int)blockDim.x + (int)threadIdx.x; k < m; k += nblk * (int)blockDim.x) {
    const uint2 e = a.list[k];
    mcc_collide(a, step, (int64_t)e.x, (int)(e.y & 7u), (int)(e.y >> 3) - 1, cnt);
  }
#pragma unroll
  for (int q = 0; q < EPIC_MCC_NPROC; q++) {
    unsigned long long v = cnt[q];
    // wave reduction, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&a.counts[q], v);
  }
}

// A context's colliding species (electrons, ions) in one launch per pass:
// blockIdx.y selects the species.
template <typename P>
struct MccPair {
  MccArgs<P> a[2];
  int np;  // species in a (1 or 2)
};

// Pass 1 of one species, its kind fixed at compile time (the electron and the
// ion null tests need different registers: each launch runs at its own
// occupancy).  Every workgroup resident; wavefront w takes chunks w, w + W, ...
template <typename P, int KIND>
__global__ __launch_bounds__(MCC_BLOCK) void mcc_select_kernel(MccArgs<P> a) {
  __shared__ uint16_t cand[MCC_BLOCK / 64][MCC_RMAX];
  // the push counts every particle it drops as non-finite (a NaN slot); the
  // count is read after that push's launch, so it is complete here
  if (a.no_dead && __builtin_amdgcn_readfirstlane((int)(a.errc[EPIC_ERRC_NAN] != 0)) == 0)
    mcc_select_body<P, KIND, false>(a, a.step, blockIdx.x, gridDim.x, cand[threadIdx.x >> 6]);
  else
    mcc_select_body<P, KIND, true>(a, a.step, blockIdx.x, gridDim.x, cand[threadIdx.x >> 6]);
}

// OCC: workgroups per CU the register allocation is held to (1: none)
template <typename P, int OCC>
__global__ __launch_bounds__(MCC_BLOCK, OCC) void mcc_kernel(MccPair<P> s) {
  const MccArgs<P>& a = s.a[blockIdx.y];
  mcc_body(a, a.step, blockIdx.x, gridDim.x);
}

// Batched: blockIdx.y selects the (simulation, species) pass; the step comes
// from the simulation's device step counter.
template <typename P>
__global__ __launch_bounds__(MCC_BLOCK) void mcc_select_batch_kernel(const MccArgs<P>* args, const int* steps) {
  __shared__ uint16_t cand[MCC_BLOCK / 64][MCC_RMAX];
  const MccArgs<P>& a = args[blockIdx.y];
  mcc_select_body<P, -1, true>(a, (uint32_t)steps[a.slot], blockIdx.x, gridDim.x, cand[threadIdx.x >> 6]);
}

template <typename P>
__global__ __launch_bounds__(MCC_BLOCK, 2) void mcc_batch_kernel(const MccArgs<P>* args, const int* steps) {
  const MccArgs<P>& a = args[blockIdx.y];
  mcc_body(a, (uint32_t)steps[a.slot], blockIdx.x, gridDim.x);
}

// Append ionisation births in parent-slot order: one workgroup scans the
// per-chunk event counts the collision pass left; an event's rank is its
// chunk's offset plus its rank within the chunk (O(chunks + events), no
// bound on the events per step but the event list).  The new electron goes
// to n_e + rank and the new ion to n_i + rank, ids next_id + rank, and both
// are deposited.
constexpr int APPEND_BLOCK = 1024;

template <typename P>
struct AppendArgs {
  const double* events;
  int* event_count;
  int cap_events;
  P *ex, *ey, *evx, *evy, *evz;
  uint32_t* eid;
  int64_t* e_n;
  int64_t e_cap;
  P *ix, *iy, *ivx, *ivy, *ivz;
  uint32_t* iid;
  int64_t* i_n;
  int64_t i_cap;
  double ion_vth;
  uint32_t step, k0, k1;
  unsigned long long* Qe;  // null: births are not deposited
  unsigned long long* Qi;
  int nx, ny;
  double x0, y0, dx, dy;
  int fix_bits;
  unsigned long long* counts;
  unsigned long long* errc;
  int slot;  // batched runs: index of the simulation's step counter
  int* chunk_n;  // per selection range: its ionisation events (scanned in place)
  int waves;     // the selection grid's wavefronts (mcc_range)
  int births;    // the electrons ionise: append their events
  int* list_n;   // the context's real-collision list counts (electrons, ions)
};

template <typename P>
__device__ __forceinline__ void mcc_births(const AppendArgs<P>& A, uint32_t step, int* scan) {
  const double* events = A.events;
  int* chunk_n = A.chunk_n;
  int* event_count = A.event_count;
  const int cap_events = A.cap_events;
  P *ex = A.ex, *ey = A.ey, *evx = A.evx, *evy = A.evy, *evz = A.evz;
  uint32_t* eid = A.eid;
  int64_t* e_n = A.e_n;
  const int64_t e_cap = A.e_cap;
  P *ix = A.ix, *iy = A.iy, *ivx = A.ivx, *ivy = A.ivy, *ivz = A.ivz;
  uint32_t* iid = A.iid;
  int64_t* i_n = A.i_n;
  const int64_t i_cap = A.i_cap;
  const double ion_vth = A.ion_vth;
  const uint32_t k0 = A.k0, k1 = A.k1;
  unsigned long long *Qe = A.Qe, *Qi = A.Qi;
  const int nx = A.nx, ny = A.ny;
  const double x0 = A.x0, y0 = A.y0, dx = A.dx, dy = A.dy;
  const int fix_bits = A.fix_bits;
  unsigned long long* counts = A.counts;
  unsigned long long* errc = A.errc;
  const int E = min(*event_count, cap_events);
  const int64_t ne = e_n[0], ni = i_n[0];
  const uint32_t ide = (uint32_t)e_n[1], idi = (uint32_t)i_n[1];
  // exclusive scan of the per-range event counts the selection pass left in
  // chunk_n (range order = slot order): each thread a contiguous run of
  // ranges, then a scan of the run totals
  const int64_t range = mcc_range(ne, A.waves);
  const int nch = (int)((ne + range - 1) / range);
  const int per = (nch + (int)blockDim.x - 1) / (int)blockDim.x;
  const int c0 = min((int)threadIdx.x * per, nch), c1 = min(c0 + per, nch);
  int run = 0;
  for (int c = c0; c < c1; c++) run += chunk_n[c];
  scan[threadIdx.x] = run;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    const int v = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  int acc = scan[threadIdx.x] - run;
  for (int c = c0; c < c1; c++) {
    const int v = chunk_n[c];
    chunk_n[c] = acc;
    acc += v;
  }
  __syncthreads();
  int dropped = 0;
  for (int k = threadIdx.x; k < E; k += blockDim.x) {
    const double* ev = events + (size_t)k * EVW;
    const int rank = chunk_n[(int64_t)ev[0] / range] + (int)ev[7];
    const int64_t se = ne + rank, si = ni + rank;
    if (se >= e_cap || si >= i_cap) {
      dropped++;
      continue;
    }
    const double px = ev[1], py = ev[2];
    const int64_t qe = pslot<P>(se), qi = pslot<P>(si);
    ex[qe] = (P)px; ey[qe] = (P)py;
    evx[qe] = (P)ev[3]; evy[qe] = (P)ev[4]; evz[se] = (P)ev[5];
    eid[se] = ide + (uint32_t)rank;
    // new ion: Maxwellian at the ion species temperature (normalised)
    const uint32_t pid = (uint32_t)ev[6];
    const u32x4 r = philox4x32_10({pid, step, RNG_MCC_NEWION, 0x1u}, k0, k1);
    const u32x4 r2 = philox4x32_10({pid, step, RNG_MCC_NEWION2, 0x1u}, k0, k1);
    double s1, s2;
    const double g0 = gauss(r.a, r.b, r.c, r.d, &s1);
    const double g2 = gauss(r2.a, r2.b, r2.c, r2.d, &s2);
    ix[qi] = (P)px; iy[qi] = (P)py;
    ivx[qi] = (P)(ion_vth * g0); ivy[qi] = (P)(ion_vth * s1); ivz[si] = (P)(ion_vth * g2);
    iid[si] = idi + (uint32_t)rank;
    if (Qe) {
      const double lx = (px - x0) / dx, ly = (py - y0) / dy;
      const int i = (int)lx, j = (int)ly;
      if (i < 0 || i + 1 >= nx || j < 0 || j + 1 >= ny) {
        atomicAdd(&errc[EPIC_ERRC_DEPOSIT_OOB], 2ull);
        continue;
      }
      const double di = lx - i, dj = ly - j, sc = ldexp(1.0, fix_bits);
      const long long one = 1ll << fix_bits;
      const long long q00 = (long long)rint((1 - di) * (1 - dj) * sc);
      const long long q10 = (long long)rint(di * (1 - dj) * sc);
      const long long q01 = (long long)rint((1 - di) * dj * sc);
      const long long q11 = one - q00 - q10 - q01;
      const size_t b = (size_t)i * ny + j;
      unsigned long long* Qs[2] = {Qe, Qi};
      for (int w = 0; w < 2; w++) {
        atomicAdd(Qs[w] + b, (unsigned long long)q00);
        atomicAdd(Qs[w] + b + ny, (unsigned long long)q10);
        atomicAdd(Qs[w] + b + 1, (unsigned long long)q01);
        atomicAdd(Qs[w] + b + ny + 1, (unsigned long long)q11);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int ok = E;
    e_n[0] = min(ne + ok, e_cap);
    e_n[1] += ok;
    i_n[0] = min(ni + ok, i_cap);
    i_n[1] += ok;
    counts[EPIC_MCC_NPROC] += (unsigned long long)ok;
    const int over = *event_count - E;
    if (over > 0) atomicAdd(&errc[EPIC_ERRC_CAPACITY], (unsigned long long)over);
    *event_count = 0;
  }
  if (dropped) atomicAdd(&errc[EPIC_ERRC_CAPACITY], (unsigned long long)dropped);
}

// The call's last MCC kernel, one workgroup per context: the births when the
// electrons ionise, then the real-collision list counts are cleared for the
// next call (the collision pass before it has read them).
template <typename P>
__device__ __forceinline__ void mcc_finish_body(const AppendArgs<P>& A, uint32_t step, int* scan) {
  if (A.births) mcc_births(A, step, scan);
  if (threadIdx.x == 0) {
    A.list_n[0] = 0;
    A.list_n[1] = 0;
  }
}

template <typename P>
__global__ __launch_bounds__(APPEND_BLOCK) void mcc_append_kernel(AppendArgs<P> A) {
  __shared__ int scan[APPEND_BLOCK];
  mcc_finish_body(A, A.step, scan);
}

// Batched: one workgroup per simulation with ionisation.
template <typename P>
__global__ __launch_bounds__(APPEND_BLOCK) void mcc_append_batch_kernel(const AppendArgs<P>* args, const int* steps) {
  __shared__ int scan[APPEND_BLOCK];
  const AppendArgs<P>& A = args[blockIdx.x];
  mcc_finish_body(A, (uint32_t)steps[A.slot], scan);
}

// The candidate draws' Philox2x32 key of a (seed, species): the first word of
// Philox4x32-10 at counter (species, 0, RNG_MCC_KEY, 0) under the seed, on the
// host (a counter no device draw uses).
static uint32_t mcc_r0_key(uint32_t k0, uint32_t k1, uint32_t sp) {
  uint32_t c[4] = {sp, 0u, RNG_MCC_KEY, 0u};
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c[0];
}

// Arguments of one species' collision pass (kind 0 electron, 1 ion).
template <typename P>
static MccArgs<P> mcc_args(Ctx* c, int s, int kind, int step) {
  MccState& M = c->mcc;
  SpeciesDev& S = c->sp[s];
  const epic_domain_desc& d = c->dom;
  MccArgs<P> a{};
  a.x = (P*)S.x; a.y = (P*)S.y; a.vx = (P*)S.vx; a.vy = (P*)S.vy; a.vz = (P*)S.vz;
  a.id = S.id; a.d_n = S.d_n;
  a.tab.sigma = M.sigma; a.tab.n = M.d.ntab;
  a.tab.loge0 = log(M.d.e_min);
  a.tab.inv_dloge = (M.d.ntab - 1) / (log(M.d.e_max) - log(M.d.e_min));
  a.kind = kind;
  unsigned en = 0;
  for (int p = 0; p < EPIC_MCC_NPROC; p++)
    if (M.d.enable[p]) en |= 1u << p;
  a.enable = en;
  a.ref_selection = M.d.ref_selection;
  const double dt = d.DT / d.W;  // seconds (collision.cpp:201)
  a.nu_max = kind == 0 ? M.nu_max_e : M.nu_max_i;
  a.p_null = -expm1(-a.nu_max * dt);
  a.r0_below = (uint64_t)std::ceil(std::ldexp(a.p_null, 53));
  a.dt = dt;
  a.vel_norm = d.vel_norm;
  a.mass = S.desc.mass;
  a.gas_mass = M.d.gas_mass;
  a.gas_density = M.d.gas_density;
  a.gas_vth = M.d.gas_temp_K > 0 ? sqrt(K_b * M.d.gas_temp_K / M.d.gas_mass) / d.vel_norm : 0.0;
  a.e_exc = M.d.e_exc_th;
  a.e_ion = M.d.e_ion_th;
  a.step = (uint32_t)step;
  a.sp = (uint32_t)s;
  a.k0 = (uint32_t)d.seed;
  a.k1 = (uint32_t)(d.seed >> 32);
  a.key_r0 = mcc_r0_key(a.k0, a.k1, a.sp);
  a.counts = M.counts;
  a.events = M.events;
  a.event_count = M.event_count;
  a.event_cap = M.event_cap;
  a.list = kind == 0 ? M.list : M.list + M.list_len_e;
  a.list_n = M.list_n + kind;
  a.chunk_ev = kind == 0 ? M.chunk_ev : nullptr;
  a.errc = c->errc;
  a.no_dead = d.bc != EPIC_BC_OPEN && c->cond.desc.empty() && !S.maybe_dead;
  return a;
}

// Arguments of the ionisation-birth append (electron species -> ion species).
template <typename P>
static AppendArgs<P> append_args(Ctx* c, int step, bool births) {
  MccState& M = c->mcc;
  const epic_domain_desc& d = c->dom;
  AppendArgs<P> A{};
  A.step = (uint32_t)step;
  A.list_n = M.list_n;
  A.births = births ? 1 : 0;
  if (!births) return A;
  SpeciesDev& S = c->sp[M.d.electron_sp];
  SpeciesDev& I = c->sp[M.d.ion_sp];
  A.events = M.events;
  A.event_count = M.event_count;
  A.cap_events = M.event_cap;
  A.chunk_n = M.chunk_ev;  // the electrons' per-range event counts (pass 1)
  A.ex = (P*)S.x; A.ey = (P*)S.y; A.evx = (P*)S.vx; A.evy = (P*)S.vy; A.evz = (P*)S.vz;
  A.eid = S.id; A.e_n = S.d_n; A.e_cap = S.cap;
  A.ix = (P*)I.x; A.iy = (P*)I.y; A.ivx = (P*)I.vx; A.ivy = (P*)I.vy; A.ivz = (P*)I.vz;
  A.iid = I.id; A.i_n = I.d_n; A.i_cap = I.cap;
  A.ion_vth = sqrt(I.desc.temp * QE / I.desc.mass) / d.vel_norm;
  A.k0 = (uint32_t)d.seed;
  A.k1 = (uint32_t)(d.seed >> 32);
  A.Qe = S.Q;
  A.Qi = I.Q;
  A.nx = d.nx; A.ny = d.ny;
  A.x0 = d.x0; A.y0 = d.y0; A.dx = d.dx; A.dy = d.dy;
  A.fix_bits = c->fix_bits;
  A.counts = M.counts;
  A.errc = c->errc;
  return A;
}

// One context's collision step: pass 1 of each colliding species, pass 2 of
// both in one launch, then the finish kernel (births, list counts cleared).
template <typename P>
static hipError_t mcc_launch(Ctx* c, int step) {
  const epic_mcc_desc& d = c->mcc.d;
  MccPair<P> s{};
  int np = 0;
  bool births = false;
  if (d.electron_sp >= 0) {
    const MccArgs<P> a = mcc_args<P>(c, d.electron_sp, 0, step);
    if (a.nu_max > 0) {
      s.a[np++] = a;
      births = (a.enable & 4) != 0;
    }
  }
  if (d.ion_sp >= 0 && (d.enable[EPIC_MCC_I_ELASTIC] || d.enable[EPIC_MCC_I_CX])) {
    const MccArgs<P> a = mcc_args<P>(c, d.ion_sp, 1, step);
    if (a.nu_max > 0) {
      s.a[np++] = a;
    }
  }
  if (np == 0) return hipSuccess;
  s.np = np;
  // pass 1 per species: every workgroup resident, each wavefront one range
  // of the species (EPIC_MCC_SEL_GRID=0: ranges of MCC_RANGE_BATCH slots, one
  // workgroup per four), then pass 2 of both species
  static int sel_per_cu[2][2] = {{0, 0}, {0, 0}};
  int waves_e = 0;
  for (int k = 0; k < np; k++) {
    MccArgs<P>& a = s.a[k];
    const void* kern = a.kind == 0 ? (const void*)mcc_select_kernel<P, 0> : (const void*)mcc_select_kernel<P, 1>;
    int& per_cu = sel_per_cu[sizeof(P) == 8][a.kind];
    if (per_cu == 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, MCC_BLOCK, 0) != hipSuccess ||
                        per_cu < 1))
      per_cu = 1;
    const int64_t cap = std::max<int64_t>(c->sp[a.sp].cap, 1);
    int g1;
    if (c->mcc_sel_grid == 0) {
      a.waves = 0;
      g1 = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cu * 16,
                                                       ((cap + MCC_RANGE_BATCH - 1) / MCC_RANGE_BATCH + 3) / 4));
    } else {
      g1 = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cu * per_cu, (cap + 255) / 256));
      a.waves = g1 * (MCC_BLOCK / 64);
    }
    if (a.kind == 0) {
      waves_e = a.waves;
      hipLaunchKernelGGL((mcc_select_kernel<P, 0>), dim3(g1), dim3(MCC_BLOCK), 0, c->stream, a);
    } else {
      hipLaunchKernelGGL((mcc_select_kernel<P, 1>), dim3(g1), dim3(MCC_BLOCK), 0, c->stream, a);
    }
    EPIC_HIP_TRY(hipGetLastError());
  }
  if (c->mcc_occ == 1)
    hipLaunchKernelGGL((mcc_kernel<P, 1>), dim3(c->num_cu, np), dim3(MCC_BLOCK), 0, c->stream, s);
  else
    hipLaunchKernelGGL((mcc_kernel<P, 2>), dim3(c->num_cu * 2, np), dim3(MCC_BLOCK), 0, c->stream, s);
  EPIC_HIP_TRY(hipGetLastError());
  AppendArgs<P> A = append_args<P>(c, step, births);
  A.waves = waves_e;
  if (!c->deposit_valid) A.Qe = A.Qi = nullptr;
  hipLaunchKernelGGL(mcc_append_kernel<P>, dim3(1), dim3(APPEND_BLOCK), 0, c->stream, A);
  EPIC_HIP_TRY(hipGetLastError());
  return hipSuccess;
}

size_t mcc_args_bytes() { return sizeof(MccArgs<float>); }
size_t append_args_bytes() { return sizeof(AppendArgs<float>); }

void mcc_batch_collect(Ctx* c, int slot, std::vector<unsigned char>& args, std::vector<unsigned char>& app) {
  if (!c->mcc.on || c->dom.precision != EPIC_FP32) return;
  const epic_mcc_desc& d = c->mcc.d;
  auto put = [](std::vector<unsigned char>& v, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    v.insert(v.end(), b, b + n);
  };
  bool any = false, births = false;
  if (d.electron_sp >= 0) {
    MccArgs<float> a = mcc_args<float>(c, d.electron_sp, 0, 0);
    a.slot = slot;
    a.waves = 0;
    if (a.nu_max > 0) {
      put(args, &a, sizeof(a));
      any = true;
      births = (a.enable & 4) != 0;
    }
  }
  if (d.ion_sp >= 0 && (d.enable[EPIC_MCC_I_ELASTIC] || d.enable[EPIC_MCC_I_CX])) {
    MccArgs<float> a = mcc_args<float>(c, d.ion_sp, 1, 0);
    a.slot = slot;
    a.waves = 0;
    if (a.nu_max > 0) {
      put(args, &a, sizeof(a));
      any = true;
    }
  }
  if (any) {
    // every simulation with a collision pass gets a finish workgroup (its
    // list counts are cleared there)
    AppendArgs<float> A = append_args<float>(c, 0, births);
    A.slot = slot;
    A.waves = 0;
    put(app, &A, sizeof(A));
  }
}

// All collision passes of a batch in one launch per pass (blockIdx.y =
// pass), then every simulation's finish workgroup in one launch.
hipError_t launch_mcc_batch(hipStream_t st, int num_cu, const void* d_args, int nargs, const void* d_app, int napp,
                            const int* d_steps) {
  if (nargs > 0) {
    const int gx = std::max(4, std::min(num_cu * 8, (num_cu * 32 + nargs - 1) / nargs));
    hipLaunchKernelGGL(mcc_select_batch_kernel<float>, dim3(gx, nargs), dim3(MCC_BLOCK), 0, st,
                       (const MccArgs<float>*)d_args, d_steps);
    EPIC_HIP_TRY(hipGetLastError());
    const int g2 = std::max(1, std::min(num_cu * 2, (num_cu * 4 + nargs - 1) / nargs));
    hipLaunchKernelGGL(mcc_batch_kernel<float>, dim3(g2, nargs), dim3(MCC_BLOCK), 0, st,
                       (const MccArgs<float>*)d_args, d_steps);
    EPIC_HIP_TRY(hipGetLastError());
  }
  if (napp > 0) {
    hipLaunchKernelGGL(mcc_append_batch_kernel<float>, dim3(napp), dim3(APPEND_BLOCK), 0, st,
                       (const AppendArgs<float>*)d_app, d_steps);
    EPIC_HIP_TRY(hipGetLastError());
  }
  return hipSuccess;
}

hipError_t launch_mcc(Ctx* c, int step) {
  if (!c->mcc.on) return hipSuccess;
  return c->dom.precision == EPIC_FP64 ? mcc_launch<double>(c, step) : mcc_launch<float>(c, step);
}

// nu_max over the table for a particle of mass m (s^-1), 5% margin for the
// interpolation between table points and for ion-neutral relative speeds.
double mcc_nu_max(const epic_mcc_desc& d, const std::vector<double>& sig, double m, bool electron) {
  double best = 0;
  const int n = d.ntab;
  const double l0 = log(d.e_min), l1 = log(d.e_max);
  for (int k = 0; k < n; k++) {
    const double e = exp(l0 + (l1 - l0) * k / (n - 1));
    double s = 0;
    if (electron) {
      for (int p = 0; p < 3; p++)
        if (d.enable[p]) s += sig[(size_t)p * n + k];
    } else {
      for (int p = 3; p < 5; p++)
        if (d.enable[p]) s += sig[(size_t)p * n + k];
    }
    const double v = sqrt(2.0 * e * eV / m);
    best = std::max(best, d.gas_density * s * v);
  }
  // above e_max the cross section is held: nu keeps growing with v; cover
  // speeds up to 4x the table's top speed
  double stop = 0;
  for (int p = electron ? 0 : 3; p < (electron ? 3 : 5); p++)
    if (d.enable[p]) stop += sig[(size_t)p * n + n - 1];
  best = std::max(best, d.gas_density * stop * sqrt(2.0 * d.e_max * eV / m) * 4.0);
  return best * 1.05;
}

}  // namespace epic
