"""Null-collision MCC on the GPU against the reference's direct method.

The oracle restates CollisionHandler::handle_collisions (collision.cpp:190-
207) with its 5,000,000-point tables and mt19937 draws; the GPU runs the
null-collision scheme with Philox streams on log-spaced tables of the same
analytic cross sections.  Per-process counts must agree within 3 sigma
(sigma^2 = sum of the two Poisson counts); energy removed by thresholds is
checked against the electrons' kinetic-energy drop; runs are repeatable."""
import ctypes
import math

import numpy as np
import pytest

from epic2d_mcc_amd import _lib, config, xsec
from oracle_api import OracleSim, oracle

pytestmark = pytest.mark.gpu

ME = config.ME
GAS = 2e21


def _setup(N, Te_load, nsp_ion_N=1):
    specs = [config.SpeciesSpec("electron", ME, N, 1.0, -1, 1.0),
             config.SpeciesSpec("ion", xsec.M_AR, nsp_ion_N, 0.026, 1, 0.0)]
    st = config.derive(16, 16, 0.0, specs, 1e15, 0.05, bc="pbc", capacity_factor=1.5)
    st.species[1].capacity = int(N * 0.6) + 16
    return st


def _electrons(st, N, Te, seed):
    rng = np.random.default_rng(seed)
    d = st.domain
    sig = np.sqrt(Te * config.QE / ME) / d.vel_norm
    return (rng.random(N) * 16, rng.random(N) * 16, rng.normal(0, sig, N), rng.normal(0, sig, N),
            rng.normal(0, sig, N))


def _ke_ev(v, vel_norm):
    return float(0.5 * ME * (v[2] ** 2 + v[3] ** 2 + v[4] ** 2).sum() * vel_norm ** 2 / config.eV)


def test_counts_match_reference_direct_method():
    print("[mcc] counts vs direct method", flush=True)
    N, K = 200000, 150
    st = _setup(N, 8.0)
    d = st.domain
    e = _electrons(st, N, 8.0, 1)
    # GPU: null collision, reference analytic e-Ar tables
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, *e)
    ctx.upload(1, [1.0], [1.0], [0.0], [0.0], [0.0])
    ctx.mcc_setup(xsec.mcc_desc(xsec.reference_e_ar(), gas_density=GAS, gas_temp_K=0.0,
                                enable=(1, 1, 1, 0, 0)))
    for k in range(K):
        ctx.collide(k)
    g = ctx.mcc_stats()["count"]
    # oracle: the reference's direct method
    o = OracleSim(d, st.species, gas=(1, 1, 1, GAS))
    o.set_particles(0, *e)
    o.set_particles(1, [1.0], [1.0], [0.0], [0.0], [0.0])
    oracle().orc_mcc_build_tables(o.h)
    import ctypes
    cnt = (ctypes.c_long * 3)(0, 0, 0)
    el = ctypes.c_double()
    for k in range(K):
        oracle().orc_collide_direct(o.h, 0, 1, cnt, ctypes.byref(el))
    c = list(cnt)
    print(f"  gpu {g[:3]}  oracle {c}")
    for p in range(3):
        sigma = np.sqrt(g[p] + c[p])
        assert abs(g[p] - c[p]) <= 3 * sigma + 1, (p, g[p], c[p])
    assert g[0] > 10000 and g[1] > 500 and g[2] > 200
    # births: one electron and one ion per ionisation
    assert ctx.count(0) == N + g[2]
    assert ctx.count(1) == 1 + g[2]


def _u53(a, b):
    return float(((int(a) << 21) ^ (int(b) >> 11)) * 2.0 ** -53)


def _philox4(ctr, key):
    C = (ctypes.c_uint32 * 4)(*[int(v) & 0xffffffff for v in ctr])
    K = (ctypes.c_uint32 * 2)(*[int(v) & 0xffffffff for v in key])
    O = (ctypes.c_uint32 * 4)()
    oracle().orc_philox4x32_10(C, K, O)
    return list(O)


def _philox2(ctr, key):
    C = (ctypes.c_uint32 * 2)(*[int(v) & 0xffffffff for v in ctr])
    O = (ctypes.c_uint32 * 2)()
    oracle().orc_philox2x32_10(C, int(key) & 0xffffffff, O)
    return list(O)


def test_collision_selection_is_the_keyed_draws():
    """One collision step, elastic only: particle id collides exactly when
    R0 < P_null and R1 P_null < 1 - exp(-nu(E) dt), with R0 from Philox2x32-10
    at counter (id, step) under the key Philox4x32-10 gives at counter
    (species, 0, 8, 0) under the seed, R1 from Philox4x32-10 at counter (id,
    step, 4, species), nu(E) = n_g sigma(E) v from the log-spaced table and
    P_null = 1 - exp(-nu_max dt).  The draws come from the oracle's Philox
    restatements (pinned to the Random123 known answers); the set of particles
    whose velocity changed must equal the predicted set.  Particles within
    1e-12 (relative) of either threshold are left out of the comparison."""
    print("[mcc] exact selection from the keyed draws", flush=True)
    N, step, gas = 30000, 7, 2e22
    st = _setup(N, 8.0)
    d = st.domain
    e = _electrons(st, N, 8.0, 11)
    table = xsec.synthetic_ar()
    e_min, e_max, tab = table
    m = xsec.mcc_desc(table, gas_density=gas, gas_temp_K=0.0, enable=(1, 0, 0, 0, 0))
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, *e)
    ctx.upload(1, [1.0], [1.0], [0.0], [0.0], [0.0])
    ctx.mcc_setup(m)
    x0, y0, vx0, vy0, vz0, ids = ctx.download(0)
    ctx.collide(step)
    _, _, vx1, vy1, vz1, ids1 = ctx.download(0)
    ctx.close()
    assert np.array_equal(ids, ids1)
    hit = (vx1 != vx0) | (vy1 != vy0) | (vz1 != vz0)
    # P_null from nu_max over the table (and 4x the top speed), 5% margin
    n = tab.shape[1]
    l0, l1 = math.log(e_min), math.log(e_max)
    best = 0.0
    for k in range(n):
        en = math.exp(l0 + (l1 - l0) * k / (n - 1))
        best = max(best, gas * tab[0][k] * math.sqrt(2.0 * en * config.eV / ME))
    best = max(best, gas * tab[0][n - 1] * math.sqrt(2.0 * e_max * config.eV / ME) * 4.0)
    dt = d.DT / d.W
    p_null = -math.expm1(-best * 1.05 * dt)
    assert 0.1 < p_null < 0.6
    k0, k1 = d.seed & 0xffffffff, d.seed >> 32
    key = _philox4([0, 0, 8, 0], [k0, k1])[0]
    loge0, inv = math.log(e_min), (n - 1) / (math.log(e_max) - math.log(e_min))
    want = np.zeros(len(ids), dtype=bool)
    keep = np.ones(len(ids), dtype=bool)
    for i, pid in enumerate(ids):
        r0 = _philox2([pid, step], key)
        R0 = _u53(r0[0], r0[1])
        if abs(R0 - p_null) < 1e-12 * p_null:
            keep[i] = False
        if R0 >= p_null:
            continue
        r = _philox4([pid, step, 4, 0], [k0, k1])
        R1 = _u53(r[2], r[3])
        v_sqr = (vx0[i] ** 2 + vy0[i] ** 2 + vz0[i] ** 2) * d.vel_norm ** 2
        en = 0.5 * ME * v_sqr / config.eV
        t = (math.log(en) - loge0) * inv
        if t <= 0:
            sig = tab[0][0]
        elif t >= n - 1:
            sig = tab[0][n - 1]
        else:
            kk = int(t)
            f = t - kk
            sig = tab[0][kk] * (1 - f) + tab[0][kk + 1] * f
        pc = -math.expm1(-gas * sig * math.sqrt(v_sqr) * dt)
        if abs(R1 * p_null - pc) < 1e-12 * pc:
            keep[i] = False
        want[i] = R1 * p_null < pc
    print(f"  p_null {p_null:.3f}: {int(want.sum())} predicted, {int(hit.sum())} collided, "
          f"{int((~keep).sum())} at a threshold")
    assert want.sum() > 1000
    assert (~keep).sum() < 5
    assert np.array_equal(hit[keep], want[keep])


def test_energy_removed_equals_thresholds():
    """Electron KE drop = 11.5 eV per excitation + 15.8 eV per ionisation (the
    ejected electron keeps its share), up to the 2 m_e/M elastic recoil."""
    print("[mcc] energy bookkeeping", flush=True)
    N, K = 100000, 100
    st = _setup(N, 8.0)
    d = st.domain
    e = _electrons(st, N, 8.0, 2)
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, *e)
    ctx.upload(1, [1.0], [1.0], [0.0], [0.0], [0.0])
    ctx.mcc_setup(xsec.mcc_desc(xsec.reference_e_ar(), gas_density=GAS, gas_temp_K=0.0, enable=(1, 1, 1, 0, 0)))
    ke0 = _ke_ev(e, d.vel_norm)
    for k in range(K):
        ctx.collide(k)
    after = ctx.download(0)
    ke1 = _ke_ev(after, d.vel_norm)
    st_ = ctx.mcc_stats()
    loss = st_["eloss_ev"]
    print(f"  KE drop {ke0 - ke1:.1f} eV, thresholds {loss:.1f} eV, counts {st_['count'][:3]}")
    assert loss > 0
    # elastic recoil: ~2 m_e/M of the colliding energy per elastic collision
    recoil = 2 * ME / xsec.M_AR * st_["count"][0] * 8.0 * 1.5
    assert abs((ke0 - ke1) - loss) <= 0.01 * loss + 3 * recoil


def test_mcc_is_deterministic():
    N = 50000
    st = _setup(N, 8.0)
    e = _electrons(st, N, 8.0, 3)
    outs = []
    for rep in range(2):
        ctx = _lib.Context(st.domain, st.species)
        ctx.upload(0, *e)
        ctx.upload(1, [1.0], [1.0], [0.0], [0.0], [0.0])
        ctx.mcc_setup(xsec.mcc_desc(xsec.reference_e_ar(), gas_density=GAS, gas_temp_K=0.0,
                                    enable=(1, 1, 1, 0, 0)))
        for k in range(30):
            ctx.collide(k)
        outs.append((ctx.download(0), ctx.mcc_stats()))
        ctx.close()
    for k in range(6):
        assert np.array_equal(outs[0][0][k], outs[1][0][k])
    assert outs[0][1] == outs[1][1]


def test_ion_neutral_rates_match_monte_carlo_expectation():
    """Ar+ - Ar elastic + charge exchange against 300 K neutrals: counts per
    step against n_g <sigma(E_rel) g> dt estimated by numpy sampling."""
    print("[mcc] ion-neutral", flush=True)
    N, K = 200000, 50
    specs = [config.SpeciesSpec("electron", ME, 16, 1.0, -1, 1.0),
             config.SpeciesSpec("ion", xsec.M_AR, N, 0.026, 1, 0.0)]
    st = config.derive(16, 16, 0.0, specs, 1e15, 0.05, bc="pbc")
    d = st.domain
    rng = np.random.default_rng(4)
    kT = 0.026 * config.QE
    sig_i = np.sqrt(kT / xsec.M_AR)
    vi = rng.normal(0, sig_i, (N, 3))
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, np.full(16, 1.0), np.full(16, 1.0), np.zeros(16), np.zeros(16), np.zeros(16))
    ctx.upload(1, rng.random(N) * 16, rng.random(N) * 16, vi[:, 0] / d.vel_norm, vi[:, 1] / d.vel_norm,
               vi[:, 2] / d.vel_norm)
    tab = xsec.synthetic_ar()
    ctx.mcc_setup(xsec.mcc_desc(tab, electron_sp=-1, ion_sp=1, gas_density=1e21, gas_temp_K=300.0,
                                enable=(0, 0, 0, 1, 1)))
    for k in range(K):
        ctx.collide(k)
    g = ctx.mcc_stats()["count"]
    # expectation: ions stay thermal at ~300 K (CX and elastic with 300 K gas)
    vn = rng.normal(0, np.sqrt(config.K_b * 300.0 / xsec.M_AR), (N, 3))
    gr = np.linalg.norm(vi - vn, axis=1)
    E = 0.5 * xsec.M_AR * gr ** 2 / config.eV
    e_min, e_max, T = tab
    lg = np.log(xsec.log_grid(e_min, e_max, T.shape[1]))
    s_el = np.interp(np.log(np.clip(E, e_min, e_max)), lg, T[_lib.MCC_I_ELASTIC])
    s_cx = np.interp(np.log(np.clip(E, e_min, e_max)), lg, T[_lib.MCC_I_CX])
    dt = d.DT / d.W
    exp_el = (1e21 * s_el * gr * dt).sum() * K
    exp_cx = (1e21 * s_cx * gr * dt).sum() * K
    print(f"  gpu elastic {g[3]} cx {g[4]}; expected {exp_el:.0f} {exp_cx:.0f}")
    assert abs(g[3] - exp_el) < 4 * np.sqrt(exp_el) + 0.02 * exp_el
    assert abs(g[4] - exp_cx) < 4 * np.sqrt(exp_cx) + 0.02 * exp_cx


def test_mcc_device_bytes_match_estimate():
    """epic_mcc_setup adds exactly what epic_mcc_estimate_bytes states
    (tables, event list, candidate lists sized by species capacity); a
    repeated setup replaces rather than adds."""
    st = _setup(5000, 8.0)
    ctx = _lib.Context(st.domain, st.species)
    base = ctx.device_bytes()
    m = xsec.mcc_desc(xsec.reference_e_ar(), gas_density=GAS, gas_temp_K=0.0, enable=(1, 1, 1, 0, 0))
    ctx.mcc_setup(m)
    want = _lib.mcc_estimate_bytes(m, st.species)
    assert ctx.device_bytes() - base == want
    ctx.mcc_setup(m)
    assert ctx.device_bytes() - base == want
    ctx.close()


def test_candidate_lists_independent_of_layout():
    """The two-pass collision step (candidate selection and null test per
    512-slot chunk, then one real collision per lane) gives each particle
    the outcome of its own (id, step) draws: the same particles, reordered in
    memory by the cell sort, end with the same velocities."""
    print("[mcc] layout independence", flush=True)
    N = 70001  # a partial last chunk
    st = _setup(N, 8.0)
    e = _electrons(st, N, 8.0, 3)
    m = xsec.mcc_desc(xsec.reference_e_ar(), gas_density=GAS, gas_temp_K=0.0, enable=(1, 1, 0, 0, 0))
    outs = []
    for sort in (False, True):
        ctx = _lib.Context(st.domain, st.species)
        ctx.upload(0, *e)
        ctx.upload(1, [1.0], [1.0], [0.0], [0.0], [0.0])
        if sort:
            ctx.sort(0)  # reorders the slots by cell; ids travel with the particles
        ctx.mcc_setup(m)
        for k in range(20):
            ctx.collide(k)
        x, y, vx, vy, vz, ids = ctx.download(0)
        srt = np.argsort(ids)
        outs.append((vx[srt], vy[srt], vz[srt], ctx.mcc_stats()["count"].copy()))
        ctx.close()
    for k in range(3):
        assert np.array_equal(outs[0][k], outs[1][k])
    assert np.array_equal(outs[0][3], outs[1][3])
    assert outs[0][3][0] > 0


def test_coupled_loop_matches_direct_method_1000_steps():
    """The whole loop with collisions on both sides, 1000 steps: periodic
    16x16 cells, 25,600 electrons loaded at 8 eV and 25,600 static Ar+
    (ionfixed, main.cpp:419), argon at 2e21 m^-3 with the reference's e-Ar
    cross sections, FP64, PCG.  Each step scatters, solves, pushes the
    electrons and collides them (main.cpp:395-423); the electrons cool by
    excitation and ionisation while ionisation adds electrons and ions.  The
    GPU (null collision, Philox per (id, step)) and the oracle (the
    reference's direct method on its global mt19937) draw different numbers,
    so the runs agree in distribution: per-process counts and the energy
    removed by thresholds within 3 sigma (sigma^2 = the sum of both runs'
    Poisson variances), and the births equal the ionisations on each side."""
    print("[mcc] coupled loop, 1000 steps", flush=True)
    N, K = 25600, 1000
    specs = [config.SpeciesSpec("electron", ME, N, 1.0, -1, 1.0),
             config.SpeciesSpec("ion", xsec.M_AR, N, 0.026, 1, 0.0)]
    st = config.derive(16, 16, 0.0, specs, 1e15, 0.05, bc="pbc", tolerance=1e-8, tol_relative=0, ionfixed=1,
                       capacity_factor=1.5)
    d = st.domain
    rng = np.random.default_rng(21)
    sig = np.sqrt(8.0 * config.QE / ME) / d.vel_norm
    e = (rng.random(N) * 16, rng.random(N) * 16, rng.normal(0, sig, N), rng.normal(0, sig, N),
         rng.normal(0, sig, N))
    i = (rng.random(N) * 16, rng.random(N) * 16, np.zeros(N), np.zeros(N), np.zeros(N))
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, *e)
    ctx.upload(1, *i)
    ctx.mcc_setup(xsec.mcc_desc(xsec.reference_e_ar(), gas_density=GAS, gas_temp_K=0.0, enable=(1, 1, 1, 0, 0)))
    ctx.initial_field()
    for ts in range(K):
        ctx.step(ts)
    gs = ctx.mcc_stats()
    g = np.array(gs["count"][:3], dtype=np.float64)
    g_loss = gs["eloss_ev"]
    n_e, n_i = ctx.count(0), ctx.count(1)
    assert not ctx.errors().any(), ctx.errors()
    ctx.close()
    o = OracleSim(d, st.species, gas=(1, 1, 1, GAS))
    o.build_mcc_tables()
    o.set_particles(0, *e)
    o.set_particles(1, *i)
    o.initial_field()
    c = np.zeros(3)
    c_loss = 0.0
    for ts in range(K):
        o.step(pushed=[1, 0])
        cnt, loss = o.collide_direct()
        c += cnt
        c_loss += loss
    print(f"  gpu {g.astype(int).tolist()} loss {g_loss:.0f} eV; oracle {c.astype(int).tolist()} loss {c_loss:.0f} eV",
          flush=True)
    for p in range(3):
        assert abs(g[p] - c[p]) <= 3 * np.sqrt(g[p] + c[p]), (p, g[p], c[p])
    var_loss = 11.5 ** 2 * (g[1] + c[1]) + 15.8 ** 2 * (g[2] + c[2])
    assert abs(g_loss - c_loss) <= 3 * np.sqrt(var_loss)
    assert g_loss == pytest.approx(11.5 * g[1] + 15.8 * g[2], rel=1e-9)
    assert n_e == N + g[2] and n_i == N + g[2]
    assert o.o.orc_count(o.h, 0) == N + c[2] and o.o.orc_count(o.h, 1) == N + c[2]
    assert c[1] > 1000 and c[2] > 500


def _ion_only(N, gas_temp_K, enable, v3):
    specs = [config.SpeciesSpec("electron", ME, 16, 1.0, -1, 1.0),
             config.SpeciesSpec("ion", xsec.M_AR, N, 0.026, 1, 0.0)]
    st = config.derive(16, 16, 0.0, specs, 1e15, 0.05, bc="pbc")
    d = st.domain
    rng = np.random.default_rng(13)
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, np.full(16, 1.0), np.full(16, 1.0), np.zeros(16), np.zeros(16), np.zeros(16))
    v = np.tile(np.asarray(v3, dtype=np.float64) / d.vel_norm, (N, 1))
    ctx.upload(1, rng.random(N) * 16, rng.random(N) * 16, v[:, 0], v[:, 1], v[:, 2])
    ctx.mcc_setup(xsec.mcc_desc(xsec.synthetic_ar(), electron_sp=-1, ion_sp=1, gas_density=1e24,
                                gas_temp_K=gas_temp_K, enable=enable))
    return ctx, d, v


def test_ion_charge_exchange_takes_the_neutral_maxwellian():
    """Resonant charge exchange hands the ion the neutral's velocity: ions
    loaded as a cold beam (every ion at 20 thermal speeds along x) that
    exchange charge with 300 K argon leave with velocities that are the
    300 K Maxwellian, N(0, kT/M) per component (KS), independent of the
    beam; the rest keep the beam velocity exactly."""
    print("[mcc] Ar+ charge exchange", flush=True)
    N = 200000
    s_n = np.sqrt(config.K_b * 300.0 / xsec.M_AR)
    ctx, d, v = _ion_only(N, 300.0, (0, 0, 0, 0, 1), (20 * s_n, 0.0, 0.0))
    ctx.collide(0)
    x, y, vx, vy, vz, ids = ctx.download(1)
    vv = np.stack([vx, vy, vz], axis=1) * d.vel_norm
    changed = np.any(vv != v[ids] * d.vel_norm, axis=1)
    ncx = int(np.count_nonzero(changed))
    assert ncx == ctx.mcc_stats()["count"][4] and ncx > 2000
    from math import erf, sqrt
    for k in range(3):
        z = np.sort(vv[changed, k] / s_n)
        cdf = 0.5 * (1 + np.array([erf(t / sqrt(2)) for t in z]))
        n = len(z)
        dks = max(np.max(np.arange(1, n + 1) / n - cdf), np.max(cdf - np.arange(n) / n))
        assert dks * np.sqrt(n) < 1.95, (k, dks)   # KS at the 0.1 % level
    ctx.close()


def test_ion_elastic_conserves_relative_speed():
    """Ar+ - Ar elastic scattering off neutrals at rest (gas_temp_K = 0):
    equal masses put the centre of mass at v/2 and the relative speed |v|
    is kept, so each scattered ion ends at v/2 + (|v|/2) n with n a unit
    vector, isotropic in the centre-of-mass frame (cos(v, n) uniform)."""
    print("[mcc] Ar+ elastic", flush=True)
    N = 200000
    s_n = np.sqrt(config.K_b * 300.0 / xsec.M_AR)
    ctx, d, v = _ion_only(N, 0.0, (0, 0, 0, 1, 0), (3 * s_n, 4 * s_n, 0.0))
    ctx.collide(0)
    x, y, vx, vy, vz, ids = ctx.download(1)
    v1 = np.stack([vx, vy, vz], axis=1)
    v0 = v[ids]
    changed = np.any(v1 != v0, axis=1)
    nel = int(np.count_nonzero(changed))
    assert nel == ctx.mcc_stats()["count"][3] and nel > 2000
    half = np.linalg.norm(v0[changed], axis=1) / 2
    nvec = v1[changed] - v0[changed] / 2
    assert np.allclose(np.linalg.norm(nvec, axis=1), half, rtol=1e-12, atol=0)
    c = np.sum(nvec * v0[changed], axis=1) / (half * 2 * half)
    u = np.sort((c + 1) / 2)
    n = len(u)
    dks = max(np.max(np.arange(1, n + 1) / n - u), np.max(u - np.arange(n) / n))
    assert dks * np.sqrt(n) < 1.95
    ctx.close()


def test_many_ionisations_per_step_appended_in_parent_order():
    """More ionisation events in one step than the former 16384-event
    limit: 2,000,000 electrons at 40 eV in 1e22 m^-3 argon ionise ~5e4
    times in one collision call.  Every event gives one electron and one ion
    (no CAPACITY count), appended in parent-slot order: the k-th new
    electron sits at the position of the k-th ionising parent in slot order,
    and ids continue in that order."""
    print("[mcc] many ionisations per step", flush=True)
    N = 2000000
    st = _setup(N, 40.0)
    d = st.domain
    rng = np.random.default_rng(31)
    speed = np.sqrt(2 * 40.0 * config.QE / ME) / d.vel_norm
    c = rng.uniform(-1, 1, N)
    ph = rng.uniform(0, 2 * np.pi, N)
    sn = np.sqrt(1 - c * c)
    e = (rng.random(N) * 16, rng.random(N) * 16, speed * sn * np.cos(ph), speed * sn * np.sin(ph), speed * c)
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, *e)
    ctx.upload(1, [1.0], [1.0], [0.0], [0.0], [0.0])
    ctx.mcc_setup(xsec.mcc_desc(xsec.reference_e_ar(), gas_density=1e22, gas_temp_K=0.0, enable=(1, 1, 1, 0, 0)))
    ctx.collide(0)
    st_ = ctx.mcc_stats()
    nion = st_["count"][2]
    print(f"  ionisations {nion}, births {st_['births']}", flush=True)
    assert nion > 20000
    assert st_["births"] == nion
    assert not ctx.errors().any(), ctx.errors()
    x, y, vx, vy, vz, ids = ctx.download(0)
    assert len(x) == N + nion
    # parents: the loaded electrons whose position the births share
    pos = {(a, b): k for k, (a, b) in enumerate(zip(e[0].tolist(), e[1].tolist()))}
    parents = np.array([pos[(a, b)] for a, b in zip(x[N:].tolist(), y[N:].tolist())])
    assert np.all(np.diff(parents) > 0)
    assert np.array_equal(ids[N:], np.arange(N, N + nion, dtype=np.uint32))
    xi, yi, _, _, _, idi = ctx.download(1)
    assert len(xi) == 1 + nion
    assert np.array_equal(xi[1:], x[N:]) and np.array_equal(yi[1:], y[N:])
    ctx.close()


@pytest.mark.parametrize("how", ["upload", "push", "push+clear"])
def test_absorbed_slots_never_collide(how):
    """A slot holding an absorbed particle (NaN x) takes no collision, on
    periodic walls where the selection pass skips reading the candidates' x
    when the species provably holds no such slot.  The NaN comes with an
    upload, or from a push that drops particles with a non-finite velocity
    (counted on the device as EPIC_ERRC_NAN), or from such a push whose count
    was cleared before the collision step.  Elastic only: a selected absorbed
    slot would add to the collision count, which must equal the number of
    live particles whose velocity changed."""
    print(f"[mcc] absorbed slots never collide ({how})", flush=True)
    N, gas = 30000, 2e22
    st = _setup(N, 8.0)
    d = st.domain
    m = xsec.mcc_desc(xsec.synthetic_ar(), gas_density=gas, gas_temp_K=0.0, enable=(1, 0, 0, 0, 0))
    x, y, vx, vy, vz = [np.array(a) for a in _electrons(st, N, 8.0, 5)]
    dead = np.zeros(N, dtype=bool)
    dead[::3] = True
    if how == "upload":
        x[dead] = np.nan
    else:
        vx[dead] = np.nan
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, x, y, vx, vy, vz)
    ctx.upload(1, [1.0], [1.0], [0.0], [0.0], [0.0])
    ctx.mcc_setup(m)
    if how != "upload":
        ctx.push(0)
        assert int(ctx.errors()[5]) == int(dead.sum())  # EPIC_ERRC_NAN
        if how == "push+clear":
            ctx.clear_errors()
            assert int(ctx.errors()[5]) == 0
    # the download returns the live particles only
    _, _, vx0, vy0, vz0, ids0 = ctx.download(0)
    assert len(ids0) == N - int(dead.sum())
    ctx.collide(7)
    _, _, vx1, vy1, vz1, ids1 = ctx.download(0)
    count = ctx.mcc_stats()["count"]
    ctx.close()
    assert np.array_equal(ids0, ids1)
    hit = (vx1 != vx0) | (vy1 != vy0) | (vz1 != vz0)
    assert hit.sum() > 1000
    assert int(count[0]) == int(hit.sum())
