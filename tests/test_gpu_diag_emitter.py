"""Energy / momentum diagnostics against the oracle (FP64, 1e-12) and the
emitter injection (counts, placement, velocity moments)."""
import numpy as np
import pytest

from epic2d_mcc_amd import _lib, config
from harness import reference_load, small_setup

pytestmark = pytest.mark.gpu


def test_diagnostics_match_oracle():
    st = small_setup(N=20000, Nx=32, Ny=32, tol=1e-10)
    o, parts = reference_load(st)
    ctx = _lib.Context(st.domain, st.species)
    for s, p in enumerate(parts):
        ctx.upload(s, *p)
    ctx.initial_field()
    o.initial_field()
    for ts in range(5):
        ctx.step(ts)
        o.step()
    # the reference stores KE after the push with the fields of that step
    for s in range(2):
        for ns in range(2):
            a, b = ctx.ke(s, ns), o.ke(s, ns)
            assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b) + 1e-300), (s, ns, a, b)
            a, b = ctx.momentum(s, ns), o.momentum(s, ns)
            assert np.all(np.abs(a - b) <= 1e-10 * np.max(np.abs(b)) + 1e-300), (s, ns, a, b)
    assert abs(ctx.pe(0) - o.pe(0)) <= 1e-12 * abs(o.pe(0))


def test_emitter_injection():
    """Emitter::inject (emitter.cpp:5-36): numparticle per step for each of
    the two species at shared positions uniform on [x0,x1] x [y0,y1];
    velocity SampleVel(temp) + drift, vz = 0."""
    st = small_setup(N=1000, Nx=32, Ny=32, capacity_factor=3.0)
    d = st.domain
    ctx = _lib.Context(d, st.species)
    o, parts = reference_load(st)
    for s, p in enumerate(parts):
        ctx.upload(s, *p)
    e = _lib.EmitterDesc(x0=2.0, y0=10.0, x1=3.0, y1=20.0, temp=1.0, numparticle=200, vdx=2.0, vdy=0.0,
                         species_idx1=0, species_idx2=1)
    ctx.add_emitter(e)
    for ts in range(5):
        ctx.inject(ts)
    assert ctx.count(0) == 1000 + 5 * 200 and ctx.count(1) == 1000 + 5 * 200
    g0 = ctx.download(0)
    g1 = ctx.download(1)
    new0 = [a[1000:] for a in g0]
    new1 = [a[1000:] for a in g1]
    assert np.all((new0[0] >= 2.0) & (new0[0] <= 3.0) & (new0[1] >= 10.0) & (new0[1] <= 20.0))
    assert np.array_equal(new0[0], new1[0]) and np.array_equal(new0[1], new1[1])
    assert np.all(new0[4] == 0.0)
    # drift vdx*vel_norm normalised -> 2; spread sqrt(kT/m)/vel_norm = 1 for electrons at 1 eV
    assert abs(new0[2].mean() - 2.0) < 0.1 and abs(new0[2].std() - 1.0) < 0.1
    # ions at the same temperature: spread sqrt(m_e/m_i)
    assert abs(new1[3].std() - np.sqrt(1 / 1836.0)) < 0.005
    # ids continue after the loaded ones
    assert np.array_equal(new0[5], np.arange(1000, 2000, dtype=np.uint32))


def test_fused_step_with_emitter_equals_discrete():
    """In the fused step the injected particles are deposited by the
    injection kernel; the result equals inject -> scatter -> ... exactly."""
    st = small_setup(N=3000, Nx=32, Ny=32, capacity_factor=3.0, tol=1e-10)
    o, parts = reference_load(st)
    e = _lib.EmitterDesc(x0=1.0, y0=5.0, x1=2.0, y1=25.0, temp=1.0, numparticle=100, vdx=3.0, vdy=0.0,
                         species_idx1=0, species_idx2=1)
    a = _lib.Context(st.domain, st.species)
    b = _lib.Context(st.domain, st.species)
    for c in (a, b):
        for s, p in enumerate(parts):
            c.upload(s, *p)
        c.add_emitter(e)
        c.initial_field()
    for ts in range(4):
        a.step(ts)
        b.inject(ts)
        for s in range(2):
            b.scatter(s)
        b.compute_rho()
        b.solve()
        b.efield()
        for s in range(2):
            b.push(s)
    for s in range(2):
        ga, gb = a.download(s), b.download(s)
        assert len(ga[0]) == len(gb[0])
        for k in range(6):
            assert np.array_equal(ga[k], gb[k]), (s, k)


def _ks2(a, b):
    """Two-sample Kolmogorov-Smirnov statistic and its asymptotic p-value."""
    a, b = np.sort(a), np.sort(b)
    allv = np.concatenate([a, b])
    fa = np.searchsorted(a, allv, side="right") / len(a)
    fb = np.searchsorted(b, allv, side="right") / len(b)
    d = float(np.max(np.abs(fa - fb)))
    ne = len(a) * len(b) / (len(a) + len(b))
    lam = (np.sqrt(ne) + 0.12 + 0.11 / np.sqrt(ne)) * d
    p = 2 * sum((-1) ** (j - 1) * np.exp(-2 * j * j * lam * lam) for j in range(1, 101))
    return d, float(min(max(p, 0.0), 1.0))


def test_emitter_distribution_matches_oracle():
    """The device injection against the oracle's restatement of
    Emitter::inject (emitter.cpp:5-36: positions uniform on the emitter box,
    v = SampleVel(temp) + drift, normalised, vz = 0).  The two draw different
    numbers (Philox keyed by (emitter, step, k) here; the reference's
    positions come from std::random_device, the oracle's from a seeded
    stream), so they are compared in distribution: two-sample KS on x, y,
    vx, vy for each target species; the Irwin-Hall bound |v - drift| <=
    3 sqrt(kT/m) on every device draw; vx and vy uncorrelated."""
    print("[emitter] distribution vs oracle", flush=True)
    n, steps = 20000, 5
    st = small_setup(N=16, Nx=32, Ny=32, capacity_factor=(16 + n * steps) / 16 + 1)
    d = st.domain
    e = _lib.EmitterDesc(x0=2.0, y0=4.0, x1=6.0, y1=28.0, temp=3.0, numparticle=n, vdx=1.5, vdy=-0.5,
                         species_idx1=0, species_idx2=1)
    ctx = _lib.Context(d, st.species)
    for s in range(2):
        ctx.upload(s, [], [], [], [], [])
    ctx.add_emitter(e)
    for ts in range(steps):
        ctx.inject(ts)
    assert not ctx.errors().any()
    from oracle_api import OracleSim
    o = OracleSim(d, st.species)
    for s in range(2):
        o.set_particles(s, [], [], [], [], [])
    for ts in range(steps):
        o.inject(e, 1000 + ts)
    for s in range(2):
        g = ctx.download(s)
        c = o.particles(s)
        assert len(g[0]) == len(c[0]) == n * steps
        sig = np.sqrt(3.0 * config.QE / st.species[s].mass) / d.vel_norm
        for k, name in enumerate(("x", "y", "vx", "vy")):
            dks, p = _ks2(g[k], c[k])
            print(f"  species {s} {name}: KS {dks:.4f} p {p:.3f}", flush=True)
            assert p > 1e-3, (s, name, dks, p)
        assert np.all(np.abs(g[2] - 1.5) <= 3 * sig * (1 + 1e-6))
        assert np.all(np.abs(g[3] + 0.5) <= 3 * sig * (1 + 1e-6))
        assert np.all(g[4] == 0.0)
        r = np.corrcoef(g[2], g[3])[0, 1]
        assert abs(r) < 5 / np.sqrt(len(g[2])), r
        assert np.all((g[0] >= 2.0) & (g[0] < 6.0) & (g[1] >= 4.0) & (g[1] < 28.0))
    ctx.close()
