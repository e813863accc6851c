"""FP32 particle storage (the large workloads): the fused step equals the
discrete operator sequence bit for bit (integer deposit), and the state
tracks the FP64 oracle within FP32 tolerances over the first steps."""
import numpy as np
import pytest

from epic2d_mcc_amd import _lib
from harness import reference_load, rel, small_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bc", ["pbc", "open"])
def test_fp32_fused_equals_discrete(bc):
    """Grid above the whole-grid LDS size (128x128): the FP32 hot kernel."""
    print(f"[fp32] fused vs discrete bc={bc}", flush=True)
    st = small_setup(N=300000, Nx=128, Ny=128, precision=_lib.FP32, bc=bc, vsx_beam=3.0 if bc == "open" else 0.0,
                     solver=_lib.SOLVER_PCG_MIXED, poisson_op=_lib.OP_DIRICHLET_WALL, tol=1e-10)
    st.domain.tol_relative = 1
    st.domain.precond = _lib.PRECOND_SPECTRAL
    o, parts = reference_load(st)
    a = _lib.Context(st.domain, st.species)
    b = _lib.Context(st.domain, st.species)
    for s, p in enumerate(parts):
        a.upload(s, *p)
        b.upload(s, *p)
    a.initial_field()
    b.initial_field()
    for ts in range(4):
        a.step(ts)
        for s in range(2):
            b.scatter(s)
        b.compute_rho()
        b.solve()
        b.efield()
        for s in range(2):
            b.push(s)
    for s in range(2):
        ga, gb = a.download(s), b.download(s)
        for k in range(4):
            assert np.array_equal(ga[k], gb[k])
    for s in range(2):
        a.scatter(s)
        b.scatter(s)
    a.compute_rho()
    b.compute_rho()
    assert np.array_equal(a.get_field(0), b.get_field(0))


def test_fp32_tracks_fp64_oracle():
    print("[fp32] vs oracle", flush=True)
    st = small_setup(N=300000, Nx=128, Ny=128, precision=_lib.FP32, tol=1e-12)
    o, parts = reference_load(st)
    ctx = _lib.Context(st.domain, st.species)
    for s, p in enumerate(parts):
        ctx.upload(s, *p)
    ctx.initial_field()
    o.initial_field()
    for ts in range(3):
        ctx.step(ts)
        o.step()
    # rho: integer deposit of FP32 positions with 2^-14 cell weights; the
    # field error (~1e-3) shows in the ion velocities, which are ~260x
    # smaller than the electron ones
    er = rel(ctx.get_field(0), o.grid(0))
    ep = rel(ctx.get_field(1), o.grid(1))
    print(f"  rho {er:.2e} phi {ep:.2e}")
    assert er < 2e-3 and ep < 2e-3
    for s in range(2):
        g = ctx.download(s)
        c = o.particles(s)
        errs = [rel(g[k], c[k]) for k in range(4)]
        print(f"  species {s}: x {errs[0]:.2e} y {errs[1]:.2e} vx {errs[2]:.2e} vy {errs[3]:.2e}")
        assert errs[0] < 1e-5 and errs[1] < 1e-5
        assert errs[2] < 2e-3 and errs[3] < 2e-3


def _run_env(st, parts, env, steps=3):
    """A context created under the given kernel-shape environment (read at
    epic_create), after `steps` fused steps: particles and rho."""
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = _lib.Context(st.domain, st.species)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for s, p in enumerate(parts):
        ctx.upload(s, *p)
    ctx.initial_field()
    for ts in range(steps):
        ctx.step(ts)
    out = [ctx.download(s)[:4] for s in range(len(parts))], ctx.get_field(0).copy()
    ctx.close()
    return out


_SHAPES = [("open", {"EPIC_PUSH_MULTI": "0"}), ("open", {"EPIC_PUSH_VARIANT": "1"}),
           ("open", {"EPIC_PUSH_VARIANT": "2"}), ("open", {"EPIC_PUSH_VARIANT": "3"}),
           ("open", {"EPIC_PUSH_VARIANT": "4"}), ("open", {"EPIC_PUSH_MAP": "0"}),
           ("open", {"EPIC_PUSH_GRID": "2"}), ("open", {"EPIC_PUSH_MULTI_WAVES": "1"}),
           ("pbc", {"EPIC_PUSH_MULTI": "0"}),
           ("reflective", {"EPIC_PUSH_MULTI": "0"})]


@pytest.mark.parametrize("bc,env", _SHAPES)
def test_fp32_kernel_shapes_identical(bc, env):
    """Every FP32 push shape (one launch for all species or one per species,
    particles per thread, occupancy target, chunk walk, grid size) gives the
    default's particles and rho bit for bit; the one-launch and per-species
    kernels also for periodic and reflecting walls."""
    print(f"[fp32] shape {bc} {env}", flush=True)
    if bc == "open":
        st = small_setup(N=200000, Nx=128, Ny=128, precision=_lib.FP32, bc=bc, vsx_beam=3.0,
                         solver=_lib.SOLVER_PCG_MIXED, poisson_op=_lib.OP_DIRICHLET_WALL, tol=1e-10)
        st.domain.tol_relative = 1
        st.domain.precond = _lib.PRECOND_SPECTRAL
    else:
        st = small_setup(N=200000, Nx=128, Ny=128, precision=_lib.FP32, bc=bc, vsx_beam=3.0,
                         solver=_lib.SOLVER_PCG, tol=1e-8)
    o, parts = reference_load(st)
    ref_parts, ref_rho = _run_env(st, parts, {})
    got_parts, got_rho = _run_env(st, parts, env)
    for s in range(2):
        for k in range(4):
            assert np.array_equal(ref_parts[s][k], got_parts[s][k], equal_nan=True), (s, k)
    assert np.array_equal(ref_rho, got_rho)


def test_long_runs_on_one_node_do_not_wrap():
    """16 particles per thread (EPIC_PUSH_VARIANT=4) all sitting exactly on
    one node: each adds 2^28 fixed-point units to that node, so a thread's
    run sums to 2^32, past a 32-bit accumulator.  Electrons and ions share
    the node, so rho = 0, E = 0 and nothing moves; the fused deposit must
    still give each species' full charge at that node and nothing elsewhere,
    equal to the default four-per-thread kernel's."""
    print("[fp32] 16-particle runs on one node", flush=True)
    n = 64 * 16 * 8
    st = small_setup(N=n, Nx=128, Ny=128, precision=_lib.FP32, tol=1e-8)
    x = np.full(n, 40.0)
    y = np.full(n, 50.0)
    z = np.zeros(n)
    parts = [(x, y, z, z, z), (x, y, z, z, z)]
    d = st.domain
    expect = n * st.species[0].spwt / (d.dx * d.dy * d.L * d.L * d.density)
    dens = []
    for env in ({"EPIC_PUSH_VARIANT": "4"}, {}):
        import os
        old = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            ctx = _lib.Context(st.domain, st.species)
        finally:
            for k, v in old.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        for s, p in enumerate(parts):
            ctx.upload(s, *p)
        ctx.initial_field()
        ctx.step(0)   # push + fused deposit of the (unmoved) positions
        ctx.step(1)   # its den
        den = [ctx.get_field(10 + s) for s in range(2)]
        assert not ctx.errors().any()
        ctx.close()
        for s in range(2):
            assert den[s][40, 50] == pytest.approx(expect, rel=1e-12), (env, s, den[s][40, 50], expect)
            rest = den[s].copy()
            rest[40, 50] = 0.0
            assert not rest.any(), env
        dens.append(den)
    for s in range(2):
        assert np.array_equal(dens[0][s], dens[1][s])


@pytest.mark.parametrize("bc", ["pbc", "open", "reflective"])
def test_fp32_positions_stay_below_upper_wall(bc):
    """FP64 positions just below the upper walls (x0 + Lx - 1e-9) round to
    the wall itself in FP32; the reference never holds a live particle there
    (its loader and push keep x < x0 + Lx).  Stored in FP32 they must stay
    below it: the stand-alone deposit and rewind see them inside the grid
    (no out-of-grid count) and every particle's charge is deposited."""
    print(f"[fp32] upper-wall rounding bc={bc}", flush=True)
    st = small_setup(N=4096, Nx=128, Ny=128, precision=_lib.FP32, bc=bc, tol=1e-8)
    d = st.domain
    Lx, Ly = (d.nx - 1) * d.dx, (d.ny - 1) * d.dy
    rng = np.random.default_rng(11)
    n = 4096
    x = rng.random(n) * Lx
    y = rng.random(n) * Ly
    x[:64] = Lx - 1e-9
    y[64:128] = Ly - 1e-9
    x[128:192] = Lx - 1e-9
    y[128:192] = Ly - 1e-9
    z = np.zeros(n)
    ctx = _lib.Context(d, st.species)
    ctx.upload(0, x, y, z, z, z)
    ctx.upload(1, y, x, z, z, z)
    for s in range(2):
        g = ctx.download(s)
        assert len(g[0]) == n
        assert np.all(g[0] < Lx) and np.all(g[1] < Ly)
    norm = d.dx * d.dy * d.L * d.L * d.density
    w = np.ones((d.nx, d.ny))
    if bc != "pbc":
        w[0, :] = w[-1, :] = w[:, 0] = w[:, -1] = 2.0
        w[0, 0] = w[0, -1] = w[-1, 0] = w[-1, -1] = 4.0
    for s in range(2):
        ctx.scatter(s)
        den = ctx.get_field(10 + s)
        if bc == "pbc":
            tot = float(den[:-1, :-1].sum())  # the folded copies counted once
        else:
            tot = float(np.sum(den * w))
        assert tot == pytest.approx(n * st.species[s].spwt / norm, rel=1e-12), s
    ctx.initial_field()
    ctx.step(0)
    assert not ctx.errors().any(), ctx.errors()
    ctx.close()


def test_periodic_deposit_across_the_boundary():
    """Periodic walls, every chunk straddling the boundary: the slots
    alternate between particles just inside the low edges and just inside
    the high edges, moving outward, so that a quarter of them wrap in the
    first push.  The fused push's deposit (its window placed in cell offsets
    modulo the period) gives the rho that a separate scatter of the pushed
    positions gives, bit for bit."""
    print("[fp32] periodic deposit across the boundary", flush=True)
    n = 64 * 4 * 256
    st = small_setup(N=n, Nx=128, Ny=128, precision=_lib.FP32, bc="pbc", tol=1e-8)
    d = st.domain
    Lx, Ly = (d.nx - 1) * d.dx, (d.ny - 1) * d.dy
    rng = np.random.default_rng(7)
    k = np.arange(n)
    hi = (k % 2) == 1
    # within 0.3 cells of the low or the high edge in x; y spread, with every
    # fourth slot also at an edge in y
    x = np.where(hi, Lx - 0.3 * d.dx * rng.random(n), 0.3 * d.dx * rng.random(n))
    y = rng.random(n) * Ly
    yedge = (k % 4) >= 2
    y = np.where(yedge & hi, Ly - 0.3 * d.dy * rng.random(n), np.where(yedge, 0.3 * d.dy * rng.random(n), y))
    # outward at up to ~0.6 cells per step (in units of the push's kx)
    a = _lib.Context(st.domain, st.species)
    step_len = 0.6 * d.dx / (d.vel_norm / (d.L * d.W) * d.DT)
    vx = np.where(hi, 1.0, -1.0) * step_len * rng.random(n)
    vy = np.where(yedge, np.where(hi, 1.0, -1.0) * step_len * rng.random(n), 0.0)
    z = np.zeros(n)
    parts = [(x, y, vx, vy, z), (x.copy(), y.copy(), z, z, z)]
    for s, p in enumerate(parts):
        a.upload(s, *p)
    a.step(0)          # push of the uploaded state with the fused deposit
    pushed = [a.download(s) for s in range(2)]
    wrapped = np.sum((~hi) & (pushed[0][0] > Lx / 2)) + np.sum(hi & (pushed[0][0] < Lx / 2))
    print(f"  electrons wrapped in x: {wrapped} of {n}", flush=True)
    assert wrapped > n // 8
    a.step(1)          # its rho, from the fused deposit
    rho_fused = a.get_field(0)
    assert not a.errors().any()
    b = _lib.Context(st.domain, st.species)
    for s in range(2):
        g = pushed[s]
        b.upload(s, g[0], g[1], g[2], g[3], g[4])
        b.scatter(s)
    b.compute_rho()
    assert np.array_equal(rho_fused, b.get_field(0))
    a.close()
    b.close()
