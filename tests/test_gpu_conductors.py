"""Internal conductors on the device (epic_set_conductors) against the oracle:
absorption in the push (species.cpp:84-98, 111-124, 152-165), the SOR's
rho := voltage nodes (grid.cpp:68-85, electrode.cpp:4-17), the circular
grid's electrodes existing only once an SOR solve has run, and the loader's
rejection of positions on a conductor (init.cpp:71-96)."""
import numpy as np
import pytest

from epic2d_mcc_amd import _lib
from harness import rel, small_setup
from oracle_api import OracleSim

pytestmark = pytest.mark.gpu

RECT = {"kind": _lib.COND_RECT, "min_x": 2, "max_x": 5, "min_y": 10, "max_y": 13, "voltage": 3.0}
CIRC = {"kind": _lib.COND_CIRCULAR, "x_center": 9, "y_center": 7, "electrode_number": 4, "grid_radius": 3.5,
        "electrode_radius": 1.2, "voltage": -2.0}


def _pair(st, conds):
    arr = _lib.conductors(conds)
    o = OracleSim(st.domain, st.species)
    o.set_conductors(arr)
    for s in range(len(st.species)):
        assert o.init_species(s) == 0
    ctx = _lib.Context(st.domain, st.species)
    ctx.set_conductors(arr)
    for s in range(len(st.species)):
        ctx.upload(s, *o.particles(s))
    return o, ctx


def _compare(o, ctx, tag, ftol, ptol):
    for which, name in ((0, "rho"), (1, "phi"), (2, "efx"), (3, "efy")):
        r = rel(ctx.get_field(which), o.grid(which))
        assert r < ftol, f"{tag}: {name} rel {r:.3e}"
    for s in range(2):
        g = ctx.download(s)
        c = o.particles(s)
        assert len(g[0]) == len(c[0]), f"{tag}: species {s} count {len(g[0])} vs {len(c[0])}"
        for k in range(4):
            r = rel(g[k], c[k])
            assert r < ptol, f"{tag}: species {s} component {k} rel {r:.3e}"


@pytest.mark.parametrize("bc", ["pbc", "open", "reflective"])
def test_pcg_absorbs_on_the_rectangle_only(bc):
    """Reference PCG (bitwise path): the rectangle absorbs from the first
    push, the circular grid never has electrodes, and no rho is written."""
    print(f"[conductors] pcg {bc}", flush=True)
    st = small_setup(N=3000, bc=bc, tol=1e-10, vsx_beam=2.0 if bc != "pbc" else 0.0)
    o, ctx = _pair(st, [RECT, CIRC])
    n0 = [len(o.particles(s)[0]) for s in range(2)]
    ctx.initial_field()
    o.initial_field()
    _compare(o, ctx, "initial", 1e-12, 1e-12)
    for ts in range(8):
        itg = ctx.step(ts)
        ito = o.step()
        assert itg == ito
        _compare(o, ctx, f"step {ts}", 1e-12, 1e-12)
    e = o.particles(0)
    assert len(e[0]) < n0[0]
    # electrons sit inside the circular grid's disks: they were never absorbed
    x, y = e[0], e[1]
    inside = np.zeros(len(x), dtype=bool)
    for k in range(4):
        t = 2 * np.pi * k / 4
        inside |= np.hypot(x - (9 + 3.5 * np.cos(t)), y - (7 + 3.5 * np.sin(t))) <= 1.2
    assert np.count_nonzero(inside) > 0
    assert not np.any(ctx.errors())


@pytest.mark.parametrize("bc", ["open", "pbc"])
def test_sor_writes_voltages_and_electrodes_absorb(bc):
    """SOR: rho carries the conductor voltages after each solve, and the
    circular grid's electrodes absorb from the first step.  Red-black and
    lexicographic SOR agree at convergence, so phi and the particles are
    compared at that level."""
    print(f"[conductors] sor {bc}", flush=True)
    st = small_setup(N=3000, bc=bc, solver=_lib.SOLVER_GS, tol=1e-12, maxit=200000,
                     vsx_beam=2.0 if bc != "pbc" else 0.0)
    o, ctx = _pair(st, [RECT, CIRC])
    n0 = [len(o.particles(s)[0]) for s in range(2)]
    ctx.initial_field()
    o.initial_field()
    # the voltage nodes hold the voltages exactly; the deposit elsewhere
    # agrees to the integer deposit's ~1e-14
    ro, rg = o.grid(0), ctx.get_field(0)
    marked = np.isin(ro, [RECT["voltage"], CIRC["voltage"]])
    assert np.count_nonzero(marked) > 20
    assert np.array_equal(rg[marked], ro[marked])
    _compare(o, ctx, "initial", 1e-9, 1e-9)
    for ts in range(5):
        ctx.step(ts)
        o.step()
        _compare(o, ctx, f"step {ts}", 1e-8, 1e-8)
    assert len(o.particles(0)[0]) < n0[0]


def test_fp32_with_conductors():
    """FP32 storage: the conductor test runs in FP64 on the FP32 positions;
    counts agree to a few boundary cases and states to the FP32 path's
    tolerances."""
    print("[conductors] fp32", flush=True)
    st = small_setup(N=20000, Nx=32, Ny=32, bc="open", precision=_lib.FP32, vsx_beam=2.0, tol=1e-8)
    conds = [{"kind": _lib.COND_RECT, "min_x": 4, "max_x": 10, "min_y": 20, "max_y": 26, "voltage": 1.0},
             {"kind": _lib.COND_CIRCULAR, "x_center": 18, "y_center": 14, "electrode_number": 5,
              "grid_radius": 6.0, "electrode_radius": 2.0, "voltage": -1.0}]
    o, ctx = _pair(st, conds)
    ctx.initial_field()
    o.initial_field()
    for ts in range(3):
        ctx.step(ts)
        o.step()
    for s in range(2):
        g = ctx.download(s)
        c = o.particles(s)
        assert abs(len(g[0]) - len(c[0])) <= 3
        if len(g[0]) == len(c[0]):
            assert rel(g[0], c[0]) < 1e-5 and rel(g[1], c[1]) < 1e-5
    assert len(o.particles(0)[0]) < 20000


def test_conductor_errors():
    print("[conductors] errors", flush=True)
    st = small_setup(N=100, bc="open", solver=_lib.SOLVER_GS)
    ctx = _lib.Context(st.domain, st.species)
    bad = _lib.conductors([{"kind": _lib.COND_RECT, "min_x": 10, "max_x": 20, "min_y": 2, "max_y": 4,
                            "voltage": 1.0}])
    with pytest.raises(_lib.EpicError, match="EPIC_ERR_ARG"):
        ctx.set_conductors(bad)
    ctx.set_conductors(_lib.conductors([]))
    ctx.close()


WIDE = [{"kind": _lib.COND_RECT, "min_x": 20, "max_x": 40, "min_y": 70, "max_y": 100, "voltage": 3.0},
        {"kind": _lib.COND_CIRCULAR, "x_center": 80, "y_center": 50, "electrode_number": 6, "grid_radius": 25.0,
         "electrode_radius": 7.0, "voltage": -2.0}]


@pytest.mark.parametrize("bc", ["pbc", "open", "reflective"])
def test_windowed_fp64_push_with_conductors_bitwise(bc):
    """129 x 129 nodes (the dedicated FP64 push kernel, with its conductor
    test): from the oracle's own E field, after an SOR solve (electrodes
    live), every surviving particle is bit-identical to the oracle's and the
    same particles are erased."""
    print(f"[conductors] windowed fp64 {bc}", flush=True)
    st = small_setup(N=60000, Nx=128, Ny=128, bc=bc, solver=_lib.SOLVER_GS, tol=1e-6, maxit=20000,
                     vsx_beam=2.0 if bc != "pbc" else 0.0)
    o, ctx = _pair(st, WIDE)
    for s in range(2):
        o.scatter(s)
        ctx.scatter(s)
    o.compute_rho()
    ctx.compute_rho()
    o.solve()
    ctx.solve()
    o.efield()
    ctx.set_field(2, o.grid(2))
    ctx.set_field(3, o.grid(3))
    n0 = len(o.particles(0)[0])
    for ts in range(3):
        for s in range(2):
            ctx.push(s)
            o.push(s)
        for s in range(2):
            g = ctx.download(s)
            c = o.particles(s)
            assert len(g[0]) == len(c[0]), f"push {ts}: species {s} count {len(g[0])} vs {len(c[0])}"
            for k in range(4):
                assert np.array_equal(g[k], c[k]), f"push {ts}: species {s} component {k}"
    assert len(o.particles(0)[0]) < n0


def _fp32_run(st, env, steps=3):
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
    ctx.set_conductors(_lib.conductors(WIDE))
    o = OracleSim(st.domain, st.species)
    o.set_conductors(_lib.conductors(WIDE))
    for s in range(2):
        assert o.init_species(s) == 0
        ctx.upload(s, *o.particles(s))
    ctx.initial_field()
    for ts in range(steps):
        ctx.step(ts)
    out = [ctx.download(s)[:4] for s in range(2)], ctx.get_field(0).copy()
    ctx.close()
    return out, o


@pytest.mark.parametrize("bc", ["open", "pbc"])
def test_fp32_conductor_kernels_identical(bc):
    """FP32 windowed pushes with conductors: the one-launch kernel and the
    per-species kernel give the same particles and rho bit for bit, so does
    the exact test in every cell (no cell bitmap), and the particles track
    the FP64 oracle."""
    print(f"[conductors] fp32 kernels {bc}", flush=True)
    st = small_setup(N=100000, Nx=128, Ny=128, bc=bc, precision=_lib.FP32, solver=_lib.SOLVER_GS, tol=1e-6,
                     maxit=20000, vsx_beam=2.0 if bc != "pbc" else 0.0)
    (a, ra), o = _fp32_run(st, {})
    (b, rb), _ = _fp32_run(st, {"EPIC_PUSH_MULTI": "0"})
    (m, rm), _ = _fp32_run(st, {"EPIC_COND_CELL_MASK": "0"})
    for s in range(2):
        for k in range(4):
            assert np.array_equal(a[s][k], b[s][k], equal_nan=True), (s, k)
            assert np.array_equal(a[s][k], m[s][k], equal_nan=True), ("no bitmap", s, k)
    assert np.array_equal(ra, rb)
    assert np.array_equal(ra, rm)
    o.initial_field()
    for ts in range(3):
        o.step()
    for s in range(2):
        c = o.particles(s)
        assert abs(len(a[s][0]) - len(c[0])) <= 3


def test_batch_member_rejects_conductors():
    """The batched kernels have no conductor test, so a batch member refuses
    a conductor list (an empty one is accepted)."""
    print("[conductors] batch member", flush=True)
    st = small_setup(N=1000, Nx=32, Ny=32, precision=_lib.FP32, solver=_lib.SOLVER_PCG_MIXED, tol=1e-10)
    st.domain.tol_relative = 1
    st.domain.precond = _lib.PRECOND_SPECTRAL
    B = _lib.Batch()
    mid, m = B.add(st.domain, st.species)
    with pytest.raises(_lib.EpicError, match="EPIC_ERR_ARG"):
        m.set_conductors(_lib.conductors([RECT]))
    m.set_conductors(_lib.conductors([]))
    B.close()
