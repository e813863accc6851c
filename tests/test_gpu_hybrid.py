"""The Boltzmann-electron SOR (EPIC_SOLVER_GS_BOLTZMANN, gsfluidsolver
field.cpp:7-68) on the device against the oracle and a manufactured
solution.  The device sweeps red-black, the reference lexicographically:
both converge to the same solution, so phi is compared at the solver's
convergence level."""
import numpy as np
import pytest

from epic2d_mcc_amd import _lib
from harness import reference_load, rel, small_setup
from oracle_api import OracleSim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Nx,Ny", [(20, 14), (96, 80)])
def test_boltzmann_sor_manufactured(Nx, Ny):
    print(f"[hybrid] manufactured {Nx}x{Ny}", flush=True)
    st = small_setup(N=10, Nx=Nx, Ny=Ny, solver=_lib.SOLVER_GS_BOLTZMANN, tol=1e-12, maxit=400000)
    d = st.domain
    nx, ny = d.nx, d.ny
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    phi_star = 0.25 * np.sin(np.pi * i / (nx - 1)) * np.sin(np.pi * j / (ny - 1)) - \
        0.08 * np.sin(2 * np.pi * i / (nx - 1)) * np.sin(3 * np.pi * j / (ny - 1))
    lap = np.zeros_like(phi_star)
    lap[1:-1, 1:-1] = (phi_star[2:, 1:-1] - 2 * phi_star[1:-1, 1:-1] + phi_star[:-2, 1:-1]) + \
                      (phi_star[1:-1, 2:] - 2 * phi_star[1:-1, 1:-1] + phi_star[1:-1, :-2])
    s = (d.L * d.L) / (d.LDe * d.LDe)
    rho = -lap / s + np.exp(phi_star)
    ctx = _lib.Context(d, st.species)
    ctx.set_field(0, rho)
    it, err = ctx.solve()
    assert err < 1e-12 and it % 50 == 1
    assert np.allclose(ctx.get_field(1).reshape(nx, ny), phi_star, atol=1e-9)
    o = OracleSim(d, st.species)
    o.set_grid(0, rho)
    o.solve()
    same = np.array_equal(np.asarray(ctx.get_field(0)).reshape(nx, ny), o.grid(0))
    assert same  # boundary rho written alike
    assert rel(ctx.get_field(1), o.grid(1)) < 1e-9
    ctx.close()


def test_boltzmann_sor_steps_track_oracle():
    """Ions pushed against Boltzmann electrons: fused steps follow the
    oracle's loop at the SOR convergence level."""
    print("[hybrid] steps", flush=True)
    st = small_setup(N=2000, Nx=16, Ny=16, bc="open", solver=_lib.SOLVER_GS_BOLTZMANN, tol=1e-12, maxit=400000)
    o, parts = reference_load(st)
    ctx = _lib.Context(st.domain, st.species)
    for s, p in enumerate(parts):
        ctx.upload(s, *p)
    ctx.initial_field()
    o.initial_field()
    for ts in range(4):
        ctx.step(ts)
        o.step()
        assert rel(ctx.get_field(1), o.grid(1)) < 1e-8, ts
        for s in range(2):
            g = ctx.download(s)
            c = o.particles(s)
            assert len(g[0]) == len(c[0])
            assert rel(g[0], c[0]) < 1e-9 and rel(g[2], c[2]) < 1e-7


@pytest.mark.parametrize("Nx,Ny", [(16, 16), (20, 12)])
def test_newton_cg_matches_oracle(Nx, Ny):
    """nrpcgsolver on the device (one workgroup, workgroup dot products)
    against the oracle's sequential restatement: the CG iteration counts
    agree and phi agrees to rounding."""
    print(f"[hybrid] newton-cg {Nx}x{Ny}", flush=True)
    st = small_setup(N=10, Nx=Nx, Ny=Ny, solver=_lib.SOLVER_NEWTON_CG)
    d = st.domain
    rng = np.random.default_rng(6)
    rho = rng.normal(scale=0.3, size=(d.nx, d.ny))
    ctx = _lib.Context(d, st.species)
    ctx.set_field(0, rho)
    itg, err = ctx.solve()
    o = OracleSim(d, st.species)
    o.set_grid(0, rho)
    ito, _ = o.solve()
    assert abs(itg - ito) <= 2, (itg, ito)
    assert err == 0.0
    assert rel(ctx.get_field(1), o.grid(1)) < 1e-9
    ctx.close()


def test_newton_cg_steps_track_oracle():
    print("[hybrid] newton-cg steps", flush=True)
    st = small_setup(N=2000, Nx=16, Ny=16, bc="pbc", solver=_lib.SOLVER_NEWTON_CG)
    o, parts = reference_load(st)
    ctx = _lib.Context(st.domain, st.species)
    for s, p in enumerate(parts):
        ctx.upload(s, *p)
    ctx.initial_field()
    o.initial_field()
    for ts in range(4):
        ctx.step(ts)
        o.step()
        assert rel(ctx.get_field(1), o.grid(1)) < 1e-8, ts
        for s in range(2):
            g = ctx.download(s)
            c = o.particles(s)
            assert len(g[0]) == len(c[0])
            assert rel(g[0], c[0]) < 1e-9 and rel(g[2], c[2]) < 1e-7
