"""Pins of the oracle's Boltzmann-electron solvers: the SOR
(FieldSolve::gsfluidsolver, field.cpp:7-68) by a manufactured solution of
the nonlinear problem and the closed form of the boundary rho it writes; the
Newton-CG (FieldSolve::nrpcgsolver, field.cpp:141-210) bit for bit against
the same loop on the reference's own dense Matrix/vec/cg() templates."""
import numpy as np
import pytest

from epic2d_mcc_amd import _lib, config
from harness import small_setup
from oracle_api import OracleSim


def _lap(phi, dx2, dy2):
    lap = np.zeros_like(phi)
    lap[1:-1, 1:-1] = (phi[2:, 1:-1] - 2 * phi[1:-1, 1:-1] + phi[:-2, 1:-1]) / dx2 + \
                      (phi[1:-1, 2:] - 2 * phi[1:-1, 1:-1] + phi[1:-1, :-2]) / dy2
    return lap


def test_gsfluid_manufactured_solution():
    # phi* zero on the boundary; rho = -lap(phi*)/s + exp(phi*) makes phi*
    # the solution of lap phi = -(rho - exp(phi)) s
    st = small_setup(N=1, Nx=20, Ny=14, solver=_lib.SOLVER_GS_BOLTZMANN, tol=1e-12, maxit=200000)
    d = st.domain
    nx, ny = d.nx, d.ny
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    # amplitudes kept small: the pointwise SOR factor 1 - 1.4 (1 + e^phi s/4)
    # must stay inside (-1, 1), as it must for the reference
    phi_star = 0.25 * np.sin(np.pi * i / (nx - 1)) * np.sin(np.pi * j / (ny - 1)) - \
        0.08 * np.sin(2 * np.pi * i / (nx - 1)) * np.sin(3 * np.pi * j / (ny - 1))
    s = (d.L * d.L) / (d.LDe * d.LDe)
    rho = -_lap(phi_star, d.dx ** 2, d.dy ** 2) / s + np.exp(phi_star)
    o = OracleSim(d, st.species)
    o.set_grid(0, rho)
    it, err = o.solve()
    assert err < 1e-12 and it % 50 == 1
    assert np.allclose(o.grid(1), phi_star, atol=1e-9)
    # a linear SOR (no exp term) on the same rho lands elsewhere
    lin = small_setup(N=1, Nx=20, Ny=14, solver=_lib.SOLVER_GS, tol=1e-12, maxit=200000)
    o2 = OracleSim(lin.domain, lin.species)
    o2.set_grid(0, rho)
    o2.solve()
    assert not np.allclose(o2.grid(1), phi_star, atol=1e-3)


def test_gsfluid_boundary_rho():
    # field.cpp:23-33: bottom, left and right edges 0; the top edge
    # 10000 kT/e at 1 eV, its corners reset to 0 by the second loop
    st = small_setup(N=1, Nx=10, Ny=8, solver=_lib.SOLVER_GS_BOLTZMANN, tol=1e-6, maxit=100)
    d = st.domain
    o = OracleSim(d, st.species)
    o.set_grid(0, np.full((d.nx, d.ny), 0.25))
    o.solve()
    rho = o.grid(0)
    top = 10000 * (config.QE / (config.K_b * (config.QE / config.K_b)))
    assert np.all(rho[1:-1, -1] == top) and abs(top - 10000) < 1e-9
    assert np.all(rho[:, 0] == 0) and np.all(rho[0, :] == 0) and np.all(rho[-1, :] == 0)
    assert np.all(rho[1:-1, 1:-1] == 0.25)


@pytest.mark.parametrize("Nx,Ny", [(11, 11), (13, 9)])
def test_newton_cg_bitwise_equal_to_reference_templates(Nx, Ny):
    # nrpcgsolver restated matrix-free == the same Newton loop on the
    # reference's own dense Matrix/vec/cg() (oracle/_ref), bit for bit,
    # including the nx-stride matrix on a non-square grid
    from oracle_api import dp, ref_pcg_lib
    r = ref_pcg_lib()
    if r is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    st = small_setup(N=1, Nx=Nx, Ny=Ny, solver=_lib.SOLVER_NEWTON_CG)
    d = st.domain
    rng = np.random.default_rng(4)
    rho = rng.normal(scale=0.3, size=(d.nx, d.ny))
    o = OracleSim(d, st.species)
    o.set_grid(0, rho)
    o.solve()
    want = np.empty(d.nx * d.ny)
    s = (d.L * d.L) / (d.LDe * d.LDe)
    r.ref_nrpcg(d.nx, d.ny, dp(np.ascontiguousarray(rho)), d.dx * d.dx, s, dp(want))
    assert np.array_equal(o.grid(1).ravel(), want)
    assert np.any(want != 0)
