"""epic_solve_potential_noalloc (the sine-transform mixed solve with its
workspace in a static device array, no allocation) against
epic_solve_potential on the same context state, called through the C-ABI:
two calls in a row, the second on a smaller grid, each within the solve's
tolerance of the allocating entry point's phi."""
import ctypes

import numpy as np
import pytest

from epic2d_mcc_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _solve_both(cells, seed):
    st = workloads.bounded_512(ppc=1, cells=cells)
    d = st.domain
    rng = np.random.default_rng(seed)
    rho = rng.standard_normal((d.nx, d.ny))
    L = _lib.lib()
    out = []
    for fn in (L.epic_solve_potential, L.epic_solve_potential_noalloc):
        ctx = _lib.Context(d, st.species)
        ctx.set_field(0, rho)
        it = ctypes.c_int()
        err = ctypes.c_double()
        rc = fn(ctx.h, ctypes.byref(it), ctypes.byref(err))
        assert rc == 0, (fn.__name__, rc, L.epic_last_error(ctx.h))
        out.append((ctx.get_field(1).copy(), it.value, err.value, ctx.errors().copy()))
        ctx.close()
    (pa, ia, ea, xa), (pb, ib, eb, xb) = out
    return d, rho, pa, pb, (ia, ea), (ib, eb), xa, xb


@pytest.mark.parametrize("cells", [(512, 64)])
def test_noalloc_solve_equals_allocating_solve(cells):
    print(f"[noalloc] grids {cells}", flush=True)
    for k, c in enumerate(cells):
        d, rho, pa, pb, ra, rb, xa, xb = _solve_both(c, 100 + k)
        # both converge to ||b - A phi|| <= 1e-10 ||b||: their phi agree
        # within the solution's conditioning times that tolerance
        diff = float(np.max(np.abs(pa - pb)) / np.max(np.abs(pa)))
        print(f"  {c}x{c} cells: corrections {ra[0]} / {rb[0]}, relres {ra[1]:.2e} / {rb[1]:.2e}, "
              f"phi rel diff {diff:.2e}", flush=True)
        assert rb[1] <= 1e-10 and ra[1] <= 1e-10
        assert diff < 1e-8
        assert not xa.any() and not xb.any()
        # the no-allocation phi solves the Dirichlet-wall problem itself
        s2 = (d.L * d.L) / (d.LDe * d.LDe)
        b = -rho[1:-1, 1:-1] * s2
        lap = pb[2:, 1:-1] + pb[:-2, 1:-1] + pb[1:-1, 2:] + pb[1:-1, :-2] - 4 * pb[1:-1, 1:-1]
        assert np.linalg.norm(b - lap) <= 2e-10 * np.linalg.norm(b)
        assert not pb[0, :].any() and not pb[-1, :].any() and not pb[:, 0].any() and not pb[:, -1].any()


def test_noalloc_rejects_oversized_grid_and_other_solvers():
    st = workloads.bounded_512(ppc=1, cells=600)
    ctx = _lib.Context(st.domain, st.species)
    L = _lib.lib()
    it, err = ctypes.c_int(), ctypes.c_double()
    assert L.epic_solve_potential_noalloc(ctx.h, ctypes.byref(it), ctypes.byref(err)) == -3
    assert b"maximum" in L.epic_last_error(ctx.h)
    ctx.close()
    st = workloads.bounded_512(ppc=1, cells=64, solver=_lib.SOLVER_PCG)
    ctx = _lib.Context(st.domain, st.species)
    assert L.epic_solve_potential_noalloc(ctx.h, ctypes.byref(it), ctypes.byref(err)) == -1
    ctx.close()
