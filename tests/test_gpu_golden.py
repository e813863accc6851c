"""The HIP path against the committed fixtures of tests/golden/: the
reference's own pcg() solutions (iterations of the reference recurrence,
phi within 1e-12), its Newton-CG loop (within 1e-9: workgroup dot products),
and 10 steps of the reference loop (fields after each step and the final
particles within 1e-12, PCG iteration counts equal)."""
import os

import numpy as np
import pytest

from epic2d_mcc_amd import _lib
from harness import rel, small_setup

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_loop_matches_fixture():
    print("[golden] 16x16 loop", flush=True)
    f = load("loop_16x16_pbc_cg.npz")
    st = small_setup(N=1000, Nx=16, Ny=16, bc="pbc", tol=1e-13)
    from harness import reference_load
    _, parts = reference_load(st)
    ctx = _lib.Context(st.domain, st.species)
    for s, p in enumerate(parts):
        ctx.upload(s, *p)
    ctx.initial_field()
    for ts in range(10):
        assert ctx.step(ts) == f["iters"][ts]
        for w in range(4):
            assert rel(ctx.get_field(w), f["fields"][ts, w]) < 1e-12, (ts, w)
    for s, key in ((0, "electrons"), (1, "ions")):
        g = ctx.download(s)
        for k in range(5):
            assert rel(g[k], f[key][k]) < 1e-12, (s, k)
    ctx.close()


def test_newton_cg_matches_reference_fixture():
    print("[golden] newton-cg", flush=True)
    f = load("ref_nrpcg_11x9.npz")
    st = small_setup(N=10, Nx=10, Ny=8, solver=_lib.SOLVER_NEWTON_CG)
    d = st.domain
    ctx = _lib.Context(d, st.species)
    ctx.set_field(0, f["rho"].reshape(d.nx, d.ny))
    ctx.solve()
    assert rel(ctx.get_field(1), f["phi"].reshape(d.nx, d.ny)) < 1e-9
    ctx.close()
