"""Stage timers (epic_timing_enable): every stage, or only the stages of an
EPIC_TIMING_ONLY mask; enabling clears the accumulated times."""
import pytest

from epic2d_mcc_amd import _lib
from harness import small_setup

pytestmark = pytest.mark.gpu


def test_stage_mask_records_only_the_listed_stages():
    print("[timing] stage mask", flush=True)
    st = small_setup(N=2000, Nx=32, Ny=32, precision=_lib.FP32, solver=_lib.SOLVER_PCG_MIXED, tol=1e-10)
    st.domain.tol_relative = 1
    st.domain.precond = _lib.PRECOND_SPECTRAL
    ctx = _lib.Context(st.domain, st.species)
    for s, sp in enumerate(st.species):
        ctx.load(s, sp.numparticle)
    ctx.initial_field()
    ctx.timing(True, stages=["push"])
    for ts in range(3):
        ctx.step(ts)
    t = ctx.timing_read()
    assert t["push"][1] >= 3 and t["push"][0] > 0
    assert all(v[1] == 0 and v[0] == 0 for k, v in t.items() if k != "push"), t
    ctx.timing(True)
    for ts in range(3, 6):
        ctx.step(ts)
    t = ctx.timing_read()
    for k in ("rho", "solve", "push"):
        assert t[k][1] >= 3 and t[k][0] > 0, (k, t)
    # the sine-transform solve launches the E-field after its convergence
    # check, inside the solve stage
    assert t["efield"][1] == 0, t
    ctx.timing(False)
    ctx.close()
