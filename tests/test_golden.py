"""The oracle against the committed fixtures of tests/golden/ (made by
tests/golden/make_fixtures.py): golden vectors of the reference's own pcg()
and Newton-CG loop on its dense templates, and oracle regression vectors (the
random stream, the cross sections, a 10-step run of the oracle's restatement
of the reference loop) that freeze the oracle's output.  These hold without
the reference tree."""
import os

import numpy as np
import pytest

from epic2d_mcc_amd import _lib
from harness import small_setup
from oracle_api import OracleSim, dp, oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


@pytest.mark.parametrize("name", ["ref_pcg_9x7.npz", "ref_pcg_12x12.npz"])
def test_oracle_pcg_equals_reference_pcg_fixture(name):
    f = load(name)
    nx, ny = int(f["nx"]), int(f["ny"])
    x = np.empty(nx * ny)
    oracle().orc_pcg(nx, ny, dp(np.ascontiguousarray(f["b"])), dp(x), int(f["max_iteration"]), float(f["tolerance"]))
    assert np.array_equal(x, f["phi"])


def test_oracle_newton_cg_equals_reference_fixture():
    f = load("ref_nrpcg_11x9.npz")
    st = small_setup(N=1, Nx=10, Ny=8, solver=_lib.SOLVER_NEWTON_CG)
    d = st.domain
    assert (d.nx, d.ny) == (int(f["nx"]), int(f["ny"])) and d.dx * d.dx == float(f["dx2"])
    o = OracleSim(d, st.species)
    o.set_grid(0, f["rho"].reshape(d.nx, d.ny))
    o.solve()
    assert np.array_equal(o.grid(1).ravel(), f["phi"])


def test_random_stream_fixture():
    u = np.empty(1000)
    oracle().orc_rnd_draws(0, 1000, dp(u))
    assert np.array_equal(u, load("mt19937_uniform_1000.npz")["u"])


def test_cross_section_fixture():
    f = load("cross_sections_64.npz")
    o = oracle()
    got = np.array([[o.orc_sigma_elastic(x), o.orc_sigma_excitation(x), o.orc_sigma_ionization(x)]
                    for x in f["energy_ev"]])
    assert np.array_equal(got, f["sigma"])


def test_oracle_loop_fixture():
    f = load("loop_16x16_pbc_cg.npz")
    st = small_setup(N=1000, Nx=16, Ny=16, bc="pbc", tol=1e-13)
    o = OracleSim(st.domain, st.species)
    for s in range(2):
        o.init_species(s, 0)
    o.initial_field()
    for ts in range(10):
        assert o.step() == f["iters"][ts]
        for w in range(4):
            assert np.array_equal(o.grid(w), f["fields"][ts, w]), (ts, w)
    assert np.array_equal(np.stack(o.particles(0)), f["electrons"])
    assert np.array_equal(np.stack(o.particles(1)), f["ions"])
