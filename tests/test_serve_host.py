"""The request stream and job descriptions of the serving loop (host code,
no GPU): arrivals are a Poisson process of the configured rate; job
parameters stay inside their ranges; a job's descriptors equal the
normalisation chain of main.cpp:151-261 (config.derive) for the same ini
values."""
import math

import numpy as np
import pytest

from epic2d_mcc_amd import _lib, config, workloads


def test_stream_is_reproducible_and_in_range():
    c = _lib.serve_config(seed=7)
    a = _lib.serve_make_jobs(c)
    b = _lib.serve_make_jobs(c)
    assert len(a) == len(b) and len(a) > 0
    for x, y in zip(a, b):
        assert bytes(x) == bytes(y)
    for j in a:
        assert 0 <= j.arrival_s < c.duration_s
        assert j.grid in (64, 128, 256, 512)
        assert c.ppc_min <= j.ppc <= c.ppc_max and c.steps_min <= j.steps <= c.steps_max
        assert c.te_min <= j.te <= c.te_max and c.ti_min <= j.ti <= c.ti_max
        assert not (j.two_stream and j.collisional)
        assert 0 <= j.vd <= c.vd_max
    t = np.array([j.arrival_s for j in a])
    assert np.all(np.diff(t) > 0)


def test_arrivals_are_poisson():
    """600-job stream: count within 4 sigma of rate*duration; exponential
    inter-arrival times (mean 1/rate, coefficient of variation ~1)."""
    c = _lib.serve_config(seed=3, rate_hz=20.0, duration_s=30.0)
    t = np.array([j.arrival_s for j in _lib.serve_make_jobs(c)])
    n = len(t)
    assert abs(n - 600) < 4 * math.sqrt(600)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.05) < 4 * 0.05 / math.sqrt(n)
    cv = gaps.std() / gaps.mean()
    assert 0.85 < cv < 1.15
    fr = np.mean([j.collisional for j in _lib.serve_make_jobs(c)])
    assert abs(fr - 0.5) < 4 * math.sqrt(0.25 / n)


def test_max_jobs_caps_the_stream():
    c = _lib.serve_config(seed=3, max_jobs=17)
    assert len(_lib.serve_make_jobs(c)) == 17


@pytest.mark.parametrize("two_stream,collisional", [(0, 0), (1, 0), (0, 1)])
def test_job_setup_matches_normalisation_chain(two_stream, collisional):
    c = _lib.serve_config(seed=1)
    j = _lib.ServeJob(index=4, grid=64, ppc=30, steps=10, two_stream=two_stream, collisional=collisional, te=2.5,
                      ti=0.05, vd=3.0 if two_stream else 0.0)
    dom, sp = _lib.serve_job_setup(j, c)
    n = 30 * 64 * 64
    mi = workloads.M_AR if collisional else workloads.M_ION
    if two_stream:
        specs = [config.SpeciesSpec("beam_plus", config.ME, n // 2, 2.5, -1, 1.0, 3.0, 0.0),
                 config.SpeciesSpec("ion", mi, n, 0.05, 1, 0.0),
                 config.SpeciesSpec("beam_minus", config.ME, n - n // 2, 2.5, -1, 1.0, -3.0, 0.0)]
    else:
        specs = [config.SpeciesSpec("electron", config.ME, n, 2.5, -1, 1.0),
                 config.SpeciesSpec("ion", mi, n, 0.05, 1, 0.0)]
    st = config.derive(64, 64, 0.0, specs, 1e13, 0.05, bc="pbc", solver=_lib.SOLVER_PCG_MIXED, tolerance=1e-10,
                       tol_relative=1, max_iteration=20000, precision=_lib.FP32, sort_interval=20,
                       precond=_lib.PRECOND_SPECTRAL, capacity_factor=1.25 if collisional else 1.0)
    for f, _ in _lib.DomainDesc._fields_:
        if f in ("seed", "reserved"):
            continue
        a, b = getattr(dom, f), getattr(st.domain, f)
        assert a == pytest.approx(b, rel=1e-15, abs=0), f
    assert len(sp) == len(st.species)
    for x, y in zip(sp, st.species):
        for f, _ in _lib.SpeciesDesc._fields_:
            a, b = getattr(x, f), getattr(y, f)
            if f == "name":
                assert bytes(a) == bytes(b)
            else:
                assert a == pytest.approx(b, rel=1e-15, abs=0), f
