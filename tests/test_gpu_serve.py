"""The request loop on the GPU: a small Poisson stream of uneven jobs
(periodic, two-stream and collisional) is admitted against a memory budget
and a member cap, packed into one batch that changes between steps, and
every job's result is the same as the job run alone, bit for bit: ||phi||,
kinetic energies and collision counts.  Everything per node and per
particle is the same arithmetic in both runs, the transform products
included (tile GEMMs whose per-member result does not depend on how many
members share a launch)."""
import ctypes

import numpy as np
import pytest

from epic2d_mcc_amd import _lib, xsec

pytestmark = pytest.mark.gpu


def _mcc_template():
    return xsec.mcc_desc(xsec.synthetic_ar(), gas_density=2e20, gas_temp_K=300.0, enable=(1, 1, 1, 1, 1))


def _cfg(tmpl, **kw):
    base = dict(seed=11, rate_hz=40.0, duration_s=0.6, grids=(32, 48, 64), ppc_min=10, ppc_max=30, steps_min=10,
                steps_max=40, gas_density=2e20, budget_bytes=60e6, max_members=4, tune_steps=1, sort_interval=7,
                mcc_template=ctypes.pointer(tmpl))
    base.update(kw)
    return _lib.serve_config(**base)


def _alone(job, cfg, tmpl):
    """The job by itself: same set-up as the request loop's admission."""
    dom, sp = _lib.serve_job_setup(job, cfg)
    B = _lib.Batch()
    mid, ctx = B.add(dom, sp)
    for s, d in enumerate(sp):
        ctx.load(s, d.numparticle)
        ctx.sort(s)
    if job.collisional:
        m = _lib.MccDesc.from_buffer_copy(tmpl)
        m.electron_sp, m.ion_sp = 0, 1
        m.gas_density = cfg.gas_density
        m.gas_mass = xsec.M_AR
        ctx.mcc_setup(m)
    B.init(mid)
    while B.member_steps(mid) < job.steps:
        B.step(1)
    ke = []
    for s in range(3):
        k = ctx.ke(s, 0) if s < len(sp) else [0.0, 0.0, 0.0]
        ke.append(float(k[0]) + float(k[1]) + float(k[2]) if s < len(sp) else 0.0)
    s2 = 0.0
    for v in ctx.get_field(1).ravel().tolist():  # the loop's summation order
        s2 += v * v
    coll = int(sum(ctx.mcc_stats()["count"])) if job.collisional else 0
    B.close()
    return ke, float(np.sqrt(s2)), coll


def test_stream_results_equal_jobs_alone():
    tmpl = _mcc_template()
    cfg = _cfg(tmpl)
    rep, jobs = _lib.serve_run(cfg)
    print(f"[serve] {rep}", flush=True)
    assert rep["jobs"] == rep["jobs_done"] == len(jobs) >= 10
    assert rep["max_members"] <= cfg.max_members
    assert rep["peak_bytes"] <= cfg.budget_bytes
    assert rep["graph_captures"] >= 2
    assert rep["particle_steps"] == pytest.approx(sum(j.particles * j.steps for j in jobs))
    kinds = {(j.two_stream, j.collisional) for j in jobs}
    assert (0, 1) in kinds and (1, 0) in kinds
    for j in jobs:
        assert j.arrival_s <= j.admit_s <= j.done_s
    lat = np.array([j.done_s - j.arrival_s for j in jobs])
    assert rep["latency_p50_s"] == pytest.approx(np.percentile(lat, 50))
    assert rep["latency_p99_s"] == pytest.approx(np.percentile(lat, 99))
    for j in jobs:
        ke, phi, coll = _alone(j, cfg, tmpl)
        print(f"  job {j.index} grid {j.grid} ppc {j.ppc} steps {j.steps} ts {j.two_stream} coll {j.collisional}: "
              f"phi {j.phi_l2:.6e} / {phi:.6e}, collisions {j.collisions} / {coll}")
        assert j.phi_l2 == phi
        assert list(j.ke) == ke
        assert j.collisions == coll
        if j.collisional:
            assert coll > 0


def test_budget_forces_queueing():
    """A budget of about two jobs: later arrivals wait for memory."""
    tmpl = _mcc_template()
    cfg = _cfg(tmpl, budget_bytes=30e6, max_members=64, rate_hz=80.0, duration_s=0.3, tune_steps=0)
    rep, jobs = _lib.serve_run(cfg)
    print(f"[serve] {rep}", flush=True)
    assert rep["jobs_done"] == rep["jobs"]
    assert rep["peak_bytes"] <= cfg.budget_bytes
    assert rep["wait_p99_s"] > 0
    assert max(j.admit_s - j.arrival_s for j in jobs) > 0


def test_failed_job_is_evicted_and_the_rest_are_undisturbed():
    """One job of the stream is admitted with a non-finite ion temperature
    (fault_job), so its ions carry NaN velocities: the reference would throw
    at their next bounds-checked gather (linalg.h:473-480); here the push
    drops them and counts EPIC_ERRC_NAN, and the loop's health check evicts
    the job within health_interval steps and reports it failed.  Every other
    job completes and equals its run alone, as in a mix without the failure."""
    tmpl = _mcc_template()
    probe = _cfg(tmpl)
    cand = [j for j in _lib.serve_make_jobs(probe) if not j.collisional]
    victim = cand[1].index
    cfg = _cfg(tmpl, fault_job=victim + 1, health_interval=4)
    rep, jobs = _lib.serve_run(cfg)
    print(f"[serve] fault job {victim}: {rep}", flush=True)
    assert rep["jobs_failed"] == 1 and rep["jobs_done"] == rep["jobs"] - 1
    bad = jobs[victim]
    assert bad.failed == 1 + _lib.ERRC_NAN, bad.failed
    assert bad.errors[_lib.ERRC_NAN] > 0
    # evicted at the first health check after its first push, long before its steps ran out
    assert bad.done_s > bad.admit_s
    assert rep["particle_steps"] == pytest.approx(sum(j.particles * j.steps for j in jobs if not j.failed))
    for j in jobs:
        if j.index == victim:
            continue
        assert j.failed == 0 and not any(j.errors[k] for k in (_lib.ERRC_NAN, _lib.ERRC_CAPACITY))
        ke, phi, coll = _alone(j, cfg, tmpl)
        assert j.phi_l2 == phi, (j.index, j.phi_l2, phi)
        assert list(j.ke) == ke
        assert j.collisions == coll
