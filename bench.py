"""Benchmark: particle-steps/s of the per-timestep PIC loop on one MI355X.

Workload (default): the four-species bounded plasma -- 512x512 cells,
absorbing walls with Dirichlet phi = 0, electrons + positive ions + negative
ions (alpha 0.1) + negative beam (beta 0.05, 10 v_th,e drift), 100
particles/cell/species = 104,857,600 particles, FP32 particles,
mixed-precision PCG (FP32 iterations, FP64 refinement) to 1e-10 relative,
counting sort every 20 steps.  A step is one full pass of the loop: scatter,
rho, Poisson, E, gather+push (+fused next deposit), absorption.

Prints one JSON line (see the keys below).  --gpus N > 1 exits: this build
runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from epic2d_mcc_amd import _lib, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes the push kernel moves per particle, from rocprofv3 --pmc on this
# workload (profiles/push_f32_multi_kernel_bounded512_pmc.md): FETCH_SIZE x 2
# + WRITE_SIZE x 1 per dispatch (the factors a plain 1 GiB device copy reads
# back as), divided by the 104,857,600 particles of a dispatch.
PUSH_TRAFFIC_BYTES_PER_PARTICLE = (881959.4 * 2 + 1747068.1) * 1024 / 104857600


def gpu_run(st, steps, warmup, solver_name, mcc=None):
    ctx = _lib.Context(st.domain, st.species)
    nsp = len(st.species)
    for s, sp in enumerate(st.species):
        ctx.load(s, sp.numparticle, vz_thermal=False)
    for s in range(nsp):
        ctx.sort(s)
    if mcc is not None:
        ctx.mcc_setup(mcc)
    ctx.initial_field()
    ctx.sync()
    print("[bench] set up: particles loaded and sorted, initial field", file=sys.stderr, flush=True)
    ts = 0
    for _ in range(warmup):
        ctx.step(ts)
        ts += 1
    ctx.sync()
    print(f"[bench] {warmup} warm-up steps done", file=sys.stderr, flush=True)
    n0 = sum(ctx.count(s) for s in range(nsp))
    per_sp0 = [ctx.count(s) for s in range(nsp)]
    # the timed steps record only the push's event pair (the roofline's
    # launch time); each further pair costs the stream a few microseconds
    ctx.timing(True, stages=["push"])
    iters = []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        iters.append(ctx.step(ts))
        ts += 1
    ctx.sync()
    t1 = time.perf_counter()
    print(f"[bench] {steps} timed steps done", file=sys.stderr, flush=True)
    tm = ctx.timing_read()
    n1 = sum(ctx.count(s) for s in range(nsp))
    per_sp1 = [ctx.count(s) for s in range(nsp)]
    coll = [int(v) for v in ctx.mcc_stats()["count"]] if mcc is not None else None
    # stage breakdown from the same number of further steps, every stage
    # timed (not part of the measurement)
    ctx.timing(True)
    for _ in range(steps):
        ctx.step(ts)
        ts += 1
    ctx.sync()
    print(f"[bench] {steps} stage-timed steps done", file=sys.stderr, flush=True)
    stages = ctx.timing_read()
    ctx.timing(False)
    errs = ctx.errors()
    ctx.close()
    return dict(t=t1 - t0, n0=n0, n1=n1, per_sp0=per_sp0, per_sp1=per_sp1, timing=tm, stages=stages, iters=iters,
                errors=errs.tolist(), collisions=coll)


def host_info():
    """The host the CPU baseline runs on: logical CPUs, the ones this process
    may use, the OpenMP thread setting and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(usable, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else usable
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "omp_num_threads": omp, "model": model,
            "threads": threads}


def cpu_baseline(st, sample_frac=1.0, steps=2, gas=None, threads=1, dense_ref=False):
    """The oracle (FP64 CPU restatement of the reference loop) on a bounded
    sample of the same workload: the full grid, `steps` steps, with
    `sample_frac` of the particles.  Particle stages (scatter, push) are
    timed on the sample and scaled to the full count; the field stages
    (rho, Poisson, E) are per step and counted once.
    threads: 1 is the reference's serial loop; more runs the oracle's
    all-cores variant (per-thread particle ranges, threaded operator
    products).
    Poisson: dense_ref times the reference's own dense pcg() with its
    threaded Matrix x vec, compiled from the reference's headers into
    oracle/_ref (pcgsolver builds the dense A, and pcg its dense
    preconditioner, on every solve: field.cpp:221-232, solvers.h:75-76);
    otherwise the matrix-free restatement of the same pcg() recurrence (dense
    PCG at 263,169 unknowns needs 1.1 TB).  Absolute tolerance = the
    workload's relative one times ||b||."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_api import OracleSim, oracle, ref_pcg_lib
    oracle().orc_set_threads(threads)
    ref = ref_pcg_lib() if dense_ref else None
    d = st.domain
    dom = _lib.DomainDesc.from_buffer_copy(d)
    dom.solver = _lib.SOLVER_PCG
    o = OracleSim(dom, st.species, gas=gas)
    if gas is not None:
        oracle().orc_mcc_build_tables(o.h)
    rng = np.random.default_rng(7)
    Lx = (d.nx - 1) * d.dx
    Ly = (d.ny - 1) * d.dy
    nsp = len(st.species)
    n_full = sum(int(sp.numparticle) for sp in st.species)
    n_samp = 0
    for s, sp in enumerate(st.species):
        n = max(1, int(sp.numparticle * sample_frac))
        n_samp += n
        sig = math.sqrt(sp.temp * 1.602176565e-19 / sp.mass) / d.vel_norm
        o.set_particles(s, rng.random(n) * Lx, rng.random(n) * Ly, rng.normal(0, sig, n) + sp.vsx,
                        rng.normal(0, sig, n) + sp.vsy, np.zeros(n))
    for s in range(nsp):
        o.scatter(s)
    o.compute_rho()
    rho = o.grid(0)
    s2 = (d.L * d.L) / (d.LDe * d.LDe)
    b = -rho[1:-1, 1:-1] * d.dx * d.dx * s2
    tol_abs = d.tolerance * float(np.sqrt((b * b).sum()))
    o.set_tolerance(tol_abs)

    def dense_solve():
        import ctypes
        rhom = -o.grid(0) * d.dx * d.dx * s2
        rhom[0, :] = rhom[-1, :] = rhom[:, 0] = rhom[:, -1] = 0.0
        rhom = np.ascontiguousarray(rhom.ravel())
        phi = np.empty_like(rhom)
        err = ctypes.c_double()
        D = ctypes.POINTER(ctypes.c_double)
        ref.ref_pcg(d.nx, d.ny, rhom.ctypes.data_as(D), phi.ctypes.data_as(D), d.max_iteration, tol_abs,
                    ctypes.byref(err))
        o.set_grid(1, phi.reshape(d.nx, d.ny))
        return -1

    t = time.perf_counter
    tp = tf = 0.0
    its = []
    start = t()
    for _ in range(steps):
        t0 = t()
        for s in range(nsp):
            o.scatter(s)
        t1 = t()
        o.compute_rho()
        if ref is not None:
            it = dense_solve()
        else:
            it, _ = o.solve()
        o.efield()
        t2 = t()
        for s in range(nsp):
            o.push(s)
        if gas is not None:
            o.collide_direct()
        t3 = t()
        tp += (t1 - t0) + (t3 - t2)
        tf += t2 - t1
        its.append(it)
    wall = t() - start
    oracle().orc_set_threads(1)
    scale = n_full / n_samp
    t_particles = tp * scale / steps
    t_fields = tf / steps
    t_step = t_particles + t_fields
    return dict(value=n_full / t_step, t=wall, iters=its, npart=n_full, nsamp=n_samp, steps=steps,
                t_particles=t_particles, t_fields=t_fields, threads=threads, dense_ref=ref is not None)


def cpu_baseline_entry(st, args, gas=None, note="", dense_ref=False):
    """cpu_baseline on one thread (the reference's serial loops: the
    reported baseline) and on every usable core (the all-cores variant)."""
    hi = host_info()
    cb = cpu_baseline(st, sample_frac=args.cpu_sample, steps=args.cpu_steps, gas=gas, dense_ref=dense_ref)
    solver = ("the reference's own dense pcg() (oracle/_ref, its Matrix x vec on %d hardware threads)"
              % (os.cpu_count() or 1) if cb["dense_ref"] else "the matrix-free pcg() restatement")
    # the reference's dense Matrix x vec runs on hardware_concurrency() threads
    # (linalg.h:397-424); everything else here is one thread
    cores = (os.cpu_count() or 1) if cb["dense_ref"] else 1
    out = {"value": cb["value"], "unit": "particle-steps/s", "cores": cores, "kind": "port",
           "sample": "%d steps of the same workload by the FP64 oracle (restatement of the reference loop) on one "
                     "thread (the Poisson matrix products excepted when it is the reference's dense pcg()): "
                     "%d of the %d particles (particle stages scaled x%.2f to the full count), %.2f s "
                     "particle + %.2f s field stages per step, Poisson by %s (%s iterations)%s; %.1f s of CPU "
                     "work" % (cb["steps"], cb["nsamp"], cb["npart"], cb["npart"] / cb["nsamp"], cb["t_particles"],
                               cb["t_fields"], solver, cb["iters"], note, cb["t"]),
           "host": {k: hi[k] for k in ("nproc", "usable_cpus", "omp_num_threads", "model")}}
    if hi["threads"] > 1 and not args.no_all_cores:
        ca = cpu_baseline(st, sample_frac=args.cpu_sample, steps=args.cpu_steps, gas=gas, threads=hi["threads"],
                          dense_ref=dense_ref)
        out["all_cores"] = {"value": ca["value"], "cores": hi["threads"], "kind": "port",
                            "t_particles": ca["t_particles"], "t_fields": ca["t_fields"],
                            "sample": "the same sample with the oracle's OpenMP variant on %d threads "
                                      "(per-thread particle ranges, threaded operator products)" % hi["threads"]}
    return out


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size (torch), timed with events:
    the practical HBM rate and the reference point for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    moved = 2 * nbytes
    del src, dst
    torch.cuda.empty_cache()
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB/s": moved / ms / 1e6}


def ensemble_run(sims, steps, warmup, tune_steps=2, eager_steps=10):
    """The batched ensemble: every member in one batch, kernel shapes tuned on
    real steps, then `steps` timed graph replays."""
    B = _lib.Batch()
    ids = []
    t0 = time.perf_counter()
    for st in sims:
        mid, ctx = B.add(st.domain, st.species)
        for s, sp in enumerate(st.species):
            ctx.load(s, sp.numparticle)
            ctx.sort(s)
        B.init(mid)  # the initial field and half-step rewind inside the first batched step
        ids.append(mid)
    B.sync()
    t_setup = time.perf_counter() - t0
    print(f"[bench] ensemble: {len(sims)} members set up in {t_setup:.1f} s", file=sys.stderr, flush=True)
    tuning = None
    if tune_steps > 0:
        tuning = B.tune(tune_steps)
        print("[bench] ensemble: kernel shapes tuned", file=sys.stderr, flush=True)
    B.step(warmup)
    B.sync()

    def count():
        return sum(B.members[m].count(s) for m, st in zip(ids, sims) for s in range(len(st.species)))

    n0 = count()
    B.timing(True)
    t0 = time.perf_counter()
    B.step(steps)
    B.sync()
    t1 = time.perf_counter()
    stats = B.stats()
    n1 = count()
    # stage breakdown from a few eager steps (a graph replay is one event pair)
    print("[bench] ensemble: timed steps done", file=sys.stderr, flush=True)
    B.configure(use_graph=False)
    B.step(1)  # the first eager step after the graph replays is not timed
    B.sync()
    B.timing(True)
    B.step(eager_steps)
    B.sync()
    eager = B.stats()
    stage = {k: v / eager_steps for k, v in zip(("rho_solve", "efield", "push", "collide"), eager["stage_ms"])}
    errs = np.zeros(_lib.NERR, dtype=np.int64)
    for m in ids:
        errs += B.members[m].errors()
    B.close()
    return dict(t=t1 - t0, n0=n0, n1=n1, stats=stats, tuning=tuning, t_setup=t_setup, errors=errs.tolist(),
                stage_ms=stage)


def sequential_run(sims, steps, precision, solver):
    """The same simulations one at a time, each on its own context (the
    plain single-simulation path), `steps` timed steps each."""
    part_steps = 0.0
    t_tot = 0.0
    iters = []
    for st in sims:
        dom = _lib.DomainDesc.from_buffer_copy(st.domain)
        dom.precision = precision
        dom.solver = solver
        ctx = _lib.Context(dom, st.species)
        for s, sp in enumerate(st.species):
            ctx.load(s, sp.numparticle)
            ctx.sort(s)
        ctx.initial_field()
        ctx.step(0)
        ctx.sync()
        n = sum(ctx.count(s) for s in range(len(st.species)))
        t0 = time.perf_counter()
        for k in range(steps):
            iters.append(ctx.step(k + 1))
        ctx.sync()
        t_tot += time.perf_counter() - t0
        part_steps += n * steps
        ctx.close()
    return dict(value=part_steps / t_tot, t=t_tot, iters_mean=float(np.mean(iters)) if iters else 0.0)


def main_ensemble(args):
    sims = workloads.ensemble_64(n=args.members)
    r = ensemble_run(sims, args.steps, args.warmup, tune_steps=args.tune_steps)
    npart = 0.5 * (r["n0"] + r["n1"])
    value = npart * args.steps / r["t"]
    tu = r["tuning"]
    push_ms = r["stage_ms"]["push"]
    achieved = 32 * npart / (push_ms / 1e3) / 1e9
    grids = {}
    for st in sims:
        grids[st.info["grid"]] = grids.get(st.info["grid"], 0) + 1
    out = {
        "metric": "particle-steps/s, batched ensemble (whole batch, one GPU)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["t"] / args.steps,
        "higher_is_better": True,
        "dtype": "fp32 particles, fp64-refined Poisson",
        "data": "synthetic (Philox-seeded uniform positions, Maxwellian velocities)",
        "config": {"workload": "%d periodic electron-ion sims, grids %s (cells: count), 20-200 ppc, Te 0.5-5 eV, "
                               "two-stream +-(0-5) v_th,e on odd members, sort every 20 steps, step in a HIP graph"
                               % (len(sims), grids),
                   "particles": int(npart), "tuning": tu, "graph_replays": r["stats"]["graph_replays"],
                   "relres_max": r["stats"]["relres_max"], "setup_s": r["t_setup"],
                   "eager_stage_ms_per_step": r["stage_ms"],
                   "device_bytes": r["stats"]["device_bytes"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": "push_batch_kernel (all members' species)",
                     "bytes_per_particle_step": 32, "push_ms_per_step": push_ms},
        "device_errors": r["errors"],
    }
    if not args.no_baseline:
        seq32 = sequential_run(sims, args.baseline_steps, _lib.FP32, _lib.SOLVER_PCG_MIXED)
        seq64 = sequential_run(sims, args.baseline_steps, _lib.FP64, _lib.SOLVER_PCG)
        out["sequential_fp32"] = {"value": seq32["value"], "steps_per_sim": args.baseline_steps,
                                  "speedup_of_batch": value / seq32["value"]}
        out["sequential_fp64_pcg"] = {"value": seq64["value"], "steps_per_sim": args.baseline_steps,
                                      "pcg_iterations_mean": seq64["iters_mean"],
                                      "speedup_of_batch": value / seq64["value"]}
    print(json.dumps(out))


def fp64_sequential_jobs(cfg, jobs, tmpl, steps):
    """The plain path: each sampled job alone on an FP64 context with the
    reference's PCG (Jacobi, FP64), `steps` timed steps after its initial
    field; returns particle-steps/s over the sample."""
    from epic2d_mcc_amd import xsec
    part_steps = 0.0
    t_tot = 0.0
    for j in jobs:
        dom, sp = _lib.serve_job_setup(j, cfg, precision=_lib.FP64, solver=_lib.SOLVER_PCG)
        ctx = _lib.Context(dom, sp)
        for s, d in enumerate(sp):
            ctx.load(s, d.numparticle)
            ctx.sort(s)
        if j.collisional:
            m = _lib.MccDesc.from_buffer_copy(tmpl)
            m.electron_sp, m.ion_sp = 0, 1
            m.gas_density = cfg.gas_density
            m.gas_mass = xsec.M_AR
            ctx.mcc_setup(m)
        ctx.initial_field()
        ctx.step(0)
        ctx.sync()
        n = sum(ctx.count(s) for s in range(len(sp)))
        t0 = time.perf_counter()
        for k in range(steps):
            ctx.step(k + 1)
        ctx.sync()
        t_tot += time.perf_counter() - t0
        part_steps += n * steps
        ctx.close()
        print(f"[bench] fp64 sample job {j.index}: {n} particles", file=sys.stderr, flush=True)
    return part_steps / t_tot


def main_serve(args):
    import ctypes
    from epic2d_mcc_amd import xsec
    tmpl = xsec.mcc_desc(xsec.synthetic_ar(), gas_density=1e21, gas_temp_K=300.0, enable=(1, 1, 1, 1, 1))
    sc = args.serve_steps_scale
    cfg = _lib.serve_config(duration_s=args.serve_duration, steps_min=max(1, round(100 * sc)),
                            steps_max=max(1, round(2000 * sc)), mcc_template=ctypes.pointer(tmpl), progress_s=20.0)
    t0 = time.perf_counter()
    rep, jobs = _lib.serve_run(cfg)
    t_run = time.perf_counter() - t0
    sample = jobs[::max(1, len(jobs) // args.serve_baseline_jobs)][:args.serve_baseline_jobs]
    base = fp64_sequential_jobs(cfg, sample, tmpl, args.baseline_steps) if not args.no_baseline else None
    grids = {}
    for j in jobs:
        grids[j.grid] = grids.get(j.grid, 0) + 1
    out = {
        "metric": "aggregate particle-steps/s of the request loop (batched FP32 + FP64-refined) and p50/p99 "
                  "time to completion",
        "value": rep["throughput"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "fp32 particles, fp64-refined Poisson",
        "data": "synthetic request stream (seeded)",
        "config": {"workload": "Poisson arrivals %.0f/s for %.0f s (%d jobs), grids %s, %d-%d ppc, %d-%d steps, "
                               "half collisional (argon, synthetic tables), half of the rest two-stream; "
                               "budget %.0f GB" % (cfg.rate_hz, cfg.duration_s, rep["jobs"], grids, cfg.ppc_min,
                                                   cfg.ppc_max, cfg.steps_min, cfg.steps_max,
                                                   cfg.budget_bytes / 1e9),
                   "steps_scale": sc},
        "latency_p50_s": rep["latency_p50_s"],
        "latency_p99_s": rep["latency_p99_s"],
        "jobs_failed": rep["jobs_failed"],
        "report": rep,
        "run_s": t_run,
    }
    if base:
        out["fp64_sequential"] = {"value": base, "jobs_sampled": len(sample), "steps_per_job": args.baseline_steps,
                                  "speedup_of_loop": rep["throughput"] / base}
    print(json.dumps(out))


def main_single(args, st, label, dtype, bytes_ps, kernel, mcc=None, gas=None, cpu_note="", dense_ref=False):
    """One simulation of the given set-up: whole-step particle-steps/s, the
    fused push's roofline, and the oracle on the host beside it."""
    r = gpu_run(st, args.steps, args.warmup, label, mcc=mcc)
    nsp = len(st.species)
    value = 0.5 * (r["n0"] + r["n1"]) * args.steps / r["t"]
    push_ms, push_calls = r["timing"]["push"]
    avg_launch_s = push_ms / max(push_calls, 1) / 1e3
    # one launch per pushed species, or one for all of them (the FP32 hot
    # kernel): particles per launch from the particle-steps over the calls
    parts_per_launch = 0.5 * (sum(r["per_sp0"]) + sum(r["per_sp1"])) * args.steps / max(push_calls, 1)
    achieved = bytes_ps * parts_per_launch / avg_launch_s / 1e9
    stage = {k: round(v[0] / max(args.steps, 1), 4) for k, v in r["stages"].items() if v[1]}
    out = {
        "metric": "particle-steps/s per GPU (+% HBM roofline)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["t"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (Philox-seeded uniform positions, Maxwellian velocities)",
        "config": {"workload": label, "particles_start": r["n0"], "particles_end": r["n1"],
                   "poisson_iterations_mean": float(np.mean(r["iters"])) if r["iters"] else 0,
                   "stage_ms_per_step": stage, "collisions": r["collisions"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel,
                     "bytes_per_particle_step": bytes_ps, "avg_launch_ms": avg_launch_s * 1e3,
                     "particles_per_launch": parts_per_launch},
        "device_errors": r["errors"],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_entry(st, args, gas=gas, note=cpu_note, dense_ref=dense_ref)
    print(json.dumps(out))


def fp64_plain_leg(args):
    """The plain FP64 path on the headline workload (FP64 particles, the
    reference's Jacobi PCG on the Dirichlet-wall operator to 1e-10): a few
    steps, for the deposit+push fraction of HBM peak (64 B per
    particle-step, event-timed push launches) and its particle-steps/s."""
    st = workloads.bounded_512(precision=_lib.FP64, solver=_lib.SOLVER_PCG)
    r = gpu_run(st, args.fp64_steps, 1, "pcg")
    push_ms, push_calls = r["timing"]["push"]
    avg_launch_s = push_ms / max(push_calls, 1) / 1e3
    parts_per_launch = 0.5 * (sum(r["per_sp0"]) + sum(r["per_sp1"])) * args.fp64_steps / max(push_calls, 1)
    achieved = 64 * parts_per_launch / avg_launch_s / 1e9
    return {"workload": "the same 512x512 four-species workload in FP64 (plain path: FP64 particles, the "
                        "reference PCG, Jacobi, FP64, to 1e-10 relative)",
            "steps": args.fp64_steps, "value": 0.5 * (r["n0"] + r["n1"]) * args.fp64_steps / r["t"],
            "ms_per_step": 1e3 * r["t"] / args.fp64_steps,
            "pcg_iterations_mean": float(np.mean(r["iters"])) if r["iters"] else 0,
            "push_kernel": "push_f64_kernel<2,true,false> (gather + push + walls + deposit, one launch per species)",
            "push_avg_launch_ms": avg_launch_s * 1e3, "push_achieved_GBs": achieved,
            "push_frac": achieved / HBM_PEAK_GBS, "device_errors": r["errors"]}


def serve_leg(args):
    """A short request stream through the serving loop (continuous batching,
    FP32 + FP64-refined) and the plain FP64 path one job at a time on a
    sample of the same jobs: p50/p99 time to completion, aggregate
    particle-steps/s, and the loop's speed-up over the FP64 path."""
    import ctypes
    from epic2d_mcc_amd import xsec
    tmpl = xsec.mcc_desc(xsec.synthetic_ar(), gas_density=1e21, gas_temp_K=300.0, enable=(1, 1, 1, 1, 1))
    sc = args.serve_steps_scale_leg
    cfg = _lib.serve_config(duration_s=args.serve_duration_leg, steps_min=max(1, round(100 * sc)),
                            steps_max=max(1, round(2000 * sc)), mcc_template=ctypes.pointer(tmpl))
    t0 = time.perf_counter()
    rep, jobs = _lib.serve_run(cfg)
    t_run = time.perf_counter() - t0
    k = max(1, args.serve_baseline_jobs_leg)
    sample = jobs[::max(1, len(jobs) // k)][:k]
    base = fp64_sequential_jobs(cfg, sample, tmpl, args.baseline_steps)
    return {"workload": "Poisson arrivals 20/s for %.0f s (%d jobs), grids 64^2-512^2, 50-400 ppc, %d-%d steps "
                        "(the 100-2000 range x%.2f), half collisional, half of the rest two-stream, 200 GB budget"
                        % (cfg.duration_s, rep["jobs"], cfg.steps_min, cfg.steps_max, sc),
            "value": rep["throughput"], "unit": "particle-steps/s", "latency_p50_s": rep["latency_p50_s"],
            "latency_p99_s": rep["latency_p99_s"], "jobs_done": rep["jobs_done"], "jobs_failed": rep["jobs_failed"],
            "max_members": rep["max_members"], "run_s": t_run,
            "fp64_plain_path": {"value": base, "jobs_sampled": len(sample), "steps_per_job": args.baseline_steps},
            "speedup_over_fp64_plain_path": rep["throughput"] / base}


def main_ops(args):
    """Device time of each operator of the loop on its own (rows the step
    fuses or skips), on the headline workload: 512x512 cells, 4 x 26.2M FP32
    particles; SOR and the FP64 push on the same grid in FP64.  Averages of
    `reps` calls between synchronisations (host clock)."""
    reps = 10

    def timed(fn, n=reps):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        ctx.sync()
        return 1e6 * (time.perf_counter() - t0) / n

    st = workloads.bounded_512()
    ctx = _lib.Context(st.domain, st.species)
    nsp = len(st.species)
    ops = {}
    n0 = st.species[0].numparticle
    t0 = time.perf_counter()
    ctx.load(0, n0)
    ctx.sync()
    ops["load (Init, Philox) per 26.2M species, us"] = 1e6 * (time.perf_counter() - t0)
    for s in range(1, nsp):
        ctx.load(s, st.species[s].numparticle)
    for s in range(nsp):
        ctx.sort(s)
    ctx.initial_field()
    for ts in range(20):
        ctx.step(ts)
    ops["scatter (stand-alone deposit) per species, us"] = timed(lambda: ctx.scatter(0))
    ops["compute_rho, us"] = timed(ctx.compute_rho)
    ops["efield, us"] = timed(ctx.efield)
    ops["rewind per species, us"] = timed(lambda: ctx.rewind(0))
    ops["push without deposit per species, us"] = timed(lambda: ctx.push(0))
    ops["sort per species (electrons, 20 steps after the last), us"] = timed(lambda: ctx.sort(0), 3)
    ops["compact per species (nothing to remove), us"] = timed(lambda: ctx.compact(0), 3)
    ops["kinetic energy per species, us"] = timed(lambda: ctx.ke(0, 0))
    ops["momentum per species, us"] = timed(lambda: ctx.momentum(0, 0))
    ops["potential energy, us"] = timed(lambda: ctx.pe(0))
    for k in range(8):
        e = _lib.EmitterDesc(x0=0.0, y0=0.0, x1=0.0, y1=(st.domain.ny - 1) * st.domain.dy, temp=1.0,
                             numparticle=1000, vdx=1.0, vdy=0.0, species_idx1=0, species_idx2=1)
        ctx.add_emitter(e)
    step = [100]

    def inj():
        ctx.inject(step[0])
        step[0] += 1
    ops["inject, 8 emitters x 1000 pairs, us"] = timed(inj)
    # internal conductors: a 64x64-node rectangle and a ring of 8 electrodes
    # of radius 20 (the generic particle kernel tests each new position)
    ctx.set_conductors(_lib.conductors([
        {"kind": _lib.COND_RECT, "min_x": 100, "max_x": 164, "min_y": 100, "max_y": 164, "voltage": 1.0},
        {"kind": _lib.COND_CIRCULAR, "x_center": 320, "y_center": 320, "electrode_number": 8,
         "grid_radius": 120.0, "electrode_radius": 20.0, "voltage": -1.0}]))
    ops["push without deposit per species, 2 conductors (9 shapes), us"] = timed(lambda: ctx.push(0))
    tsn = [200]

    def step():
        ctx.step(tsn[0])
        tsn[0] += 1
    ops["fused step with 2 conductors (all 4 species), us"] = timed(step, 20)
    ctx.set_conductors(_lib.conductors([]))
    ops["fused step without conductors (all 4 species), us"] = timed(step, 20)
    ctx.close()
    # SOR (red-black, gssolver's problem) on the same grid, FP64, 1e-6
    st64 = workloads.bounded_512(ppc=1, precision=_lib.FP64, solver=_lib.SOLVER_GS, tol=1e-6)
    ctx = _lib.Context(st64.domain, st64.species)
    for s, sp in enumerate(st64.species):
        ctx.load(s, sp.numparticle)
    for s in range(len(st64.species)):
        ctx.scatter(s)
    ctx.compute_rho()
    t0 = time.perf_counter()
    it, _ = ctx.solve()
    ctx.sync()
    ops["SOR solve 513^2 FP64 to 1e-6 RMS residual, us"] = 1e6 * (time.perf_counter() - t0)
    ops["SOR iterations"] = it
    ctx.close()
    out = {"metric": "per-operator device time (us), headline workload", "value": None, "unit": "us",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
           "config": {"workload": "bounded 512x512, 4 x 26,214,400 FP32 particles; SOR in FP64"},
           "ops": {k: (round(v, 1) if isinstance(v, float) else v) for k, v in ops.items()}}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--solver", default=None, choices=["pcg_mixed", "spectral", "pcg", "gs"],
                    help="bounded512 Poisson solver (default: pcg_mixed in fp32, the reference PCG in fp64)")
    ap.add_argument("--ppc", type=int, default=100)
    ap.add_argument("--cpu-sample", type=float, default=1.0)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="bounded512",
                    choices=["bounded512", "periodic64", "mcc256", "ensemble64", "serve", "ops"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"],
                    help="bounded512: fp64 runs the plain FP64 path (FP64 particles, the reference PCG)")
    ap.add_argument("--serve-steps-scale", type=float, default=1.0,
                    help="multiplies the jobs' 100-2000 step range")
    ap.add_argument("--serve-duration", type=float, default=30.0)
    ap.add_argument("--serve-baseline-jobs", type=int, default=12)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--tune-steps", type=int, default=2,
                    help="ensemble: timed steps per kernel-shape candidate (0: no tuning, the default shapes)")
    ap.add_argument("--sort-interval", type=int, default=20, help="bounded512: steps between counting sorts")
    ap.add_argument("--no-all-cores", action="store_true", help="skip the all-cores CPU variant")
    ap.add_argument("--no-fp64-leg", action="store_true", help="headline: skip the plain FP64 path leg")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-serve-leg", action="store_true", help="headline: skip the short serving leg")
    ap.add_argument("--serve-duration-leg", type=float, default=5.0)
    ap.add_argument("--serve-steps-scale-leg", type=float, default=0.1)
    ap.add_argument("--serve-baseline-jobs-leg", type=int, default=4)
    ap.add_argument("--baseline-steps", type=int, default=3)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also time a plain 1 GiB device copy (the reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU (the per-step loop of one simulation); "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble64":
        main_ensemble(args)
        return
    if args.workload == "ops":
        main_ops(args)
        return
    if args.workload == "serve":
        main_serve(args)
        return
    if args.workload == "periodic64":
        st = workloads.periodic_64()
        main_single(args, st, "periodic 64x64 cells, electron + ion (409,600 each), FP64 particles, reference PCG "
                              "(Jacobi, FP64) to 1e-8 relative", "fp64", 64,
                    "particle_kernel<double,2,PUSH,whole-grid LDS,deposit> (gather + push + wrap + deposit)",
                    dense_ref=True)
        return
    if args.workload == "mcc256":
        from epic2d_mcc_amd import xsec
        st = workloads.mcc_256()
        mcc = xsec.mcc_desc(xsec.synthetic_ar(), gas_density=1e21, gas_temp_K=300.0, enable=(1, 1, 1, 1, 1))
        main_single(args, st, "periodic 256x256 cells, electron + Ar+ at 1e15 m^-3, 200 ppc each (26,214,400), "
                              "FP32, Ar neutrals 1e21 m^-3 at 300 K (null-collision MCC, e-Ar elastic/excitation/"
                              "ionisation, Ar+-Ar elastic/CX), mixed Poisson 1e-10, sort every 20 steps", "fp32", 32,
                    "push_f32_multi_kernel<4,false,8,PBC> (gather + push + wrap + deposit, all species in one launch, tiled storage, out of place)", mcc=mcc, gas=(1, 1, 1, 1e21),
                    cpu_note=", electron collisions by the reference's direct method")
        return
    solvers = {"pcg_mixed": _lib.SOLVER_PCG_MIXED, "spectral": _lib.SOLVER_SPECTRAL, "pcg": _lib.SOLVER_PCG,
               "gs": _lib.SOLVER_GS}
    solver_text = {"pcg_mixed": "FP32 sine-transform refinement, FP64 residual", "spectral": "FP64 sine transform",
                   "pcg": "reference PCG (Jacobi, FP64)", "gs": "red-black SOR (FP64)"}
    if args.precision == "fp64":
        name = args.solver or "pcg"
        st = workloads.bounded_512(ppc=args.ppc, precision=_lib.FP64, solver=solvers[name])
        main_single(args, st, "bounded 512x512, 4 species x %d ppc, FP64 particles (the plain path), %s on the "
                              "Dirichlet-wall operator to 1e-10 relative" % (args.ppc, solver_text[name]), "fp64", 64,
                    "push_f64_kernel<2,true,false> (gather + push + walls + deposit)")
        return
    args.solver = args.solver or "pcg_mixed"
    solver = solvers[args.solver]
    copy_cal = copy_calibration() if args.copy_calibration else None
    st = workloads.bounded_512(ppc=args.ppc, solver=solver, sort_interval=args.sort_interval)
    r = gpu_run(st, args.steps, args.warmup, args.solver)
    nsp = len(st.species)
    part_steps = 0.5 * (r["n0"] + r["n1"]) * args.steps
    value = part_steps / r["t"]
    # dominant kernel: the fused gather + push + deposit, one launch per species
    push_ms, push_calls = r["timing"]["push"]
    avg_launch_s = push_ms / max(push_calls, 1) / 1e3
    # one launch per pushed species, or one for all of them (the FP32 hot
    # kernel): particles per launch from the particle-steps over the calls
    parts_per_launch = 0.5 * (sum(r["per_sp0"]) + sum(r["per_sp1"])) * args.steps / max(push_calls, 1)
    bytes_per_ps = 8 * 4  # read + write x, y, vx, vy in FP32
    achieved = bytes_per_ps * parts_per_launch / avg_launch_s / 1e9
    stage = {k: round(v[0] / max(args.steps, 1), 4) for k, v in r["stages"].items() if v[1]}
    out = {
        "metric": "particle-steps/s per GPU (+% HBM roofline)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["t"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (Philox-seeded uniform positions, Maxwellian velocities)",
        "config": {"workload": "bounded 512x512, 4 species x %d ppc (%d particles), open walls, "
                               "Dirichlet phi=0, Poisson by FP32 sine-transform corrections with an FP64 true "
                               "residual to 1e-10 relative, sort every %d steps" % (
                                   args.ppc, sum(int(s.numparticle) for s in st.species), args.sort_interval),
                   "solver": {"pcg_mixed": "FP32 sine-transform corrections, FP64 true residual and refinement "
                                            "(EPIC_SOLVER_PCG_MIXED + EPIC_PRECOND_SPECTRAL)",
                              "spectral": "FP64 sine-transform solve + FP64 refinement",
                              "pcg": "reference Jacobi PCG, FP64", "gs": "red-black SOR, FP64"}[args.solver],
                   "particles_start": r["n0"], "particles_end": r["n1"],
                   "poisson_iterations_mean": float(np.mean(r["iters"])) if r["iters"] else 0,
                   "stage_ms_per_step": stage},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": PUSH_TRAFFIC_BYTES_PER_PARTICLE * parts_per_launch,
                     "kernel": "push_f32_multi_kernel<4,false,8,OPEN> (CIC gather + leapfrog push + walls + deposit, all species in one launch, tiled storage, out of place)",
                     "bytes_per_particle_step": bytes_per_ps, "avg_launch_ms": avg_launch_s * 1e3,
                     "particles_per_launch": parts_per_launch},
        "device_errors": r["errors"],
    }
    if copy_cal:
        out["copy_calibration"] = copy_cal
    if not args.no_fp64_leg:
        print("[bench] plain FP64 path leg", file=sys.stderr, flush=True)
        out["fp64_plain_path"] = fp64_plain_leg(args)
    if not args.no_serve_leg:
        print("[bench] serving leg", file=sys.stderr, flush=True)
        out["serving"] = serve_leg(args)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_entry(st, args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
