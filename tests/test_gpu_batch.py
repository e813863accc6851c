"""Batched ensemble (epic_batch_*): a member advanced by the batch equals the
same simulation on its own context; it tracks the FP64 oracle; a member's
result is the same alone and in a mix of uneven grids; HIP-graph replays
equal eager steps bit for bit; tuning the kernel shapes changes no result."""
import numpy as np
import pytest

from epic2d_mcc_amd import _lib, config, xsec
from harness import reference_load, rel, small_setup

pytestmark = pytest.mark.gpu


def _fp32(N, Nx, Ny, bc="pbc", sort_interval=0, vsx_beam=0.0, op=None, capacity_factor=1.0, ion_mass=None):
    st = small_setup(N=N, Nx=Nx, Ny=Ny, bc=bc, precision=_lib.FP32, solver=_lib.SOLVER_PCG_MIXED, tol=1e-10,
                     sort_interval=sort_interval, vsx_beam=vsx_beam, poisson_op=op,
                     capacity_factor=capacity_factor, ion_mass=ion_mass)
    st.domain.tol_relative = 1
    st.domain.precond = _lib.PRECOND_SPECTRAL
    return st


def _mcc(gas=2e20):
    return xsec.mcc_desc(xsec.synthetic_ar(), gas_density=gas, gas_temp_K=300.0, enable=(1, 1, 1, 1, 1))


def _load(ctx, st, seed_shift=0):
    for s, sp in enumerate(st.species):
        ctx.load(s, sp.numparticle)


def _state(ctx, nsp):
    out = [ctx.get_field(1)]
    for s in range(nsp):
        out.extend(ctx.download(s)[:5])
    return out


def test_member_equals_single_context():
    """128x128 nodes (the single-context FP32 hot kernel): the same push,
    deposit and Poisson arithmetic (both take the transform products from
    the same tile GEMMs with the same padding), so phi and the particles are
    the same bit for bit."""
    print("[batch] member vs single context", flush=True)
    st = _fp32(200000, 128, 128, sort_interval=2)
    o, parts = reference_load(st)
    single = _lib.Context(st.domain, st.species)
    B = _lib.Batch()
    mid, mem = B.add(st.domain, st.species)
    for s, p in enumerate(parts):
        single.upload(s, *p)
        mem.upload(s, *p)
    single.initial_field()
    mem.initial_field()
    for ts in range(5):
        single.step(ts)
    B.step(5)
    B.sync()
    ep = rel(mem.get_field(1), single.get_field(1))
    print(f"  phi {ep:.2e}  relres {B.relres(mid):.2e}  bitwise {np.array_equal(mem.get_field(1), single.get_field(1))}")
    assert np.array_equal(mem.get_field(1), single.get_field(1))
    assert B.relres(mid) < 1e-10
    for s in range(2):
        a, b = mem.download(s), single.download(s)
        for k in range(6):
            assert np.array_equal(a[k], b[k]), (s, k)
    assert B.member_steps(mid) == 5
    assert all(v == 0 for v in mem.errors())
    B.close()


def test_member_tracks_fp64_oracle():
    print("[batch] member vs oracle", flush=True)
    st = _fp32(300000, 128, 128)
    st.domain.tolerance = 1e-12
    o, parts = reference_load(st)
    B = _lib.Batch(refine=3)
    mid, mem = B.add(st.domain, st.species)
    for s, p in enumerate(parts):
        mem.upload(s, *p)
    mem.initial_field()
    o.initial_field()
    for ts in range(3):
        B.step(1)
        o.step()
    er = rel(mem.get_field(0), o.grid(0))
    ep = rel(mem.get_field(1), o.grid(1))
    print(f"  rho {er:.2e} phi {ep:.2e}")
    assert er < 2e-3 and ep < 2e-3
    for s in range(2):
        g = mem.download(s)
        c = o.particles(s)
        errs = [rel(g[k], c[k]) for k in range(4)]
        print(f"  species {s}: " + " ".join(f"{e:.2e}" for e in errs))
        assert errs[0] < 1e-5 and errs[1] < 1e-5
        assert errs[2] < 2e-3 and errs[3] < 2e-3
    B.close()


def _mix():
    """Uneven members: periodic 33x33 and 65x65, an open 48x40 box with a
    drifting beam, a Dirichlet-wall 64x64 grid, a collisional 32x32 argon
    plasma."""
    out = [("pbc33", _fp32(20000, 33, 33, sort_interval=5), None),
           ("pbc65", _fp32(60000, 65, 65, sort_interval=4), None),
           ("open48x40", _fp32(30000, 48, 40, bc="open", vsx_beam=2.0, sort_interval=3), None),
           ("wall64", _fp32(40000, 64, 64, op=_lib.OP_DIRICHLET_WALL), None)]
    mcc = _fp32(30000, 32, 32, sort_interval=5, capacity_factor=1.5, ion_mass=xsec.M_AR)
    out.append(("mcc32", mcc, _mcc()))
    return out


def _run_batch(entries, steps, use_graph=True):
    B = _lib.Batch(use_graph=use_graph)
    ids = []
    for name, st, mcc in entries:
        mid, ctx = B.add(st.domain, st.species)
        _load(ctx, st)
        if mcc is not None:
            ctx.mcc_setup(mcc)
        ctx.initial_field()
        ids.append(mid)
    B.step(steps)
    B.sync()
    return B, ids


def test_same_alone_and_in_the_mix():
    """Each member's phi, kinetic energies, particles and collision counts
    after 12 steps are the same, bit for bit, alone and inside the mix: every
    stage is the same arithmetic per member, and the transform products of a
    grid size are tile GEMMs whose per-member result does not depend on how
    many members share the launch."""
    print("[batch] alone vs mixed", flush=True)
    entries = _mix()
    steps = 12
    Bm, ids = _run_batch(entries, steps)
    st_mix = Bm.stats()
    print(f"  mix: {st_mix['members']} members, {st_mix['groups']} groups, {st_mix['tiles']} tiles, "
          f"{st_mix['chunks']} chunks, relres_max {st_mix['relres_max']:.2e}")
    assert st_mix["relres_max"] < 1e-10
    for (name, st, mcc), mid in zip(entries, ids):
        Ba, (aid,) = _run_batch([(name, st, mcc)], steps)
        a = Ba.members[aid]
        m = Bm.members[mid]
        ep = rel(m.get_field(1), a.get_field(1))
        kes = []
        for s in range(2):
            ka, km = a.ke(s, 0), m.ke(s, 0)
            kes.append(rel(km, ka))
        print(f"  {name}: phi {ep:.2e} ke {max(kes):.2e}")
        assert np.array_equal(m.get_field(1), a.get_field(1))
        for s in range(2):
            assert np.array_equal(a.ke(s, 0), m.ke(s, 0))
            pa, pm = a.download(s), m.download(s)
            for k in range(6):
                assert np.array_equal(pa[k], pm[k]), (name, s, k)
        if mcc is not None:
            ca, cm = a.mcc_stats()["count"], m.mcc_stats()["count"]
            print(f"  {name}: collisions alone {list(ca)} mixed {list(cm)}")
            assert sum(ca) > 0
            assert list(ca) == list(cm)
        assert all(v == 0 for v in m.errors())
        Ba.close()
    Bm.close()


def test_graph_replays_equal_eager_steps():
    """Every replay of the captured step against an eager twin, with
    downloads (eager work and synchronisation) between replays and a sort
    (descriptor refresh) every third step."""
    print("[batch] graph vs eager", flush=True)
    entries = [("a", _fp32(40000, 64, 64, sort_interval=3), None),
               ("b", _fp32(30000, 33, 33, bc="open", vsx_beam=2.0, sort_interval=3), None),
               ("c", _fp32(30000, 32, 32, sort_interval=3, capacity_factor=1.5, ion_mass=xsec.M_AR), _mcc())]
    G, gids = _run_batch(entries, 0, use_graph=True)
    E, eids = _run_batch(entries, 0, use_graph=False)
    for k in range(8):
        G.step(1)
        E.step(1)
        for (name, st, _), gi, ei in zip(entries, gids, eids):
            sg = _state(G.members[gi], 2)
            se = _state(E.members[ei], 2)
            for u, v in zip(sg, se):
                assert np.array_equal(u, v), f"step {k} member {name}"
    sg = G.stats()
    print(f"  graph captures {sg['graph_captures']} replays {sg['graph_replays']}")
    assert sg["graph_replays"] >= 6
    assert E.stats()["graph_replays"] == 0
    G.close()
    E.close()


def test_tuning_changes_no_result():
    print("[batch] tuning", flush=True)
    entries = [("a", _fp32(60000, 96, 96, sort_interval=4), None),
               ("b", _fp32(20000, 33, 33, bc="reflective"), None)]
    A, aids = _run_batch(entries, 0)
    T, tids = _run_batch(entries, 0)
    t = T.tune(steps_per_candidate=1)
    print(f"  tuned: {t}")
    n = t["steps_used"]
    assert n == 2 * (_lib.BATCH_NPUSH + _lib.BATCH_NNODE)
    A.step(n)
    for ai, ti in zip(aids, tids):
        assert T.member_steps(ti) == n
        for u, v in zip(_state(A.members[ai], 2), _state(T.members[ti], 2)):
            assert np.array_equal(u, v)
    A.step(3)
    T.step(3)
    for ai, ti in zip(aids, tids):
        for u, v in zip(_state(A.members[ai], 2), _state(T.members[ti], 2)):
            assert np.array_equal(u, v)
    s = T.stats()
    assert s["push_variant"] == t["push_variant"] and s["node_block"] == t["node_block"]
    A.close()
    T.close()


def test_membership_changes_between_steps():
    """Continuous batching: a member joins after others have run and one
    leaves; the joiner's result equals the same simulation run alone."""
    print("[batch] join and leave", flush=True)
    e1 = _fp32(30000, 48, 48, sort_interval=4)
    e2 = _fp32(20000, 33, 33, sort_interval=5)
    e3 = _fp32(25000, 40, 40, bc="open", vsx_beam=1.0)
    B = _lib.Batch()
    i1, c1 = B.add(e1.domain, e1.species)
    _load(c1, e1)
    c1.initial_field()
    i2, c2 = B.add(e2.domain, e2.species)
    _load(c2, e2)
    c2.initial_field()
    B.step(4)
    i3, c3 = B.add(e3.domain, e3.species)
    _load(c3, e3)
    c3.initial_field()
    B.step(3)
    B.remove(i2)
    B.step(5)
    assert B.member_steps(i1) == 12 and B.member_steps(i3) == 8
    A, (a3,) = _run_batch([("e3", e3, None)], 8)
    ep = rel(B.members[i3].get_field(1), A.members[a3].get_field(1))
    print(f"  joiner phi vs alone {ep:.2e}")
    assert ep < 1e-6
    A.close()
    B.close()


def test_device_bytes_match_estimate():
    st = _fp32(10000, 40, 24)
    B = _lib.Batch()
    mid, ctx = B.add(st.domain, st.species)
    assert ctx.device_bytes() == _lib.estimate_bytes(st.domain, st.species)
    single = _lib.Context(st.domain, st.species)
    assert single.device_bytes() == _lib.estimate_bytes(st.domain, st.species)
    B.close()
