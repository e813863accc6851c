"""The benchmark workloads, expressed as the reference's ini-level inputs
(species lines name,mass,N,T,sign,frac,vsx,vsy,load) and derived through the
normalisation chain of main.cpp:151-261."""
from . import _lib, config

ME = config.ME
M_ION = 1836.0 * ME          # mass ratio 1836 (1.6724825e-27 kg)
M_AR = 6.63352090e-26        # argon


def periodic_64(precision=_lib.FP64, N=409600, solver=_lib.SOLVER_PCG, tol=1e-8, tol_relative=1, seed=1):
    """Plain method, FP64: periodic electron-ion plasma, 64x64 cells,
    1e13 m^-3, Te = 1 eV, Ti = 0.026 eV, 100 particles/cell/species, CIC,
    PCG, dt = 0.05/omega_pe.  200 steps."""
    specs = [config.SpeciesSpec("electron", ME, N, 1.0, -1, 1.0),
             config.SpeciesSpec("ion", M_ION, N, 0.026, 1, 0.0)]
    st = config.derive(64, 64, 0.0, specs, 1e13, 0.05, bc="pbc", solver=solver, tolerance=tol,
                       tol_relative=tol_relative, max_iteration=10000, precision=precision, seed=seed)
    st.info["steps"] = 200
    return st


def bounded_512(ppc=100, precision=_lib.FP32, solver=_lib.SOLVER_PCG_MIXED, sort_interval=20, seed=2,
                tol=1e-10, precond=_lib.PRECOND_SPECTRAL, cells=512):
    """Four-species bounded plasma: 512x512 cells, Dirichlet phi = 0 walls,
    absorbing walls; electrons, positive ions, negative ions (alpha = 0.1) and
    a negative beam (beta = 0.05, drift 10 v_th,e along x); 100 particles/
    cell/species (104,857,600 total); Te = 1 eV, Ti = Tn = 0.026 eV; FP32
    particles, mixed-precision PCG to 1e-10 relative; counting sort every 20
    steps.  500 steps.  `cells` scales the square grid down for parity runs
    (the species, walls and solver stay those of the 512x512 workload)."""
    N = ppc * cells * cells
    specs = [config.SpeciesSpec("electron", ME, N, 1.0, -1, 1.0),
             config.SpeciesSpec("ion", M_ION, N, 0.026, 1, 0.0),
             config.SpeciesSpec("negion", M_ION, N, 0.026, -1, 0.1),
             config.SpeciesSpec("beam", ME, N, 1.0, -1, 0.05, 10.0, 0.0)]
    st = config.derive(cells, cells, 0.0, specs, 1e13, 0.05, bc="open", solver=solver,
                       poisson_op=_lib.OP_DIRICHLET_WALL, tolerance=tol, tol_relative=1, max_iteration=20000,
                       precision=precision, sort_interval=sort_interval, seed=seed, precond=precond)
    st.info["steps"] = 500
    return st


def mcc_256(ppc=200, precision=_lib.FP32, seed=3, precond=_lib.PRECOND_SPECTRAL):
    """MCC workload: 256x256 periodic, electrons + Ar+ at 1e15 m^-3,
    200 particles/cell/species (26,214,400 total), FP32; Ar neutrals 1e21 m^-3 at
    300 K.  2000 steps."""
    N = ppc * 256 * 256
    specs = [config.SpeciesSpec("electron", ME, N, 1.0, -1, 1.0),
             config.SpeciesSpec("ion", M_AR, N, 0.026, 1, 0.0)]
    st = config.derive(256, 256, 0.0, specs, 1e15, 0.05, bc="pbc", solver=_lib.SOLVER_PCG_MIXED,
                       poisson_op=_lib.OP_REF_GHOST, tolerance=1e-10, tol_relative=1, max_iteration=20000,
                       precision=precision, sort_interval=20, seed=seed, capacity_factor=1.05, precond=precond)
    st.info["steps"] = 2000
    return st


ENSEMBLE_GRIDS = (32, 64, 128, 256)


def ensemble_member(i, rng, sort_interval=20, precision=_lib.FP32, solver=_lib.SOLVER_PCG_MIXED, seed0=4000):
    """Member i of the batched ensemble: a periodic electron-ion plasma on a
    grid drawn from 32^2, 64^2, 128^2, 256^2 cells, 20 to 200 particles/cell,
    Te 0.5 to 5 eV, Ti 0.026 eV, 1e13 m^-3.  Odd members are two-stream runs:
    the electrons are two counter-drifting halves at +-vd, vd drawn from 0 to
    5 v_th,e (the species drift is in units of L*W = v_th,e)."""
    g = int(rng.choice(ENSEMBLE_GRIDS))
    ppc = int(rng.integers(20, 201))
    te = float(rng.uniform(0.5, 5.0))
    vd = float(rng.uniform(0.0, 5.0))
    two_stream = i % 2 == 1
    n = ppc * g * g
    if two_stream:
        specs = [config.SpeciesSpec("beam_plus", ME, n // 2, te, -1, 1.0, vd, 0.0),
                 config.SpeciesSpec("ion", M_ION, n, 0.026, 1, 0.0),
                 config.SpeciesSpec("beam_minus", ME, n - n // 2, te, -1, 1.0, -vd, 0.0)]
    else:
        specs = [config.SpeciesSpec("electron", ME, n, te, -1, 1.0),
                 config.SpeciesSpec("ion", M_ION, n, 0.026, 1, 0.0)]
    st = config.derive(g, g, 0.0, specs, 1e13, 0.05, bc="pbc", solver=solver, tolerance=1e-10, tol_relative=1,
                       max_iteration=20000, precision=precision, sort_interval=sort_interval, seed=seed0 + i,
                       precond=_lib.PRECOND_SPECTRAL)
    st.info.update({"grid": g, "ppc": ppc, "Te": te, "two_stream": two_stream, "vd": vd if two_stream else 0.0})
    return st


def ensemble_64(n=64, seed=4, **kw):
    """The batched ensemble: n independent periodic sims with uneven grids
    and particle counts, everything drawn from `seed`."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [ensemble_member(i, rng, **kw) for i in range(n)]
