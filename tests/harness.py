"""Shared scenarios for the parity tests and smoke(): the same seeded inputs
go to the HIP context and to the oracle."""
import numpy as np

from epic2d_mcc_amd import config
from epic2d_mcc_amd import _lib

ME = config.ME


def small_setup(N=2000, Nx=16, Ny=16, bc="pbc", shape="CIC", solver=_lib.SOLVER_PCG, tol=1e-10,
                precision=_lib.FP64, vsx_beam=0.0, capacity_factor=1.0, maxit=10000, poisson_op=None,
                mass_ratio=1836.0, DT_coeff=0.05, sort_interval=0, ion_mass=None):
    """Electron-ion plasma at 1e13 m^-3, Te = 1 eV, Ti = 0.026 eV (the
    configuration of the 64x64 periodic workload, scaled down)."""
    specs = [config.SpeciesSpec("electron", ME, N, 1.0, -1, 1.0, vsx_beam, 0.0),
             config.SpeciesSpec("ion", ion_mass if ion_mass else mass_ratio * ME, N, 0.026, 1, 0.0, 0.0, 0.0)]
    return config.derive(Nx, Ny, 0.0, specs, 1e13, DT_coeff, bc=bc, shape=shape, solver=solver, tolerance=tol,
                         max_iteration=maxit, precision=precision, capacity_factor=capacity_factor,
                         poisson_op=poisson_op, sort_interval=sort_interval)


def reference_load(st, gauss_seed=0):
    """Particles exactly as Init::Init draws them (global mt19937(0)), from
    the oracle's restatement; returns per-species (x, y, vx, vy, vz)."""
    from oracle_api import OracleSim
    o = OracleSim(st.domain, st.species)
    for s in range(len(st.species)):
        o.init_species(s, gauss_seed)
    return o, [o.particles(s) for s in range(len(st.species))]


def rel(a, b):
    """Normwise relative difference ||a - b||_inf / ||b||_inf."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)
