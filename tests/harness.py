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


def dirichlet_solver(dom):
    """Exact solve of gssolver's discrete problem (field.cpp:89-116 at its
    fixed point): (phi[i+1,j] - 2 phi[i,j] + phi[i-1,j])/dx^2 + (phi[i,j+1] -
    2 phi[i,j] + phi[i,j-1])/dy^2 = -rho[i,j] L^2/LDe^2 on the interior nodes,
    phi = 0 on the boundary.  One sparse LU factorisation, reused per call.
    The oracle's own SOR converges to the same solution (pinned against a
    sparse direct solve in test_oracle_conductors); LU reaches it to round-off
    at once, where SOR at omega = 1.4 needs ~2e4 sweeps at 129^2 nodes."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import splu
    mx, my = dom.nx - 2, dom.ny - 2

    def lap1(m, h):
        return sp.diags([np.ones(m - 1), -2.0 * np.ones(m), np.ones(m - 1)], [-1, 0, 1]) / (h * h)

    A = (sp.kron(lap1(mx, dom.dx), sp.identity(my)) + sp.kron(sp.identity(mx), lap1(my, dom.dy))).tocsc()
    lu = splu(A)
    s = (dom.L * dom.L) / (dom.LDe * dom.LDe)

    def solve(rho):
        rho = np.asarray(rho, dtype=np.float64).reshape(dom.nx, dom.ny)
        phi = np.zeros((dom.nx, dom.ny))
        phi[1:-1, 1:-1] = lu.solve(-rho[1:-1, 1:-1].ravel() * s).reshape(mx, my)
        return phi

    return solve


class OracleDirichletLoop:
    """The oracle's loop (main.cpp:366-423: scatter, ComputeRho, Poisson, E,
    push) with the Poisson step on gssolver's Dirichlet-wall problem solved
    exactly (dirichlet_solver), for comparison with the mixed-precision
    sine-transform solve of the same operator."""

    def __init__(self, o, dom):
        self.o = o
        self.solve_phi = dirichlet_solver(dom)

    def _fields(self):
        o = self.o
        for s in range(o.nsp):
            o.scatter(s)
        o.compute_rho()
        o.set_grid(1, self.solve_phi(o.grid(0)))
        o.efield()

    def initial_field(self):
        self._fields()
        for s in range(self.o.nsp):
            self.o.rewind(s)

    def step(self):
        self._fields()
        for s in range(self.o.nsp):
            self.o.push(s)


def total_energy(src, nsp, norm_sp=0, components=3):
    """KE of every species plus PE, in the reference's normalisation
    (Compute_KE / Compute_PE, species.cpp:248-271, domain.cpp:266-286);
    `src` is a Context or an OracleSim.  components=2 sums KE_x + KE_y only:
    without collisions the reference's vz stays 0 (init.cpp:86-92), so KE_z
    is 0, and tests that carry a particle id in vz leave it out."""
    return float(sum(np.sum(src.ke(s, norm_sp)[:components]) for s in range(nsp)) + src.pe(norm_sp))
