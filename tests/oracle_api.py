"""ctypes wrapper of oracle/liboracle.so (the FP64 CPU restatement) and of
oracle/_ref/libref_pcg.so (the reference's own pcg()).  Test infrastructure
only: used as the checker, never as the thing measured."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PCG = os.path.join(ROOT, "oracle", "_ref", "libref_pcg.so")

D = ctypes.POINTER(ctypes.c_double)


def dp(a):
    return a.ctypes.data_as(D)


class OrcDomain(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("L", ctypes.c_double), ("W", ctypes.c_double),
                ("LDe", ctypes.c_double), ("wpe", ctypes.c_double), ("vel_norm", ctypes.c_double),
                ("density", ctypes.c_double), ("DT", ctypes.c_double), ("bc", ctypes.c_int), ("shape", ctypes.c_int),
                ("solver", ctypes.c_int), ("tolerance", ctypes.c_double), ("max_iteration", ctypes.c_int),
                ("elastic", ctypes.c_int), ("excitation", ctypes.c_int), ("ionization", ctypes.c_int),
                ("gas_density", ctypes.c_double)]


class OrcSpecies(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_double), ("charge", ctypes.c_double), ("spwt", ctypes.c_double),
                ("temp", ctypes.c_double), ("numparticle", ctypes.c_int), ("vsx", ctypes.c_double),
                ("vsy", ctypes.c_double), ("load_type", ctypes.c_int), ("load_n", ctypes.c_int)]


class OrcInputs(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("x0", ctypes.c_double), ("y0", ctypes.c_double),
                ("NUM_TS", ctypes.c_int), ("DT_coeff", ctypes.c_double), ("density", ctypes.c_double),
                ("norm_scheme", ctypes.c_int), ("L_scale", ctypes.c_double), ("T_scale_kind", ctypes.c_int),
                ("T_scale_value", ctypes.c_double), ("nsp", ctypes.c_int), ("mass", D),
                ("nparticles", ctypes.POINTER(ctypes.c_int)), ("temps", D),
                ("charge_signs", ctypes.POINTER(ctypes.c_int)), ("frac", D), ("vsx", D), ("vsy", D)]


class OrcSetup(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("LDe", "wpe", "LDi", "wpi", "CS", "L", "W", "vel_norm", "dx", "dy",
                                               "DT", "x0", "y0", "Lx", "Ly")] + \
               [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("normden", ctypes.c_double * 16),
                ("spwt", ctypes.c_double * 16), ("charge", ctypes.c_double * 16)]


_o = None


def oracle():
    global _o
    if _o is None:
        _o = ctypes.CDLL(ORACLE)
        o = _o
        o.orc_create.restype = ctypes.c_void_p
        o.orc_create.argtypes = [ctypes.POINTER(OrcDomain), ctypes.POINTER(OrcSpecies), ctypes.c_int]
        for name in ("orc_destroy", "orc_compute_rho", "orc_efield"):
            getattr(o, name).argtypes = [ctypes.c_void_p]
        o.orc_reseed.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        o.orc_init_species.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint]
        o.orc_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_int, D, D, D, D, D, ctypes.c_long]
        o.orc_get_particles.argtypes = [ctypes.c_void_p, ctypes.c_int, D, D, D, D, D]
        o.orc_count.argtypes = [ctypes.c_void_p, ctypes.c_int]
        o.orc_count.restype = ctypes.c_long
        o.orc_get_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, D]
        o.orc_set_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, D]
        o.orc_scatter.argtypes = [ctypes.c_void_p, ctypes.c_int]
        o.orc_solve.argtypes = [ctypes.c_void_p, D]
        o.orc_push.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        o.orc_rewind.argtypes = [ctypes.c_void_p, ctypes.c_int]
        o.orc_ke.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, D]
        o.orc_momentum.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, D]
        o.orc_pe.argtypes = [ctypes.c_void_p, ctypes.c_int]
        o.orc_pe.restype = ctypes.c_double
        o.orc_pcg.argtypes = [ctypes.c_int, ctypes.c_int, D, D, ctypes.c_int, ctypes.c_double]
        o.orc_rnd_draws.argtypes = [ctypes.c_uint, ctypes.c_int, D]
        o.orc_mt19937_nth.argtypes = [ctypes.c_uint, ctypes.c_int]
        o.orc_mt19937_nth.restype = ctypes.c_uint
        o.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        o.orc_philox2x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_uint32)]
        o.orc_compute_setup.argtypes = [ctypes.POINTER(OrcInputs), ctypes.POINTER(OrcSetup)]
        for n in ("orc_sigma_elastic", "orc_sigma_excitation", "orc_sigma_ionization"):
            getattr(o, n).argtypes = [ctypes.c_double]
            getattr(o, n).restype = ctypes.c_double
        o.orc_mcc_build_tables.argtypes = [ctypes.c_void_p]
        o.orc_collide_direct.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_long), D]
        o.orc_set_conductors.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        o.orc_is_conductor.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        o.orc_init_species.restype = ctypes.c_int
        o.orc_solve.restype = ctypes.c_int
        o.orc_last_iters.argtypes = [ctypes.c_void_p]
        o.orc_sample_vel.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, D]
        o.orc_collide_one.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, D]
        o.orc_collide_one.restype = ctypes.c_int
        o.orc_inject.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
        o.orc_set_tolerance.argtypes = [ctypes.c_void_p, ctypes.c_double]
        o.orc_set_threads.argtypes = [ctypes.c_int]
    return _o


def ref_pcg_lib():
    if not os.path.exists(REF_PCG):
        return None
    r = ctypes.CDLL(REF_PCG)
    r.ref_pcg.argtypes = [ctypes.c_int, ctypes.c_int, D, D, ctypes.c_int, ctypes.c_double, D]
    r.ref_nrpcg.argtypes = [ctypes.c_int, ctypes.c_int, D, ctypes.c_double, ctypes.c_double, D]
    return r


class OracleSim:
    """Oracle simulation built from the same descriptors as the HIP context."""

    def __init__(self, dom, species, gas=None):
        o = oracle()
        self.o = o
        d = OrcDomain(x0=dom.x0, y0=dom.y0, dx=dom.dx, dy=dom.dy, nx=dom.nx, ny=dom.ny, L=dom.L, W=dom.W,
                      LDe=dom.LDe, wpe=dom.wpe, vel_norm=dom.vel_norm, density=dom.density, DT=dom.DT, bc=dom.bc,
                      shape=dom.shape, solver=dom.solver if dom.solver in (1, 4, 5) else 0, tolerance=dom.tolerance,
                      max_iteration=dom.max_iteration, elastic=0, excitation=0, ionization=0, gas_density=0.0)
        if gas:
            d.elastic, d.excitation, d.ionization, d.gas_density = gas
        sp = (OrcSpecies * len(species))(*[OrcSpecies(mass=s.mass, charge=s.charge, spwt=s.spwt, temp=s.temp,
                                                      numparticle=int(s.numparticle), vsx=s.vsx, vsy=s.vsy,
                                                      load_type=0, load_n=0) for s in species])
        self.h = o.orc_create(ctypes.byref(d), sp, len(species))
        self.nx, self.ny = dom.nx, dom.ny
        self.nsp = len(species)

    def __del__(self):
        try:
            self.o.orc_destroy(self.h)
        except Exception:
            pass

    def set_tolerance(self, tol):
        self.o.orc_set_tolerance(self.h, tol)

    def init_species(self, s, gauss_seed=0):
        return self.o.orc_init_species(self.h, s, gauss_seed)

    def sample_vel(self, mass, temp, n):
        out = np.empty(n)
        self.o.orc_sample_vel(self.h, mass, temp, n, dp(out))
        return out

    def build_mcc_tables(self):
        self.o.orc_mcc_build_tables(self.h)

    def collide_direct(self, se=0, st=1):
        """handle_collisions once; returns (counts[3], energy removed in eV)."""
        counts = (ctypes.c_long * 3)()
        eloss = ctypes.c_double()
        self.o.orc_collide_direct(self.h, se, st, counts, ctypes.byref(eloss))
        return list(counts), eloss.value

    def collide_one(self, v3, eindex, x=0.0, y=0.0, se=0, st=1):
        v = np.ascontiguousarray(v3, dtype=np.float64).copy()
        proc = self.o.orc_collide_one(self.h, se, st, eindex, x, y, dp(v))
        return proc, v

    def inject(self, e, pos_seed):
        """e: the library's EmitterDesc (the oracle's orc_emitter has the same
        layout)."""
        self.o.orc_inject(self.h, ctypes.byref(e), pos_seed)

    def set_conductors(self, conds):
        """conds: a ctypes array of the library's epic_conductor_desc (the
        oracle's orc_conductor has the same layout)."""
        self.o.orc_set_conductors(self.h, ctypes.cast(conds, ctypes.c_void_p) if len(conds) else None, len(conds))

    def is_conductor(self, x, y):
        return bool(self.o.orc_is_conductor(self.h, x, y))

    def set_particles(self, s, x, y, vx, vy, vz):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, vx, vy, vz)]
        self.o.orc_set_particles(self.h, s, *[dp(v) for v in a], len(a[0]))

    def particles(self, s):
        n = self.o.orc_count(self.h, s)
        out = [np.empty(n) for _ in range(5)]
        self.o.orc_get_particles(self.h, s, *[dp(v) for v in out])
        return tuple(out)

    def grid(self, which):
        out = np.empty(self.nx * self.ny)
        self.o.orc_get_grid(self.h, which, dp(out))
        return out.reshape(self.nx, self.ny)

    def set_grid(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        self.o.orc_set_grid(self.h, which, dp(a))

    def scatter(self, s):
        return self.o.orc_scatter(self.h, s)

    def compute_rho(self):
        self.o.orc_compute_rho(self.h)

    def solve(self):
        err = ctypes.c_double()
        it = self.o.orc_solve(self.h, ctypes.byref(err))
        return it, err.value

    def efield(self):
        self.o.orc_efield(self.h)

    def push(self, s, sub=1):
        return self.o.orc_push(self.h, s, sub)

    def rewind(self, s):
        return self.o.orc_rewind(self.h, s)

    def ke(self, s, ns):
        out = np.zeros(3)
        self.o.orc_ke(self.h, s, ns, dp(out))
        return out

    def momentum(self, s, ns):
        out = np.zeros(3)
        self.o.orc_momentum(self.h, s, ns, dp(out))
        return out

    def pe(self, ns):
        return self.o.orc_pe(self.h, ns)

    # main.cpp:366-382 and one pass of :395-423 (no emitters, no collisions)
    def initial_field(self):
        for s in range(self.nsp):
            self.scatter(s)
        self.compute_rho()
        self.solve()
        self.efield()
        for s in range(self.nsp):
            self.rewind(s)

    def step(self, pushed=None):
        for s in range(self.nsp):
            self.scatter(s)
        self.compute_rho()
        it, _ = self.solve()
        self.efield()
        for s in range(self.nsp):
            if pushed is None or pushed[s]:
                self.push(s)
        return it
