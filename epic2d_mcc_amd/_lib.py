"""ctypes binding of libepic2d.so (include/epic2d.h).

The library is the product: every call here goes to hand-written gfx950 HIP
kernels.  There is no CPU fallback; if the library is missing the import of
this module raises.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libepic2d.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "epic2d.h")

EPIC_OK = 0
ERR_NAMES = {-1: "EPIC_ERR_ARG", -2: "EPIC_ERR_HIP", -3: "EPIC_ERR_CAPACITY", -4: "EPIC_ERR_STATE",
             -5: "EPIC_ERR_NOCONV", -6: "EPIC_ERR_DEVICE"}

BC = {"pbc": 0, "open": 1, "reflective": 2}
SHAPE = {"CIC": 0, "NGP": 1}
SOLVER_PCG, SOLVER_GS, SOLVER_PCG_MIXED, SOLVER_SPECTRAL = 0, 1, 2, 3
OP_REF_GHOST, OP_DIRICHLET_WALL = 0, 1
PRECOND_JACOBI, PRECOND_SPECTRAL = 0, 1
FP64, FP32 = 0, 1
NERR = 8
NSTAGE = 8
STAGES = ["deposit", "rho", "solve", "efield", "push", "collide", "sort", "inject"]
MCC_E_ELASTIC, MCC_E_EXCITATION, MCC_E_IONIZATION, MCC_I_ELASTIC, MCC_I_CX = range(5)
MCC_NPROC = 5


class DomainDesc(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("L", ctypes.c_double), ("W", ctypes.c_double), ("LDe", ctypes.c_double), ("wpe", ctypes.c_double),
                ("vel_norm", ctypes.c_double), ("density", ctypes.c_double), ("DT", ctypes.c_double),
                ("bc", ctypes.c_int32), ("shape", ctypes.c_int32),
                ("solver", ctypes.c_int32), ("poisson_op", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("tol_relative", ctypes.c_int32), ("max_iteration", ctypes.c_int32),
                ("sor_omega", ctypes.c_double), ("precision", ctypes.c_int32), ("sort_interval", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("precond", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class SpeciesDesc(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("mass", ctypes.c_double), ("charge", ctypes.c_double),
                ("spwt", ctypes.c_double), ("temp", ctypes.c_double), ("numparticle", ctypes.c_int64),
                ("capacity", ctypes.c_int64), ("vsx", ctypes.c_double), ("vsy", ctypes.c_double),
                ("fract_den", ctypes.c_double), ("pushed", ctypes.c_int32)]


class EmitterDesc(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("x1", ctypes.c_double), ("y1", ctypes.c_double),
                ("temp", ctypes.c_double), ("numparticle", ctypes.c_int32), ("vdx", ctypes.c_double),
                ("vdy", ctypes.c_double), ("species_idx1", ctypes.c_int32), ("species_idx2", ctypes.c_int32)]


class MccDesc(ctypes.Structure):
    _fields_ = [("electron_sp", ctypes.c_int32), ("ion_sp", ctypes.c_int32), ("gas_density", ctypes.c_double),
                ("gas_temp_K", ctypes.c_double), ("gas_mass", ctypes.c_double), ("e_exc_th", ctypes.c_double),
                ("e_ion_th", ctypes.c_double), ("ntab", ctypes.c_int32), ("e_min", ctypes.c_double),
                ("e_max", ctypes.c_double), ("sigma", ctypes.POINTER(ctypes.c_double)),
                ("enable", ctypes.c_int32 * 5), ("ref_selection", ctypes.c_int32), ("event_cap", ctypes.c_int32)]


class MccStats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int64 * 5), ("eloss_ev", ctypes.c_double), ("births", ctypes.c_int64)]


class EpicError(RuntimeError):
    pass


def header_symbols():
    """Names of the functions include/epic2d.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(epic_\w+)\s*\(", text, re.M)))


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EpicError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I64 = ctypes.c_int64
    sigs = {
        "epic_create": [ctypes.POINTER(DomainDesc), ctypes.POINTER(SpeciesDesc), ctypes.c_int, ctypes.POINTER(P)],
        "epic_destroy": [P],
        "epic_sync": [P],
        "epic_device_name": [ctypes.c_char_p, ctypes.c_int],
        "epic_upload_particles": [P, ctypes.c_int, D, D, D, D, D, I64],
        "epic_download_particles": [P, ctypes.c_int, D, D, D, D, D, ctypes.POINTER(ctypes.c_uint32), I64,
                                    ctypes.POINTER(I64)],
        "epic_count": [P, ctypes.c_int, ctypes.POINTER(I64)],
        "epic_load_species": [P, ctypes.c_int, I64, ctypes.c_int],
        "epic_scatter": [P, ctypes.c_int],
        "epic_compute_rho": [P],
        "epic_solve_potential": [P, ctypes.POINTER(ctypes.c_int), D],
        "epic_efield": [P],
        "epic_push": [P, ctypes.c_int, ctypes.c_int],
        "epic_rewind": [P, ctypes.c_int],
        "epic_inject": [P, ctypes.c_int],
        "epic_add_emitter": [P, ctypes.POINTER(EmitterDesc)],
        "epic_mcc_setup": [P, ctypes.POINTER(MccDesc)],
        "epic_collide": [P, ctypes.c_int],
        "epic_mcc_get_stats": [P, ctypes.POINTER(MccStats)],
        "epic_sort": [P, ctypes.c_int],
        "epic_compact": [P, ctypes.c_int],
        "epic_step": [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "epic_initial_field": [P],
        "epic_compute_ke": [P, ctypes.c_int, ctypes.c_int, D],
        "epic_compute_momentum": [P, ctypes.c_int, ctypes.c_int, D],
        "epic_compute_pe": [P, ctypes.c_int, D],
        "epic_get_field": [P, ctypes.c_int, D],
        "epic_set_field": [P, ctypes.c_int, D],
        "epic_error_counters": [P, ctypes.POINTER(I64)],
        "epic_clear_error_counters": [P],
        "epic_timing_enable": [P, ctypes.c_int],
        "epic_timing_read": [P, D, ctypes.POINTER(I64)],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.epic_last_error.argtypes = [P]
    L.epic_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def check(ctx, rc, what):
    if rc != EPIC_OK:
        msg = lib().epic_last_error(ctx).decode() if ctx else ""
        raise EpicError(f"{what} failed: {ERR_NAMES.get(rc, rc)} {msg}")


class Context:
    """One device context: a Domain with its Species, driven through the C-ABI."""

    def __init__(self, domain: DomainDesc, species):
        self.L = lib()
        self.domain = domain
        self.species = list(species)
        arr = (SpeciesDesc * len(self.species))(*self.species)
        h = ctypes.c_void_p()
        rc = self.L.epic_create(ctypes.byref(domain), arr, len(self.species), ctypes.byref(h))
        if rc != EPIC_OK:
            raise EpicError(f"epic_create failed: {ERR_NAMES.get(rc, rc)}")
        self.h = h
        self.nnodes = domain.nx * domain.ny

    def close(self):
        if self.h:
            self.L.epic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        check(self.h, rc, what)

    # particles
    def upload(self, s, x, y, vx, vy, vz):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, vx, vy, vz)]
        self._c(self.L.epic_upload_particles(self.h, s, *[dptr(v) for v in a], len(a[0])), "upload")

    def download(self, s, cap=None):
        if cap is None:
            cap = int(self.species[s].capacity or self.species[s].numparticle) + 1
        out = [np.empty(cap) for _ in range(5)]
        ids = np.empty(cap, dtype=np.uint32)
        n = ctypes.c_int64()
        self._c(self.L.epic_download_particles(self.h, s, *[dptr(v) for v in out],
                                               ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap,
                                               ctypes.byref(n)), "download")
        m = n.value
        return tuple(v[:m] for v in out) + (ids[:m],)

    def count(self, s):
        n = ctypes.c_int64()
        self._c(self.L.epic_count(self.h, s, ctypes.byref(n)), "count")
        return n.value

    def load(self, s, n, vz_thermal=False):
        self._c(self.L.epic_load_species(self.h, s, int(n), int(vz_thermal)), "load")

    # operators
    def scatter(self, s):
        self._c(self.L.epic_scatter(self.h, s), "scatter")

    def compute_rho(self):
        self._c(self.L.epic_compute_rho(self.h), "compute_rho")

    def solve(self):
        it = ctypes.c_int()
        err = ctypes.c_double()
        self._c(self.L.epic_solve_potential(self.h, ctypes.byref(it), ctypes.byref(err)), "solve")
        return it.value, err.value

    def efield(self):
        self._c(self.L.epic_efield(self.h), "efield")

    def push(self, s, sub_cycle=1):
        self._c(self.L.epic_push(self.h, s, sub_cycle), "push")

    def rewind(self, s):
        self._c(self.L.epic_rewind(self.h, s), "rewind")

    def step(self, ts):
        it = ctypes.c_int()
        self._c(self.L.epic_step(self.h, ts, ctypes.byref(it)), "step")
        return it.value

    def initial_field(self):
        self._c(self.L.epic_initial_field(self.h), "initial_field")

    def sort(self, s):
        self._c(self.L.epic_sort(self.h, s), "sort")

    def compact(self, s):
        self._c(self.L.epic_compact(self.h, s), "compact")

    def add_emitter(self, e: EmitterDesc):
        self._c(self.L.epic_add_emitter(self.h, ctypes.byref(e)), "add_emitter")

    def inject(self, step):
        self._c(self.L.epic_inject(self.h, step), "inject")

    def mcc_setup(self, d: MccDesc):
        self._c(self.L.epic_mcc_setup(self.h, ctypes.byref(d)), "mcc_setup")

    def collide(self, step):
        self._c(self.L.epic_collide(self.h, step), "collide")

    def mcc_stats(self):
        st = MccStats()
        self._c(self.L.epic_mcc_get_stats(self.h, ctypes.byref(st)), "mcc_stats")
        return {"count": list(st.count), "eloss_ev": st.eloss_ev, "births": st.births}

    # diagnostics
    def ke(self, s, ns):
        out = np.zeros(3)
        self._c(self.L.epic_compute_ke(self.h, s, ns, dptr(out)), "ke")
        return out

    def momentum(self, s, ns):
        out = np.zeros(3)
        self._c(self.L.epic_compute_momentum(self.h, s, ns, dptr(out)), "momentum")
        return out

    def pe(self, ns):
        out = ctypes.c_double()
        self._c(self.L.epic_compute_pe(self.h, ns, ctypes.byref(out)), "pe")
        return out.value

    def get_field(self, which):
        out = np.empty(self.nnodes)
        self._c(self.L.epic_get_field(self.h, int(which), dptr(out)), "get_field")
        return out.reshape(self.domain.nx, self.domain.ny)

    def set_field(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        self._c(self.L.epic_set_field(self.h, int(which), dptr(a)), "set_field")

    def errors(self):
        out = np.zeros(NERR, dtype=np.int64)
        self._c(self.L.epic_error_counters(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))), "errors")
        return out

    def clear_errors(self):
        self._c(self.L.epic_clear_error_counters(self.h), "clear_errors")

    def sync(self):
        self._c(self.L.epic_sync(self.h), "sync")

    def timing(self, on=True):
        self._c(self.L.epic_timing_enable(self.h, int(on)), "timing")

    def timing_read(self):
        ms = np.zeros(NSTAGE)
        calls = np.zeros(NSTAGE, dtype=np.int64)
        self._c(self.L.epic_timing_read(self.h, dptr(ms), calls.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
                "timing_read")
        return {STAGES[k]: (ms[k], int(calls[k])) for k in range(NSTAGE)}


def device_name():
    buf = ctypes.create_string_buffer(256)
    rc = lib().epic_device_name(buf, 256)
    return buf.value.decode() if rc == 0 else "unknown"
