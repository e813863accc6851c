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
SOLVER_PCG, SOLVER_GS, SOLVER_PCG_MIXED, SOLVER_SPECTRAL, SOLVER_GS_BOLTZMANN, SOLVER_NEWTON_CG = 0, 1, 2, 3, 4, 5
OP_REF_GHOST, OP_DIRICHLET_WALL = 0, 1
PRECOND_JACOBI, PRECOND_SPECTRAL = 0, 1
FP64, FP32 = 0, 1
NERR = 8
(ERRC_GATHER_OOB, ERRC_DEPOSIT_OOB, ERRC_FIXED_OVF, ERRC_CAPACITY, ERRC_NOCONV, ERRC_NAN,
 ERRC_MCC_BOUND) = range(7)
NSTAGE = 8
STAGES = ["deposit", "rho", "solve", "efield", "push", "collide", "sort", "inject"]
TIMING_ONLY = 0x10000
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


COND_CIRCULAR, COND_RECT = 0, 1


class ConductorDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("x_center", ctypes.c_int32), ("y_center", ctypes.c_int32),
                ("electrode_number", ctypes.c_int32), ("grid_radius", ctypes.c_double),
                ("electrode_radius", ctypes.c_double), ("min_x", ctypes.c_int32), ("max_x", ctypes.c_int32),
                ("min_y", ctypes.c_int32), ("max_y", ctypes.c_int32), ("voltage", ctypes.c_double)]


def conductors(items):
    """A ctypes array of ConductorDesc from dicts, e.g.
    {"kind": COND_RECT, "min_x": 4, "max_x": 8, "min_y": 4, "max_y": 6, "voltage": 2.0}."""
    arr = (ConductorDesc * len(items))()
    for k, it in enumerate(items):
        for f, v in it.items():
            setattr(arr[k], f, v)
    return arr


class MccDesc(ctypes.Structure):
    _fields_ = [("electron_sp", ctypes.c_int32), ("ion_sp", ctypes.c_int32), ("gas_density", ctypes.c_double),
                ("gas_temp_K", ctypes.c_double), ("gas_mass", ctypes.c_double), ("e_exc_th", ctypes.c_double),
                ("e_ion_th", ctypes.c_double), ("ntab", ctypes.c_int32), ("e_min", ctypes.c_double),
                ("e_max", ctypes.c_double), ("sigma", ctypes.POINTER(ctypes.c_double)),
                ("enable", ctypes.c_int32 * 5), ("ref_selection", ctypes.c_int32), ("event_cap", ctypes.c_int32)]


class MccStats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int64 * 5), ("eloss_ev", ctypes.c_double), ("births", ctypes.c_int64)]


BATCH_NPUSH = 5
BATCH_NNODE = 4
BATCH_NODE_BLOCKS = (64, 128, 256, 512)


class BatchOpts(ctypes.Structure):
    _fields_ = [("refine", ctypes.c_int), ("use_graph", ctypes.c_int), ("push_variant", ctypes.c_int),
                ("node_block", ctypes.c_int)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("push_variant", ctypes.c_int), ("node_block", ctypes.c_int),
                ("push_ms", ctypes.c_double * BATCH_NPUSH), ("node_ms", ctypes.c_double * BATCH_NNODE),
                ("steps_used", ctypes.c_int)]


class BatchStats(ctypes.Structure):
    _fields_ = [("members", ctypes.c_int), ("groups", ctypes.c_int), ("tiles", ctypes.c_int),
                ("segments", ctypes.c_int), ("chunks", ctypes.c_int), ("mcc_passes", ctypes.c_int),
                ("push_variant", ctypes.c_int), ("node_block", ctypes.c_int),
                ("steps", ctypes.c_int64), ("graph_replays", ctypes.c_int64), ("graph_captures", ctypes.c_int64),
                ("device_bytes", ctypes.c_int64), ("relres_max", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("stage_ms", ctypes.c_double * 4)]


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
        "epic_solve_potential_noalloc": [P, ctypes.POINTER(ctypes.c_int), D],
        "epic_efield": [P],
        "epic_push": [P, ctypes.c_int, ctypes.c_int],
        "epic_rewind": [P, ctypes.c_int],
        "epic_inject": [P, ctypes.c_int],
        "epic_add_emitter": [P, ctypes.POINTER(EmitterDesc)],
        "epic_set_conductors": [P, ctypes.POINTER(ConductorDesc), ctypes.c_int],
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
        "epic_estimate_bytes": [ctypes.POINTER(DomainDesc), ctypes.POINTER(SpeciesDesc), ctypes.c_int,
                                ctypes.POINTER(I64)],
        "epic_device_bytes": [P, ctypes.POINTER(I64)],
        "epic_mcc_estimate_bytes": [ctypes.POINTER(MccDesc), ctypes.POINTER(SpeciesDesc), ctypes.c_int,
                                    ctypes.POINTER(I64)],
        "epic_batch_create": [ctypes.POINTER(BatchOpts), ctypes.POINTER(P)],
        "epic_batch_destroy": [P],
        "epic_batch_add": [P, ctypes.POINTER(DomainDesc), ctypes.POINTER(SpeciesDesc), ctypes.c_int,
                           ctypes.POINTER(ctypes.c_int)],
        "epic_batch_member": [P, ctypes.c_int, ctypes.POINTER(P)],
        "epic_batch_remove": [P, ctypes.c_int],
        "epic_batch_step": [P, ctypes.c_int],
        "epic_batch_init": [P, ctypes.c_int],
        "epic_batch_sync": [P],
        "epic_batch_tune": [P, ctypes.c_int, ctypes.POINTER(BatchTuning)],
        "epic_batch_relres": [P, ctypes.c_int, D],
        "epic_batch_member_steps": [P, ctypes.c_int, ctypes.POINTER(I64)],
        "epic_batch_get_stats": [P, ctypes.POINTER(BatchStats)],
        "epic_batch_timing_enable": [P, ctypes.c_int],
        "epic_batch_configure": [P, ctypes.POINTER(BatchOpts)],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.epic_last_error.argtypes = [P]
    L.epic_last_error.restype = ctypes.c_char_p
    L.epic_batch_last_error.argtypes = [P]
    L.epic_batch_last_error.restype = ctypes.c_char_p
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

    def __init__(self, domain: DomainDesc, species, handle=None):
        self.L = lib()
        self.domain = domain
        self.species = list(species)
        self.nnodes = domain.nx * domain.ny
        if handle is not None:
            # a batch member: the batch owns and destroys it
            self.h = handle
            self.owned = False
            return
        arr = (SpeciesDesc * len(self.species))(*self.species)
        h = ctypes.c_void_p()
        rc = self.L.epic_create(ctypes.byref(domain), arr, len(self.species), ctypes.byref(h))
        if rc != EPIC_OK:
            raise EpicError(f"epic_create failed: {ERR_NAMES.get(rc, rc)}")
        self.h = h
        self.owned = True

    def close(self):
        if self.h and self.owned:
            self.L.epic_destroy(self.h)
        self.h = None

    def device_bytes(self):
        v = ctypes.c_int64()
        self._c(self.L.epic_device_bytes(self.h, ctypes.byref(v)), "device_bytes")
        return v.value

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

    def set_conductors(self, conds):
        """conds: a ctypes array of ConductorDesc (see conductors())."""
        self._c(self.L.epic_set_conductors(self.h, conds, len(conds)), "set_conductors")

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

    def timing(self, on=True, stages=None):
        """Stage timers on (every stage, or only the named ones) or off."""
        v = int(bool(on))
        if on and stages is not None:
            v = TIMING_ONLY | sum(1 << STAGES.index(n) for n in stages)
        self._c(self.L.epic_timing_enable(self.h, v), "timing")

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


def estimate_bytes(domain: DomainDesc, species):
    """Device bytes a context for this description allocates at creation."""
    arr = (SpeciesDesc * len(species))(*species)
    v = ctypes.c_int64()
    rc = lib().epic_estimate_bytes(ctypes.byref(domain), arr, len(species), ctypes.byref(v))
    if rc != EPIC_OK:
        raise EpicError(f"epic_estimate_bytes failed: {ERR_NAMES.get(rc, rc)}")
    return v.value


def mcc_estimate_bytes(d: MccDesc, species):
    """Device bytes epic_mcc_setup adds to a context of these species."""
    arr = (SpeciesDesc * len(species))(*species)
    v = ctypes.c_int64()
    rc = lib().epic_mcc_estimate_bytes(ctypes.byref(d), arr, len(species), ctypes.byref(v))
    if rc != EPIC_OK:
        raise EpicError(f"epic_mcc_estimate_bytes failed: {ERR_NAMES.get(rc, rc)}")
    return v.value


class Batch:
    """A batched ensemble (epic_batch_*): members are Contexts that the batch
    owns; every member advances one step per epic_batch_step."""

    def __init__(self, refine=2, use_graph=True, push_variant=-1, node_block=0):
        self.L = lib()
        o = BatchOpts(refine, 1 if use_graph else 0, push_variant, node_block)
        h = ctypes.c_void_p()
        rc = self.L.epic_batch_create(ctypes.byref(o), ctypes.byref(h))
        if rc != EPIC_OK:
            raise EpicError(f"epic_batch_create failed: {ERR_NAMES.get(rc, rc)}")
        self.h = h
        self.members = {}

    def _c(self, rc, what):
        if rc != EPIC_OK:
            msg = self.L.epic_batch_last_error(self.h).decode() if self.h else ""
            raise EpicError(f"{what} failed: {ERR_NAMES.get(rc, rc)} {msg}")

    def close(self):
        if self.h:
            for m in self.members.values():
                m.h = None
            self.members = {}
            self.L.epic_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, domain: DomainDesc, species):
        """New member; returns (id, Context) for its per-simulation calls."""
        species = list(species)
        arr = (SpeciesDesc * len(species))(*species)
        mid = ctypes.c_int()
        self._c(self.L.epic_batch_add(self.h, ctypes.byref(domain), arr, len(species), ctypes.byref(mid)), "batch_add")
        ch = ctypes.c_void_p()
        self._c(self.L.epic_batch_member(self.h, mid.value, ctypes.byref(ch)), "batch_member")
        ctx = Context(domain, species, handle=ch)
        self.members[mid.value] = ctx
        return mid.value, ctx

    def remove(self, mid):
        self._c(self.L.epic_batch_remove(self.h, mid), "batch_remove")
        ctx = self.members.pop(mid)
        ctx.h = None

    def step(self, n=1):
        self._c(self.L.epic_batch_step(self.h, n), "batch_step")

    def init(self, mid):
        """The member's initial field and rewind happen in the next step."""
        self._c(self.L.epic_batch_init(self.h, mid), "batch_init")

    def sync(self):
        self._c(self.L.epic_batch_sync(self.h), "batch_sync")

    def tune(self, steps_per_candidate=2):
        t = BatchTuning()
        self._c(self.L.epic_batch_tune(self.h, steps_per_candidate, ctypes.byref(t)), "batch_tune")
        return {"push_variant": t.push_variant, "node_block": t.node_block, "push_ms": list(t.push_ms),
                "node_ms": list(t.node_ms), "steps_used": t.steps_used}

    def relres(self, mid):
        v = ctypes.c_double()
        self._c(self.L.epic_batch_relres(self.h, mid, ctypes.byref(v)), "batch_relres")
        return v.value

    def member_steps(self, mid):
        v = ctypes.c_int64()
        self._c(self.L.epic_batch_member_steps(self.h, mid, ctypes.byref(v)), "batch_member_steps")
        return v.value

    def stats(self):
        st = BatchStats()
        self._c(self.L.epic_batch_get_stats(self.h, ctypes.byref(st)), "batch_stats")
        out = {k: getattr(st, k) for k, _ in BatchStats._fields_}
        out["stage_ms"] = list(st.stage_ms)
        return out

    def timing(self, on=True):
        self._c(self.L.epic_batch_timing_enable(self.h, 1 if on else 0), "batch_timing")

    def configure(self, refine=0, use_graph=True, push_variant=-1, node_block=0):
        o = BatchOpts(refine, 1 if use_graph else 0, push_variant, node_block)
        self._c(self.L.epic_batch_configure(self.h, ctypes.byref(o)), "batch_configure")


# ---- request loop (libepic2d_host.so, include/epic2d_serve.h) ---------------
HOST_LIB_PATH = os.path.join(HERE, "libepic2d_host.so")
SERVE_MAX_GRIDS = 8


class ServeConfig(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("rate_hz", ctypes.c_double), ("duration_s", ctypes.c_double),
                ("max_jobs", ctypes.c_int32), ("n_grids", ctypes.c_int32),
                ("grid_sizes", ctypes.c_int32 * SERVE_MAX_GRIDS), ("ppc_min", ctypes.c_int32),
                ("ppc_max", ctypes.c_int32), ("steps_min", ctypes.c_int32), ("steps_max", ctypes.c_int32),
                ("te_min", ctypes.c_double), ("te_max", ctypes.c_double), ("ti_min", ctypes.c_double),
                ("ti_max", ctypes.c_double), ("vd_max", ctypes.c_double), ("two_stream_fraction", ctypes.c_double),
                ("collisional_fraction", ctypes.c_double), ("density", ctypes.c_double),
                ("gas_density", ctypes.c_double), ("dt_coeff", ctypes.c_double), ("sort_interval", ctypes.c_int32),
                ("budget_bytes", ctypes.c_double), ("max_members", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("tune_steps", ctypes.c_int32), ("refine", ctypes.c_int32),
                ("mcc_template", ctypes.POINTER(MccDesc)), ("progress_s", ctypes.c_double),
                ("health_interval", ctypes.c_int32), ("fail_relres", ctypes.c_double), ("fault_job", ctypes.c_int32)]


class ServeJob(ctypes.Structure):
    _fields_ = [("index", ctypes.c_int32), ("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("grid", ctypes.c_int32), ("ppc", ctypes.c_int32),
                ("steps", ctypes.c_int32), ("collisional", ctypes.c_int32), ("two_stream", ctypes.c_int32),
                ("te", ctypes.c_double), ("ti", ctypes.c_double), ("vd", ctypes.c_double),
                ("particles", ctypes.c_int64), ("bytes", ctypes.c_int64), ("ke", ctypes.c_double * 3),
                ("phi_l2", ctypes.c_double), ("collisions", ctypes.c_int64), ("failed", ctypes.c_int32),
                ("relres", ctypes.c_double), ("errors", ctypes.c_int64 * NERR)]


class ServeReport(ctypes.Structure):
    _fields_ = [("jobs", ctypes.c_int32), ("jobs_done", ctypes.c_int32), ("wall_s", ctypes.c_double),
                ("busy_s", ctypes.c_double), ("latency_p50_s", ctypes.c_double), ("latency_p99_s", ctypes.c_double),
                ("latency_mean_s", ctypes.c_double), ("latency_max_s", ctypes.c_double),
                ("wait_p50_s", ctypes.c_double), ("wait_p99_s", ctypes.c_double),
                ("particle_steps", ctypes.c_double), ("throughput", ctypes.c_double),
                ("max_members", ctypes.c_int32), ("peak_bytes", ctypes.c_double), ("batch_steps", ctypes.c_int64),
                ("graph_captures", ctypes.c_int64), ("push_variant", ctypes.c_int32), ("node_block", ctypes.c_int32),
                ("jobs_failed", ctypes.c_int32)]


_host = None


def host_lib():
    global _host
    if _host is not None:
        return _host
    lib()
    if not os.path.exists(HOST_LIB_PATH):
        raise EpicError(f"{HOST_LIB_PATH} is not built: run __graft_entry__.build()")
    H = ctypes.CDLL(HOST_LIB_PATH)
    H.epic_serve_make_jobs.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(ServeJob), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int)]
    H.epic_serve_job_setup.argtypes = [ctypes.POINTER(ServeJob), ctypes.POINTER(ServeConfig), ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(DomainDesc), ctypes.POINTER(SpeciesDesc),
                                       ctypes.POINTER(ctypes.c_int)]
    H.epic_serve_run.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(ServeReport), ctypes.POINTER(ServeJob),
                                 ctypes.c_int]
    for f in (H.epic_serve_make_jobs, H.epic_serve_job_setup, H.epic_serve_run):
        f.restype = ctypes.c_int
    H.epic_serve_last_error.argtypes = []
    H.epic_serve_last_error.restype = ctypes.c_char_p
    P = ctypes.c_void_p
    H.epic_h5_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
    H.epic_h5_group.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    H.epic_h5_dataset_f64.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    H.epic_h5_attr_i32.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int32]
    H.epic_h5_attr_f64.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_double]
    H.epic_h5_attr_str.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32]
    H.epic_h5_close.argtypes = [P]
    for f in (H.epic_h5_create, H.epic_h5_group, H.epic_h5_dataset_f64, H.epic_h5_attr_i32, H.epic_h5_attr_f64,
              H.epic_h5_attr_str, H.epic_h5_close):
        f.restype = ctypes.c_int
    H.epic_h5_lookup3.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32]
    H.epic_h5_lookup3.restype = ctypes.c_uint32
    _host = H
    return H


class H5File:
    """The result.h5 writer (include/epic2d_h5.h).  Objects are numbers; 0 is
    the root group.  close() writes the metadata."""
    ROOT = 0

    def __init__(self, path):
        self.H = host_lib()
        self.h = ctypes.c_void_p()
        self._c(self.H.epic_h5_create(os.fsencode(path), ctypes.byref(self.h)), "create")

    def _c(self, rc, what):
        if rc != EPIC_OK:
            raise EpicError(f"epic_h5_{what} failed: {ERR_NAMES.get(rc, rc)}")

    def group(self, parent, name):
        out = ctypes.c_int()
        self._c(self.H.epic_h5_group(self.h, parent, name.encode(), ctypes.byref(out)), "group")
        return out.value

    def dataset(self, parent, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        dims = (ctypes.c_uint64 * a.ndim)(*a.shape)
        out = ctypes.c_int()
        self._c(self.H.epic_h5_dataset_f64(self.h, parent, name.encode(), a.ndim, dims, a.ctypes.data,
                                           ctypes.byref(out)), "dataset_f64")
        return out.value

    def attr_i32(self, obj, name, v):
        self._c(self.H.epic_h5_attr_i32(self.h, obj, name.encode(), int(v)), "attr_i32")

    def attr_f64(self, obj, name, v):
        self._c(self.H.epic_h5_attr_f64(self.h, obj, name.encode(), float(v)), "attr_f64")

    def attr_str(self, obj, name, s, size):
        self._c(self.H.epic_h5_attr_str(self.h, obj, name.encode(), s.encode(), size), "attr_str")

    def close(self):
        if self.h:
            h, self.h = self.h, ctypes.c_void_p()
            self._c(self.H.epic_h5_close(h), "close")


def h5_lookup3(data: bytes, initval=0):
    return host_lib().epic_h5_lookup3(data, len(data), initval)


def serve_config(**kw):
    """ServeConfig with the request stream of the serving workload
    (grids 64^2..512^2, 50-400 ppc, 100-2000 steps, 20 jobs/s for 30 s,
    200 GB budget); keyword arguments override fields."""
    grids = kw.pop("grids", (64, 128, 256, 512))
    c = ServeConfig(seed=5, rate_hz=20.0, duration_s=30.0, max_jobs=0, n_grids=len(grids), ppc_min=50, ppc_max=400,
                    steps_min=100, steps_max=2000, te_min=0.5, te_max=5.0, ti_min=0.02, ti_max=0.1, vd_max=5.0,
                    two_stream_fraction=0.5, collisional_fraction=0.5, density=1e13, gas_density=1e21,
                    dt_coeff=0.05, sort_interval=20, budget_bytes=200e9, max_members=512, use_graph=1,
                    tune_steps=2, refine=2)
    for k, g in enumerate(grids):
        c.grid_sizes[k] = g
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _serve_check(rc, what):
    if rc != EPIC_OK:
        raise EpicError(f"{what} failed: {ERR_NAMES.get(rc, rc)} {host_lib().epic_serve_last_error().decode()}")


def serve_make_jobs(cfg):
    H = host_lib()
    n = ctypes.c_int()
    _serve_check(H.epic_serve_make_jobs(ctypes.byref(cfg), None, 0, ctypes.byref(n)), "serve_make_jobs")
    arr = (ServeJob * max(n.value, 1))()
    _serve_check(H.epic_serve_make_jobs(ctypes.byref(cfg), arr, n.value, ctypes.byref(n)), "serve_make_jobs")
    return list(arr)[:n.value]


def serve_job_setup(job, cfg, precision=FP32, solver=SOLVER_PCG_MIXED):
    H = host_lib()
    dom = DomainDesc()
    sp = (SpeciesDesc * 3)()
    nsp = ctypes.c_int()
    _serve_check(H.epic_serve_job_setup(ctypes.byref(job), ctypes.byref(cfg), precision, solver, ctypes.byref(dom),
                                        sp, ctypes.byref(nsp)), "serve_job_setup")
    return dom, [SpeciesDesc.from_buffer_copy(sp[k]) for k in range(nsp.value)]


def serve_run(cfg):
    """Runs the stream in real time; returns (report dict, job records)."""
    H = host_lib()
    jobs = serve_make_jobs(cfg)
    out = (ServeJob * max(len(jobs), 1))()
    rep = ServeReport()
    _serve_check(H.epic_serve_run(ctypes.byref(cfg), ctypes.byref(rep), out, len(jobs)), "serve_run")
    return {k: getattr(rep, k) for k, _ in ServeReport._fields_}, list(out)[:len(jobs)]
