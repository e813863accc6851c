"""The C-ABI library (CPU-side checks, no GPU calls): it loads, exports every
function include/epic2d.h declares, and the ctypes mirrors of its structs
have the C layout (compared against a program compiled from the header)."""
import ctypes
import os
import subprocess

import pytest

import epic2d_mcc_amd as e
from epic2d_mcc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = e.lib()
    syms = e.header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []


def test_host_layer_built():
    host = os.path.join(ROOT, "epic2d_mcc_amd", "libepic2d_host.so")
    exe = os.path.join(ROOT, "epic2d_mcc_amd", "epic2d_run")
    assert os.path.exists(host) and os.path.exists(exe)


@pytest.mark.parametrize("name,cls", [("epic_domain_desc", _lib.DomainDesc), ("epic_species_desc", _lib.SpeciesDesc),
                                      ("epic_emitter_desc", _lib.EmitterDesc), ("epic_mcc_desc", _lib.MccDesc),
                                      ("epic_conductor_desc", _lib.ConductorDesc),
                                      ("epic_mcc_stats", _lib.MccStats), ("epic_batch_opts", _lib.BatchOpts),
                                      ("epic_batch_tuning", _lib.BatchTuning),
                                      ("epic_batch_stats", _lib.BatchStats),
                                      ("epic_serve_config", _lib.ServeConfig), ("epic_serve_job", _lib.ServeJob),
                                      ("epic_serve_report", _lib.ServeReport)])
def test_struct_layout_matches_c(tmp_path, name, cls):
    fields = [f[0] for f in cls._fields_]
    src = tmp_path / "probe.c"
    body = "".join(f'printf("%zu ", offsetof({name}, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "epic2d_serve.h"\n'
                   f'int main(void){{printf("%zu ", sizeof({name}));{body}return 0;}}\n')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(cls)] + [getattr(cls, f).offset for f in fields]
    assert got == want


def test_context_creation_fails_cleanly_without_gpu_or_bad_args():
    """Invalid descriptors are rejected before any device work."""
    L = e.lib()
    d = _lib.DomainDesc(nx=2, ny=2)
    sp = (_lib.SpeciesDesc * 1)()
    h = ctypes.c_void_p()
    assert L.epic_create(ctypes.byref(d), sp, 1, ctypes.byref(h)) == -1
    assert L.epic_create(None, sp, 1, ctypes.byref(h)) == -1


def test_batch_entry_points_reject_bad_args_without_device_work():
    L = e.lib()
    assert L.epic_batch_create(None, None) == -1
    o = _lib.BatchOpts(refine=99, use_graph=1, push_variant=-1, node_block=0)
    h = ctypes.c_void_p()
    assert L.epic_batch_create(ctypes.byref(o), ctypes.byref(h)) == -1
    o = _lib.BatchOpts(refine=2, use_graph=1, push_variant=-1, node_block=100)
    assert L.epic_batch_create(ctypes.byref(o), ctypes.byref(h)) == -1
    assert L.epic_batch_step(None, 1) == -1
    assert L.epic_batch_add(None, None, None, 0, None) == -1


def test_estimate_bytes_counts_every_creation_allocation():
    """The admission estimate is the sum of epic_create's allocations: a
    hand count for a small two-species FP32 description."""
    L = e.lib()
    d = _lib.DomainDesc(nx=33, ny=17, precision=_lib.FP32)
    sp = (_lib.SpeciesDesc * 2)()
    sp[0].numparticle = 1000
    sp[1].numparticle = 10
    sp[1].capacity = 64
    v = ctypes.c_int64()
    assert L.epic_estimate_bytes(ctypes.byref(d), sp, 2, ctypes.byref(v)) == 0
    nn, ncell = 33 * 17, 32 * 16
    nkey = 2 * 1 * 256  # sort keys: 16x16-cell tiles, 2 x 1 of them
    ncu = 256
    hip_ok = True
    try:
        buf = ctypes.create_string_buffer(256)
        hip_ok = L.epic_device_name(buf, 256) == 0
        if hip_ok:
            ncu = int(buf.value.decode().split(",")[-1].split()[0])
    except Exception:
        hip_ok = False
    want = 0
    for cap in (1000, 64):
        pad = (cap + 255) // 256 * 256  # whole 256-slot tiles
        want += 10 * pad * 4 + 2 * pad * 4 + 2 * nn * 8 + 16
    want += 4 * nn * 8 + 2 * nn * 4 + ncell * 8 * 4 + 2 * nn * 8 + 2 * nn * 8 + 2 * nn * 8
    want += 16 * max(ncu, 1024) * 8 + 64 + 16
    want += 2 * (nkey + 2) * 4 + 2 * 1000 * 4
    blk = 1000 // 2048 + 4
    want += blk * 4 + (max(nkey + 2, blk) // 1024 + 4) * 8
    want += 8 * ncu * 4 * 8 + 64 + 8 * 8
    assert v.value == want


def test_host_library_exports_serving_entry_points():
    import re
    text = open(os.path.join(ROOT, "include", "epic2d_serve.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(epic_\w+)\s*\(", text, re.M)))
    assert len(names) == 4
    H = _lib.host_lib()
    assert [n for n in names if not hasattr(H, n)] == []


def test_host_library_exports_h5_writer_entry_points():
    import re
    text = open(os.path.join(ROOT, "include", "epic2d_h5.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|uint32_t)\s+(epic_\w+)\s*\(", text, re.M)))
    assert len(names) == 8
    H = _lib.host_lib()
    assert [n for n in names if not hasattr(H, n)] == []
