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
                                      ("epic_mcc_stats", _lib.MccStats)])
def test_struct_layout_matches_c(tmp_path, name, cls):
    fields = [f[0] for f in cls._fields_]
    src = tmp_path / "probe.c"
    body = "".join(f'printf("%zu ", offsetof({name}, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "epic2d.h"\n'
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
