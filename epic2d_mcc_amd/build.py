"""Builds libepic2d.so (the HIP kernels and the C-ABI) and the host tools.

hipcc cross-compiles gfx950 code objects on a machine without a GPU.  Every
HIP source is compiled with -ffp-contract=off (the FP64 path reproduces the
reference's operation order) and with packed FP32 instructions disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libepic2d.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EPIC_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{msg}")
    return r


def build(verbose=True, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = []

    def compile_one(f):
        src = os.path.join(CSRC, f)
        obj = os.path.join(BUILD, f.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj in ex.map(compile_one, srcs):
            objs.append(obj)
            if verbose:
                print("compiled", os.path.basename(obj), flush=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("linked", LIB, flush=True)
    # host-side C++ classes and the ePIC++ driver (reference API mirror)
    if os.path.isdir(HOST) and any(f.endswith(".cpp") for f in os.listdir(HOST)):
        build_host(verbose)
    return LIB


def build_host(verbose=True):
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp") and f != "main.cpp"]
    hostlib = os.path.join(HERE, "libepic2d_host.so")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", HOST]
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", *inc, "-o", hostlib + ".tmp", *srcs,
          f"-L{HERE}", "-lepic2d", "-Wl,-rpath,$ORIGIN"])
    os.replace(hostlib + ".tmp", hostlib)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        exe = os.path.join(HERE, "epic2d_run")
        _run(["g++", "-std=c++17", "-O2", *inc, "-o", exe + ".tmp", main, f"-L{HERE}", "-lepic2d_host",
              "-lepic2d", "-Wl,-rpath,$ORIGIN"])
        os.replace(exe + ".tmp", exe)
    if verbose:
        print("built host library and driver", flush=True)


if __name__ == "__main__":
    build()
    sys.exit(0)
