"""Reads HDF5 files through h5dump, the HDF5 library's own command-line
reader (installed with the image's HDF5 1.10 under /opt/conda/bin): an
independent reader for the result.h5 writer.  Test infrastructure only."""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def available():
    return H5DUMP is not None


def run(*args):
    r = subprocess.run([H5DUMP, *map(str, args)], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise AssertionError(f"h5dump {' '.join(map(str, args))} failed ({r.returncode}): {r.stderr[-2000:]}")
    return r.stdout


def contents(path):
    """{object path: 'group' | 'dataset'} from h5dump -n."""
    out = {}
    for line in run("-n", path).splitlines():
        m = re.match(r"\s*(group|dataset)\s+(\S+)\s*$", line)
        if m:
            out[m.group(2)] = m.group(1)
    return out


def _dims(header):
    m = re.search(r"DATASPACE\s+SIMPLE\s*\{\s*\(\s*([0-9,\s]*)\)", header)
    if not m:
        raise AssertionError(f"no simple dataspace in:\n{header}")
    return tuple(int(v) for v in m.group(1).replace(" ", "").split(",") if v)


def dataset(path, name):
    """(values as a float64 array of the dataset's shape, datatype text)."""
    header = run("-H", "-d", name, path)
    dims = _dims(header)
    dtype = re.search(r"DATATYPE\s+(\S+)", header).group(1)
    n = int(np.prod(dims)) if dims else 1
    if n == 0:
        return np.zeros(dims), dtype
    with tempfile.TemporaryDirectory() as d:
        raw = os.path.join(d, "raw.bin")
        run("-d", name, "-b", "LE", "-o", raw, path)
        a = np.fromfile(raw, dtype="<f8")
    assert a.size == n, (name, a.size, dims)
    return a.reshape(dims), dtype


def attribute(path, name):
    """(value, datatype text) of a scalar attribute; floats printed with 17
    significant digits, so they read back exactly."""
    out = run("-m", "%.17g", "-a", name, path)
    dtype = re.search(r"DATATYPE\s+(H5T_\w+)", out).group(1)
    m = re.search(r"\(0\):\s*(.*)\s*$", out.split("DATA {", 1)[1].strip().splitlines()[0])
    text = m.group(1).strip()
    if dtype == "H5T_STRING":
        size = int(re.search(r"STRSIZE\s+(\d+);", out).group(1))
        return text.strip('"'), f"H5T_STRING[{size}]"
    if "F64" in dtype:
        return float(text), dtype
    return int(text), dtype
