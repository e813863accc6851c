"""The result.h5 writer (include/epic2d_h5.h, host/h5lite.cpp) on the CPU.

Pins:
  - lookup3, the format's metadata checksum, against the self-test values of
    Bob Jenkins' published lookup3.c;
  - the superblock and object-header fields the HDF5 File Format
    Specification fixes (signature, versions, sizes of offsets and lengths,
    base address, undefined extension address, end-of-file address, the
    checksums) on a tiny file;
  - the whole reference layout (output.cpp:4-315), read back by the HDF5
    library's own h5dump: object tree, dataset shapes and types, every value
    bit for bit, every attribute exactly.
The last pin needs h5dump (the image's HDF5 1.10 install); without it that
test is skipped."""
import struct

import numpy as np
import pytest

import h5dump
from epic2d_mcc_amd import _lib

UNDEF = 0xFFFFFFFFFFFFFFFF


def test_lookup3_known_answers():
    # lookup3.c (2006), driver5: hashlittle over "" and the 30-byte string
    assert _lib.h5_lookup3(b"", 0) == 0xDEADBEEF
    assert _lib.h5_lookup3(b"", 0xDEADBEEF) == 0xBD5B7DDE
    s = b"Four score and seven years ago"
    assert _lib.h5_lookup3(s, 0) == 0x17770551
    assert _lib.h5_lookup3(s, 1) == 0xCD628161


def _tiny(path):
    f = _lib.H5File(path)
    g = f.group(f.ROOT, "metadata")
    f.attr_i32(g, "NC", 17)
    d = f.group(f.ROOT, "fielddata")
    p = f.group(d, "pot")
    f.dataset(p, "0", np.arange(6.0).reshape(2, 3))
    f.close()


def test_superblock_and_object_header_fields(tmp_path):
    path = tmp_path / "t.h5"
    _tiny(path)
    b = path.read_bytes()
    # superblock version 2 (HDF5 File Format Specification 3.0, II.A)
    assert b[:8] == b"\x89HDF\r\n\x1a\n"
    version, so, sl, flags = b[8], b[9], b[10], b[11]
    assert (version, so, sl, flags) == (2, 8, 8, 0)
    base, ext, eof, root = struct.unpack_from("<4Q", b, 12)
    assert base == 0 and ext == UNDEF
    assert eof == len(b)
    assert struct.unpack_from("<I", b, 44)[0] == _lib.h5_lookup3(b[:44])
    # root group object header, version 2 (IV.A.1.b): signature, version,
    # flags (4-byte chunk size, attribute phase change stored), checksum
    assert 48 <= root < eof
    assert b[root:root + 4] == b"OHDR" and b[root + 4] == 2
    oflags = b[root + 5]
    assert oflags & 0x03 == 2
    off = root + 6 + (4 if oflags & 0x10 else 0)
    size = struct.unpack_from("<I", b, off)[0]
    end = off + 4 + size
    assert struct.unpack_from("<I", b, end)[0] == _lib.h5_lookup3(b[root:end])
    # its messages: Link Info (0x02), Group Info (0x0A), then one Link (0x06) per child
    types = []
    k = off + 4
    while k < end:
        t, n = b[k], struct.unpack_from("<H", b, k + 1)[0]
        types.append(t)
        if t == 0x06:
            ln = b[k + 4:k + 4 + n]
            assert ln[0] == 1 and ln[1] == 0  # version 1, hard link, 1-byte name length
            name = ln[3:3 + ln[2]].decode()
            addr = struct.unpack_from("<Q", ln, 3 + ln[2])[0]
            assert name in ("metadata", "fielddata") and b[addr:addr + 4] == b"OHDR"
        k += 4 + n
    assert k == end and types == [0x02, 0x0A, 0x06, 0x06]


def _layout(path, nx, ny, nsp, rng):
    """The reference's result.h5 for a synthetic run, through the writer the
    way host Output writes it; returns what was written."""
    f = _lib.H5File(path)
    R = f.ROOT
    fd, tv, md, ms = (f.group(R, n) for n in ("fielddata", "time_var", "metadata", "metadata_species"))
    meta = {"NC": (nx, "i"), "NUM_TS": (10, "i"), "write_int": (5, "i"), "write_int_phase": (10, "i"),
            "DT_coeff": (0.05, "f"), "density": (1e13, "f"), "save_fig": (0, "i"), "norm_scheme": (1, "i"),
            "sub_cycle_interval": (1, "i"), "LDe": (rng.random(), "f"), "LDi": (rng.random(), "f"),
            "wpe": (rng.random() * 1e9, "f"), "wpi": (rng.random() * 1e7, "f"), "spno": (nsp, "i"),
            "GAS_DENSITY": (0.0, "f")}
    for k, (v, t) in meta.items():
        (f.attr_i32 if t == "i" else f.attr_f64)(md, k, v)
    names = ["electron", "ion", "negion"][:nsp]
    spmeta = {}
    for s, name in enumerate(names):
        g = f.group(ms, name)
        f.attr_str(g, "name", name, 1)
        vals = {"mass": rng.random() * 1e-27, "charge": (-1) ** (s + 1) * 1.602176565e-19, "spwt": rng.random() * 1e6,
                "temperature": rng.random(), "density": rng.random()}
        for k, v in vals.items():
            f.attr_f64(g, k, v)
        f.attr_i32(g, "num_particles", 1000 + s)
        f.attr_f64(g, "streaming_velocity", -3.5 * s)
        spmeta[name] = dict(vals, num_particles=1000 + s, streaming_velocity=-3.5 * s)
    data = {}
    for ts in (0, 5, 10):
        for fld in ("pot", "efieldx", "efieldy"):
            a = rng.standard_normal((nx, ny))
            f.dataset(f.group(fd, fld), str(ts), a)
            data[f"/fielddata/{fld}/{ts}"] = a
        for name in names[:2]:
            a = rng.standard_normal((nx, ny))
            f.dataset(f.group(fd, "den_" + name), str(ts), a)
            data[f"/fielddata/den_{name}/{ts}"] = a
    for ts in (0, 10):
        for s, name in enumerate(names):
            g = f.group(R, "particle_" + name)
            n = 0 if (ts == 10 and s == 1) else 50 + 7 * s  # all absorbed: an N x 2 set with N = 0
            for kind in ("pos", "vel"):
                a = rng.standard_normal((n, 2))
                f.dataset(g, f"{kind}{ts}", a)
                data[f"/particle_{name}/{kind}{ts}"] = a
    ke = rng.standard_normal((3, 3 * nsp + 2))
    f.dataset(tv, "kinetic_energy", ke)
    data["/time_var/kinetic_energy"] = ke
    f.close()
    return meta, spmeta, data


@pytest.mark.skipif(not h5dump.available(), reason="no h5dump in this image")
def test_reference_layout_reads_back_through_the_hdf5_library(tmp_path):
    path = tmp_path / "result.h5"
    rng = np.random.default_rng(3)
    nx, ny, nsp = 9, 7, 3  # non-square, so the (nx, ny) shape of a field dataset is checked
    meta, spmeta, data = _layout(path, nx, ny, nsp, rng)
    tree = h5dump.contents(path)
    groups = {"/", "/fielddata", "/time_var", "/metadata", "/metadata_species", "/fielddata/pot",
              "/fielddata/efieldx", "/fielddata/efieldy", "/fielddata/den_electron", "/fielddata/den_ion",
              "/metadata_species/electron", "/metadata_species/ion", "/metadata_species/negion",
              "/particle_electron", "/particle_ion", "/particle_negion"}
    assert {k for k, v in tree.items() if v == "group"} == groups
    assert {k for k, v in tree.items() if v == "dataset"} == set(data)
    for name, a in data.items():
        got, dtype = h5dump.dataset(path, name)
        assert dtype == "H5T_IEEE_F64LE"
        assert got.shape == a.shape, name
        assert np.array_equal(got.view(np.uint64), a.view(np.uint64)) if a.size else True, name
    for k, (v, t) in meta.items():
        got, dtype = h5dump.attribute(path, f"/metadata/{k}")
        assert dtype == ("H5T_STD_I32LE" if t == "i" else "H5T_IEEE_F64LE"), (k, dtype)
        assert got == v, (k, got, v)
    for name, vals in spmeta.items():
        got, dtype = h5dump.attribute(path, f"/metadata_species/{name}/name")
        assert (got, dtype) == (name[0], "H5T_STRING[1]")  # PredType::C_S1 keeps one character
        for k, v in vals.items():
            got, dtype = h5dump.attribute(path, f"/metadata_species/{name}/{k}")
            assert got == v, (name, k, got, v)


@pytest.mark.skipif(not h5dump.available(), reason="no h5dump in this image")
def test_large_group_and_long_names(tmp_path):
    """A field group with a dataset per write (here 300 links, stored
    compactly in one object header) and names past 255 bytes."""
    path = tmp_path / "big.h5"
    f = _lib.H5File(path)
    g = f.group(f.group(f.ROOT, "fielddata"), "pot")
    for ts in range(300):
        f.dataset(g, str(ts), np.full((3, 2), float(ts)))
    long = "x" * 300
    f.dataset(f.ROOT, long, np.arange(4.0))
    f.close()
    tree = h5dump.contents(path)
    assert sum(1 for k in tree if k.startswith("/fielddata/pot/")) == 300
    got, _ = h5dump.dataset(path, "/fielddata/pot/299")
    assert np.array_equal(got, np.full((3, 2), 299.0))
    got, _ = h5dump.dataset(path, "/" + long)
    assert np.array_equal(got, np.arange(4.0))


def test_writer_rejects_bad_requests(tmp_path):
    f = _lib.H5File(tmp_path / "r.h5")
    g = f.group(f.ROOT, "a")
    assert f.group(f.ROOT, "a") == g  # an existing group is returned
    f.dataset(g, "d", np.zeros(3))
    with pytest.raises(_lib.EpicError):
        f.dataset(g, "d", np.zeros(3))  # name taken
    with pytest.raises(_lib.EpicError):
        f.group(g, "d")  # a dataset, not a group
    with pytest.raises(_lib.EpicError):
        f.group(g, "b/c")
    with pytest.raises(_lib.EpicError):
        f.dataset(g, "e", np.zeros((1, 1, 1, 1, 1)))  # rank above 4
    f.attr_i32(g, "n", 1)
    with pytest.raises(_lib.EpicError):
        f.attr_f64(g, "n", 2.0)
    f.close()
    with pytest.raises(_lib.EpicError):
        _lib.H5File(tmp_path / "missing_dir" / "x.h5")
