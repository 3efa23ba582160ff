"""The real-data drivers' HDF5 input stages (row f3).

CPU: the reference's own readers -- the APS driver's read_data and the
standard driver's read_ProjWeightData / read_AngleTimeReconList, compiled
from their sources under oracle/_ref -- read small HDF5 files written here,
and the oracle's restatements of their arithmetic (oracle.aps_normalize,
oracle.std_downsample) must give the same arrays from the same hyperslabs
(white and dark frames 1 .. N-2, rows from datafile_row0, the channels'
remainder split off both sides): weights bit-exact, projections within one
float ulp (the double log).  The product library's exports are checked too.
GPU: the product's readers (include/mbir_hdf5.h) against the reference
readers on the same files.
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from conftest import ROOT

REF = os.path.join(ROOT, "oracle", "_ref")
APS_SO = os.path.join(REF, "libhdfio_aps_ref.so")
STD_SO = os.path.join(REF, "libhdfio_std_ref.so")
PROD_SO = os.path.join(ROOT, "phasecon_amd", "libmbir4d_hdf5.so")

needs_ref = pytest.mark.skipif(not (os.path.exists(APS_SO) and os.path.exists(STD_SO)),
                               reason="reference HDF5 readers not built (make -C oracle ref)")
p_ = lambda a: a.ctypes.data_as(C.c_void_p)


def devnull_file():
    """A C FILE* on /dev/null for the readers' messages."""
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    return C.c_void_p(libc.fopen(b"/dev/null", b"w"))


def _lib(path):
    L = C.CDLL(path)
    L.h5x_mpi_up()
    return L


def write(L, path, create, name, arr):
    arr = np.ascontiguousarray(arr)
    dims = np.array(arr.shape, np.int64)
    kind = 0 if arr.dtype == np.uint16 else 1
    rc = L.h5x_write(path.encode(), int(create), name.encode(), arr.ndim, p_(dims), kind, p_(arr))
    assert rc == 0, f"writing {name} to {path}"


# APS fixture: 7 views x 20 rows x 70 channels; proj_cols 16 -> ratio 4,
# 6 remainder channels (3 off each side); proj_rows 4 -> 16 file rows from 2
APS = dict(views=7, rows=20, chans=70, proj_rows=4, proj_cols=16, row0=2, start=2, num=4,
           whites=5, darks=4)


def aps_files(L, d, seed=0):
    r = np.random.default_rng(seed)
    s = APS
    data = r.integers(800, 30000, (s["views"], s["rows"], s["chans"])).astype(np.uint16)
    white = r.integers(30000, 40000, (s["whites"], s["rows"], s["chans"])).astype(np.uint16)
    dark = r.integers(80, 120, (s["darks"], s["rows"], s["chans"])).astype(np.uint16)
    paths = [os.path.join(d, n) for n in ("data.h5", "white.h5", "dark.h5")]
    write(L, paths[0], 1, "/exchange/data", data)
    write(L, paths[1], 1, "/exchange/data_white", white)
    write(L, paths[2], 1, "/exchange/data_dark", dark)
    return paths, data, white, dark


def ref_read_aps(L, paths, s=APS):
    n = s["num"] * s["proj_rows"] * s["proj_cols"]
    proj = np.zeros(n, np.float32)
    wts = np.zeros(n, np.float32)
    rc = L.read_data(*[p.encode() for p in paths], p_(proj), p_(wts), s["row0"], s["proj_rows"],
                     s["proj_cols"], s["start"], s["num"], devnull_file())
    assert rc == 0
    shape = (s["num"], s["proj_rows"], s["proj_cols"])
    return proj.reshape(shape), wts.reshape(shape)


def aps_slabs(data, white, dark, s=APS):
    ratio = (s["chans"] - s["chans"] % s["proj_cols"]) // s["proj_cols"]
    c0 = (s["chans"] % s["proj_cols"]) // 2
    rows = slice(s["row0"], s["row0"] + s["proj_rows"] * ratio)
    chans = slice(c0, c0 + s["proj_cols"] * ratio)
    return (data[s["start"]:s["start"] + s["num"], rows, chans], white[1:-1, rows, chans],
            dark[1:-1, rows, chans], ratio)


@needs_ref
def test_aps_oracle_matches_reference_read_data():
    from oracle.oracle import aps_normalize
    L = _lib(APS_SO)
    with tempfile.TemporaryDirectory() as d:
        paths, data, white, dark = aps_files(L, d)
        rp, rw = ref_read_aps(L, paths)
    p, wh, dk, ratio = aps_slabs(data, white, dark)
    op, ow = aps_normalize(p, wh, dk, ratio, bh_coef=0.5)
    assert np.array_equal(ow, rw)
    np.testing.assert_array_max_ulp(op, rp, maxulp=1)


STD = dict(views=6, rows=14, chans=37, proj_rows=3, proj_cols=9, row0=1, start=1, num=4)


def std_file(L, path, seed=1):
    r = np.random.default_rng(seed)
    s = STD
    proj = (r.random((s["views"], s["rows"], s["chans"])) * 3).astype(np.float32)
    wts = r.random((s["views"], s["rows"], s["chans"])).astype(np.float32)
    ang = (r.random(s["views"]) * 3).astype(np.float32)
    tim = np.arange(s["views"], dtype=np.float32) * 0.5
    write(L, path, 1, "/projections", proj)
    write(L, path, 0, "/weights", wts)
    write(L, path, 0, "/angles", ang)
    write(L, path, 0, "/times", tim)
    return proj, wts, ang, tim


def ref_read_std(L, path, s=STD):
    n = s["num"] * s["proj_rows"] * s["proj_cols"]
    proj, wts = np.zeros(n, np.float32), np.zeros(n, np.float32)
    ang, tim = np.zeros(s["num"], np.float32), np.zeros(s["num"], np.float32)
    log = devnull_file()
    assert L.read_ProjWeightData(path.encode(), p_(proj), p_(wts), s["row0"], s["proj_rows"],
                                 s["proj_cols"], s["start"], s["num"], log) == 0
    assert L.read_AngleTimeReconList(path.encode(), p_(ang), p_(tim), s["start"], s["num"],
                                     log) == 0
    shape = (s["num"], s["proj_rows"], s["proj_cols"])
    return proj.reshape(shape), wts.reshape(shape), ang, tim


@needs_ref
def test_std_oracle_matches_reference_readers():
    from oracle.oracle import std_downsample
    L = _lib(STD_SO)
    s = STD
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "std.h5")
        proj, wts, ang, tim = std_file(L, path)
        rp, rw, ra, rt = ref_read_std(L, path)
    ratio = (s["chans"] - s["chans"] % s["proj_cols"]) // s["proj_cols"]
    c0 = (s["chans"] % s["proj_cols"]) // 2
    sl = (slice(s["start"], s["start"] + s["num"]),
          slice(s["row0"], s["row0"] + s["proj_rows"] * ratio),
          slice(c0, c0 + s["proj_cols"] * ratio))
    op, ow = std_downsample(proj[sl], wts[sl], ratio, 0.0)
    assert np.array_equal(ow, rw) and np.array_equal(op, rp)
    assert np.array_equal(ra, ang[sl[0]]) and np.array_equal(rt, tim[sl[0]])


@pytest.mark.skipif(not os.path.exists(PROD_SO), reason="libmbir4d_hdf5.so not built")
def test_product_hdf5_library_exports():
    import re
    L = C.CDLL(PROD_SO)
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "mbir_hdf5.h")).read(),
                  flags=re.S)
    names = set(re.findall(r"\b(mbir_[a-z0-9_]+)\s*\(", text))
    assert names == {"mbir_read_aps_hdf5", "mbir_read_std_hdf5"}
    for n in names:
        assert hasattr(L, n), n
