"""The product's HDF5 input stages (include/mbir_hdf5.h: file reading on the
host, normalisation on the GPU) against the reference drivers' own readers
(read_data; read_ProjWeightData + read_AngleTimeReconList, compiled from
their sources under oracle/_ref) on the same HDF5 files: weights and the
standard format's projections, angles and times bit-exact; APS projections
within one float ulp (the device's double log against the host's).  Also
the errors the product adds: views beyond the data, a missing dataset."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import test_hdf5_io as H

pytestmark = [pytest.mark.gpu, H.needs_ref]


def _prod():
    L = C.CDLL(H.PROD_SO)
    L.mbir_read_aps_hdf5.restype = C.c_int32
    L.mbir_read_std_hdf5.restype = C.c_int32
    return L


def _aps(L, paths, s, start=None, num=None):
    start = s["start"] if start is None else start
    num = s["num"] if num is None else num
    n = num * s["proj_rows"] * s["proj_cols"]
    proj, wts = np.zeros(n, np.float32), np.zeros(n, np.float32)
    rc = L.mbir_read_aps_hdf5(*[p.encode() for p in paths], H.p_(proj), H.p_(wts), s["row0"],
                              s["proj_rows"], s["proj_cols"], start, num, H.devnull_file())
    shape = (num, s["proj_rows"], s["proj_cols"])
    return rc, proj.reshape(shape), wts.reshape(shape)


def test_aps_reader_matches_reference():
    R = H._lib(H.APS_SO)
    P = _prod()
    with tempfile.TemporaryDirectory() as d:
        paths, _, _, _ = H.aps_files(R, d, seed=3)
        print("APS reader: 7 views x 20 rows x 70 channels -> 4 x 4 x 16", flush=True)
        rp, rw = H.ref_read_aps(R, paths)
        rc, p, w = _aps(P, paths, H.APS)
        assert rc == 0
        assert np.array_equal(w, rw)
        np.testing.assert_array_max_ulp(p, rp, maxulp=1)
        assert np.mean(p == rp) >= 0.99
        # views beyond the data's frames: refused (the reference reads past them)
        rc, _, _ = _aps(P, paths, H.APS, start=5, num=4)
        assert rc == -1


def test_std_reader_matches_reference():
    R = H._lib(H.STD_SO)
    P = _prod()
    s = H.STD
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "std.h5")
        H.std_file(R, path, seed=4)
        print("standard reader: 6 views x 14 rows x 37 channels -> 4 x 3 x 9", flush=True)
        rp, rw, ra, rt = H.ref_read_std(R, path)
        n = s["num"] * s["proj_rows"] * s["proj_cols"]
        proj, wts = np.zeros(n, np.float32), np.zeros(n, np.float32)
        ang, tim = np.zeros(s["num"], np.float32), np.zeros(s["num"], np.float32)
        rc = P.mbir_read_std_hdf5(path.encode(), H.p_(proj), H.p_(wts), H.p_(ang), H.p_(tim),
                                  s["row0"], s["proj_rows"], s["proj_cols"], s["start"],
                                  s["num"], H.devnull_file())
        assert rc == 0
        assert np.array_equal(proj.reshape(rp.shape), rp)
        assert np.array_equal(wts.reshape(rw.shape), rw)
        assert np.array_equal(ang, ra) and np.array_equal(tim, rt)
        # a file without /weights
        bad = os.path.join(d, "bad.h5")
        H.write(R, bad, 1, "/projections", np.zeros((2, 4, 9), np.float32))
        rc = P.mbir_read_std_hdf5(bad.encode(), H.p_(proj), H.p_(wts), H.p_(ang), H.p_(tim), 0,
                                  1, 9, 0, 1, H.devnull_file())
        assert rc == -1
