"""Voxel-selection criterion of non-homogeneous ICD in the oracle (vsc,
vsc_select): against the reference's own ComputeVSC and
VSC_based_Voxel_Line_Select (the reference build, where present), and
against what fixes them whatever the code: the filter's table in
XT_NHICD.c:128-132, its impulse response and its truncation at the edges,
and the selection's count on distinct values."""
import numpy as np
import pytest

from oracle import oracle as O

# HammingWindow of XT_NHICD.c:128-132
TABLE = np.array([[0.0064, 0.0432, 0.0800, 0.0432, 0.0064],
                  [0.0432, 0.2916, 0.5400, 0.2916, 0.0432],
                  [0.0800, 0.5400, 1.0000, 0.5400, 0.0800],
                  [0.0432, 0.2916, 0.5400, 0.2916, 0.0432],
                  [0.0064, 0.0432, 0.0800, 0.0432, 0.0064]])


def test_vsc_impulse_response_is_the_reference_table():
    m = np.zeros((9, 11), np.float32)
    m[4, 5] = 1.0
    f = O.vsc(m)
    assert np.allclose(f[2:7, 3:8], TABLE, atol=1e-7)
    f[2:7, 3:8] = 0
    assert not f.any()


def test_vsc_truncates_at_the_edges():
    f = O.vsc(np.ones((8, 8), np.float32))
    edge = 1.0 + 0.54 + 0.08
    assert np.isclose(f[0, 0], edge * edge, rtol=1e-6)
    assert np.isclose(f[4, 4], 2.24 * 2.24, rtol=1e-6)   # the full window's sum
    assert np.isclose(f[0, 4], edge * 2.24, rtol=1e-6)


def test_vsc_select_takes_the_top_tenth():
    rng = np.random.default_rng(3)
    filt = rng.random((16, 16)).astype(np.float32)
    sel = O.vsc_select(filt, 200)
    assert sel.sum() == 20
    assert filt[sel].min() > filt[~sel].max()
    assert not O.vsc_select(filt, 9).any()      # floor(9 / 10) = 0 lines


needs_ref = pytest.mark.skipif(not O.reference_available(), reason="reference build absent")


@needs_ref
def test_vsc_matches_reference(tiny_scan):
    R = O.Reference(tiny_scan)
    rng = np.random.default_rng(5)
    mag = (rng.random((R.N_y, R.N_x)) ** 3).astype(np.float32)
    ref = R.vsc(mag)
    R.close()
    assert np.allclose(O.vsc(mag), ref, rtol=1e-6, atol=1e-7)


@needs_ref
def test_vsc_select_matches_reference(tiny_scan):
    R = O.Reference(tiny_scan)
    rng = np.random.default_rng(6)
    mag = (rng.random((R.N_time, R.N_y, R.N_x)) ** 2).astype(np.float32)
    count = (R.N_y * R.N_x * 3) // 4
    ref = R.vsc_select(mag, count)
    R.close()
    for t in range(mag.shape[0]):
        ours = O.vsc_select(O.vsc(mag[t]), count)
        assert ours.sum() == count // 10
        assert np.array_equal(ours, ref[t])
