"""Ring-offset and noise-variance updates of the oracle (orc_offset_update,
orc_variance_update): against the reference's own update_Sinogram_Offset /
update_variance_parameter (the reference build, where present) and against
what each update must satisfy whatever the code: mode 1 leaves no weighted
residual at any detector element (sum_p lambda e = 0, lambda the
generalized-Huber weight), mode 2 the same up to one common constant with
sum d = 0, mode 3 keeps <window_m, d> = 0 for every constraint window; the
variance update gives the mean weighted squared residual when every entry is
selected."""
import numpy as np
import pytest

from oracle import oracle as O

T, DELTA = 2.0, 0.5


def _case(N_p=24, N_r=32, N_t=4, seed=0):
    rng = np.random.default_rng(seed)
    w = (rng.random((N_p, N_r, N_t)) * 100 + 10).astype(np.float32)
    e = (rng.standard_normal((N_p, N_r, N_t)) / np.sqrt(w) * 1.6 + 0.01).astype(np.float32)
    d = (rng.standard_normal((N_r, N_t)) * 0.01).astype(np.float32)
    sel = (np.abs(e * np.sqrt(w)) < T).astype(np.uint8)
    return w, e, d, sel


def _lam(w, e, sel):
    w, e = w.astype(np.float64), e.astype(np.float64)
    return np.where(sel > 0, w, T * DELTA * np.sqrt(w) / np.abs(e))


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_offset_update_properties(mode):
    w, e, d, sel = _case()
    lam = _lam(w, e, sel)
    oe, od, osel = O.offset_update(mode, w, e, d, sel, T, DELTA)
    # e and d move together: e + d is unchanged
    assert np.allclose(oe + od[None], e.astype(np.float64) + d[None], atol=1e-12)
    # the residual the step leaves, weighted with the weights it used
    r = (lam * oe).sum(axis=0)
    if mode == 1:
        assert np.max(np.abs(r)) <= 1e-9 * np.max(np.abs((lam * e).sum(axis=0)))
    elif mode == 2:
        assert abs(od.sum()) <= 1e-12 * np.abs(od).sum()
        assert np.ptp(r) <= 1e-9 * np.max(np.abs((lam * e).sum(axis=0)))
    else:
        win = O.offset_windows(*d.shape)
        assert win.shape[0] > 1
        assert np.max(np.abs((win * od[None]).sum(axis=(1, 2)))) <= 1e-10 * np.abs(od).sum()
    assert np.array_equal(osel, (np.abs(oe * np.sqrt(w)) < T).astype(np.uint8))


def test_offset_windows_closed_form():
    """Triangular windows: N_r = 16, N_t = 4 -> sizes 8 x 4, steps 4 x 2."""
    win = O.offset_windows(16, 4)
    assert win.shape == (4 * 2, 16, 4)
    # first window: (k+1) for k < 4, 8-k after, times (l+1) for l < 2, 4-l after
    a = np.array([1, 2, 3, 4, 4, 3, 2, 1] + [0] * 8, float)
    b = np.array([1, 2, 2, 1], float)
    assert np.array_equal(win[0], np.outer(a, b))


def test_variance_update_all_selected():
    w, e, _, _ = _case(seed=1)
    sel = np.ones_like(w, np.uint8)
    var0 = 1.7
    v, ow, _ = O.variance_update(var0, w, e, sel, T, DELTA)
    raw = w.astype(np.float64) * var0
    assert np.isclose(v, np.mean(e.astype(np.float64) ** 2 * raw), rtol=1e-12)
    assert np.allclose(ow, raw / v, rtol=1e-12)


needs_ref = pytest.mark.skipif(not O.reference_available(), reason="reference build absent")


@needs_ref
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_offset_update_matches_reference(mode, tiny_scan):
    R = O.Reference(tiny_scan, quad_convex=0, remove_rings=mode, huber_T=T, huber_delta=DELTA)
    N_p, N_r, N_t = R.shape_e
    w, e, d, sel = _case(N_p, N_r, N_t)
    re_, rd, rs = R.offset_update(mode, w, e, d, sel)
    R.close()
    oe, od, os_ = O.offset_update(mode, w, e, d, sel, T, DELTA)
    assert np.max(np.abs(re_ - oe)) <= 1e-6 * np.max(np.abs(e))
    assert np.max(np.abs(rd - od)) <= 1e-6 * max(np.max(np.abs(od)), 1e-3)
    assert np.array_equal(rs, os_)


@needs_ref
def test_variance_update_matches_reference(tiny_scan):
    R = O.Reference(tiny_scan, quad_convex=0, remove_rings=0, huber_T=T, huber_delta=DELTA)
    w, e, _, sel = _case(*R.shape_e, seed=2)
    rv, rw, rs = R.variance_update(1.3, w, e, sel)
    R.close()
    ov, ow, os_ = O.variance_update(1.3, w, e, sel, T, DELTA)
    assert abs(rv - ov) <= 1e-8 * ov
    assert np.max(np.abs(rw - ow) / ow) <= 1e-6
    assert np.array_equal(rs, os_)
