"""GPU ring-offset and noise-variance updates (mbir_hip_offset_update,
mbir_hip_variance_update) against the oracle (orc_offset_update,
orc_variance_update, pinned against the reference's own routines): the same
error sinogram, weights and offset in, e, d, w and the variance out within
1e-6 relative (FP64 sums on both sides, FP32 storage)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

T, DELTA = 2.0, 0.5


def _stage(scan):
    from phasecon_amd.mbir import Stage
    return Stage(scan, quad_convex=0, huber_T=T, huber_delta=DELTA)


def _state(st, seed):
    rng = np.random.default_rng(seed)
    w = st.get_weight()
    e = (rng.standard_normal(st.shape_e) / np.sqrt(w) * 1.6 + 0.01).astype(np.float32)
    st.set_error(e)
    return w, e


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_offset_update_matches_oracle(mode, small4d_scan):
    st = _stage(small4d_scan)
    w, e = _state(st, mode)
    d0 = np.zeros((st.g.N_r, st.g.N_t))
    sel = (np.abs(e.astype(np.float64) * np.sqrt(w.astype(np.float64))) < T).astype(np.uint8)
    oe, od, _ = O.offset_update(mode, w, e, d0, sel, T, DELTA)
    for k in range(2):  # a second update from the first one's result
        st.offset_update(mode)
        if k == 0:
            ge, gd = st.get_error(), st.get_offset()
            print(f"mode {mode}: max |e - oracle| {np.max(np.abs(ge - oe)):.3g}, "
                  f"max |d - oracle| {np.max(np.abs(gd - od)):.3g} (max |d| {np.max(np.abs(od)):.3g})",
                  flush=True)
            assert np.max(np.abs(ge - oe)) <= 1e-6 * np.max(np.abs(e))
            assert np.max(np.abs(gd - od)) <= 1e-6 * max(np.max(np.abs(od)), 1e-3)
            sel = (np.abs(oe * np.sqrt(w.astype(np.float64))) < T).astype(np.uint8)
            oe, od, _ = O.offset_update(mode, w, oe, od, sel, T, DELTA)
    ge, gd = st.get_error(), st.get_offset()
    st.close()
    assert np.max(np.abs(ge - oe)) <= 2e-6 * np.max(np.abs(e))
    assert np.max(np.abs(gd - od)) <= 2e-6 * max(np.max(np.abs(od)), 1e-3)


def test_variance_update_matches_oracle(small4d_scan):
    st = _stage(small4d_scan)
    w, e = _state(st, 7)
    sel = (np.abs(e.astype(np.float64) * np.sqrt(w.astype(np.float64))) < T).astype(np.uint8)
    ov, ow, _ = O.variance_update(1.0, w, e, sel, T, DELTA)
    v = st.variance_update()
    gw = st.get_weight()
    st.close()
    print(f"variance gpu {v:.12g} oracle {ov:.12g}", flush=True)
    assert abs(v - ov) <= 1e-9 * ov
    assert np.max(np.abs(gw - ow) / ow) <= 1e-6
