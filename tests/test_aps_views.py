"""Input stages of the APS driver: the time-interlaced view schedule
(mbir_interlaced_views, host C) against the APS program's own
gen_interlaced_views_0_to_Inf (XT_Views.c, the reference build), and the
oracle's flat-field normalisation (read_data, XT_HDFIO.c:138-178; that file
needs libhdf5, absent here, so the oracle is pinned by closed forms) --
the GPU kernel is checked against it in test_gpu_aps.py."""
import numpy as np
import pytest

from oracle import oracle as O

CASES = [
    (4, 256, 256, 0, 0.0, 0.0),
    (8, 720, 720, 5, 0.0, 0.0),
    (16, 1024, 300, 0, 0.01, 0.25),    # views closer than rot_speed * min_acq_time dropped
    (2, 90, 200, 17, 0.001, 3.0),
    (1, 64, 64, 0, 0.0, 0.0),
]


@pytest.mark.skipif(not O.views_reference_available(), reason="reference build absent")
@pytest.mark.parametrize("case", CASES)
def test_interlaced_views_match_reference(case):
    from phasecon_amd.mbir import interlaced_views
    K, n_theta, n, start, acq, speed = case
    rv, rt, rn = O.reference_interlaced_views(K, n_theta, n, start, acq, speed)
    v, t = interlaced_views(K, n_theta, n, start, acq, speed)
    assert rn == start + n
    assert np.array_equal(v, rv) and np.array_equal(t, rt)


def test_interlaced_views_closed_form():
    """K = 4, N_theta = 16 (L = 4): the first 16 candidates sweep 0..15 with
    the sub-sequence offsets bit-reversed over two bits (0, 2, 1, 3)."""
    from phasecon_amd.mbir import interlaced_views
    v, t = interlaced_views(4, 16, 16)
    idx = np.array([4 * i + [0, 2, 1, 3][i // 4] for i in range(16)], np.float32)
    assert np.array_equal(t, idx / np.float32(4))
    assert np.array_equal(v, idx * np.float32(np.pi / 16))
    from phasecon_amd.synth import interlaced_views as synth_views
    sv, st = synth_views(4, 16, 16)
    assert np.allclose(sv, v) and np.allclose(st, t)


def test_aps_closed_form():
    """Uniform white 1000 and dark 100 frames, projections constant over
    each 2 x 2 block: t = log(900 / (p - 100)) exactly as written."""
    rng = np.random.default_rng(0)
    rows, cols, ratio = 3, 5, 2
    block = rng.integers(150, 1000, (4, rows, cols))
    proj = np.repeat(np.repeat(block, ratio, 1), ratio, 2).astype(np.uint16)
    white = np.full((3, rows * ratio, cols * ratio), 1000, np.uint16)
    dark = np.full((2, rows * ratio, cols * ratio), 100, np.uint16)
    p, w = O.aps_normalize(proj, white, dark, ratio)
    tt = np.log(900.0 / (block - 100.0))
    assert np.array_equal(w, (4 * (block - 100)).astype(np.float32))
    # the quotient is rounded to float before the log: |dt| <= 2^-23
    assert np.allclose(p, 0.5 * tt * tt + tt, rtol=1e-6, atol=5e-7)


def test_aps_ratio_one_and_frame_means():
    """ratio 1: the weight is |proj - mean dark| and t = log(|white - dark|
    / weight) per pixel; means of two frames are exact half-integers."""
    rng = np.random.default_rng(1)
    proj = rng.integers(0, 4000, (2, 4, 6)).astype(np.uint16)
    white = np.stack([np.full((4, 6), 3000), np.full((4, 6), 3001)]).astype(np.uint16)
    dark = np.stack([np.full((4, 6), 10), np.full((4, 6), 13)]).astype(np.uint16)
    p, w = O.aps_normalize(proj, white, dark, 1, bh_coef=0.0)
    assert np.array_equal(w, np.abs(proj.astype(np.float32) - np.float32(11.5)))
    with np.errstate(divide="ignore"):
        tt = np.log(np.float32(2989.0) / w.astype(np.float64))
    assert np.array_equal(np.isfinite(p), np.isfinite(tt))
    assert np.allclose(p[np.isfinite(p)], tt[np.isfinite(tt)], rtol=1e-6, atol=5e-7)


def test_std_downsample_closed_form():
    """Standard format: blocks of equal values keep their value (b = 0),
    the weights add up; with b the polynomial b t^2 + t."""
    rng = np.random.default_rng(2)
    block = rng.random((3, 4, 5)).astype(np.float32)
    wblk = rng.random((3, 4, 5)).astype(np.float32)
    up = lambda a: np.repeat(np.repeat(a, 2, 1), 2, 2)
    p, w = O.std_downsample(up(block), up(wblk), 2)
    assert np.array_equal(p, block)
    assert np.allclose(w, 4 * wblk, rtol=1e-6)
    p, _ = O.std_downsample(up(block), up(wblk), 2, bh_coef=0.5)
    t = block.astype(np.float64)
    assert np.allclose(p, 0.5 * t * t + t, rtol=1e-6)
