"""The GPU flat-field normalisation (mbir_hip_aps_normalize) against the
oracle's restatement of read_data (XT_HDFIO.c:138-178) on the same raw
counts: weights bit-exact (float sums in the reference's order);
projections bit-exact except where the device's and the host's double log
round differently, and never by more than one float ulp."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,rows,cols,ratio,nw,nd", [
    (3, 8, 12, 1, 2, 2),
    (16, 32, 48, 2, 5, 3),
    (90, 32, 64, 4, 8, 8),
])
def test_aps_normalize_matches_oracle(n, rows, cols, ratio, nw, nd):
    from oracle.oracle import aps_normalize as ref
    from phasecon_amd.mbir import aps_normalize
    rng = np.random.default_rng(n)
    R, Cc = rows * ratio, cols * ratio
    dark = rng.integers(90, 110, (nd, R, Cc)).astype(np.uint16)
    white = rng.integers(20000, 40000, (nw, R, Cc)).astype(np.uint16)
    proj = rng.integers(200, 20000, (n, R, Cc)).astype(np.uint16)
    proj[0, :ratio, :ratio] = 100                      # a block at the dark level
    if nd == 2:
        dark[:, :ratio, :ratio] = 100                  # its weight is exactly 0 -> t = inf
    print(f"aps n={n} rows={rows} cols={cols} ratio={ratio} white={nw} dark={nd}", flush=True)
    p, w = aps_normalize(proj, white, dark, ratio)
    rp, rw = ref(proj, white, dark, ratio)
    assert np.array_equal(w, rw)
    fin = np.isfinite(rp)
    assert np.array_equal(np.isfinite(p), fin)
    assert np.array_equal(p[~fin], rp[~fin])
    exact = np.mean(p[fin] == rp[fin])
    print(f"projections bit-exact {exact:.6f}", flush=True)
    np.testing.assert_array_max_ulp(p[fin], rp[fin], maxulp=1)
    assert exact >= 0.999


@pytest.mark.parametrize("n,rows,cols,ratio,bh", [(4, 8, 12, 1, 0.0), (20, 16, 32, 4, 0.0),
                                                   (6, 8, 16, 2, 0.5)])
def test_std_downsample_matches_oracle(n, rows, cols, ratio, bh):
    """The standard-format down-sampling: bit-exact (float sums in the
    reference's order, an exact double polynomial rounded once)."""
    from oracle.oracle import std_downsample as ref
    from phasecon_amd.mbir import std_downsample
    rng = np.random.default_rng(n)
    proj = rng.random((n, rows * ratio, cols * ratio)).astype(np.float32) * 3
    wts = rng.random((n, rows * ratio, cols * ratio)).astype(np.float32)
    print(f"std n={n} rows={rows} cols={cols} ratio={ratio} bh={bh}", flush=True)
    p, w = std_downsample(proj, wts, ratio, bh)
    rp, rw = ref(proj, wts, ratio, bh)
    assert np.array_equal(w, rw)
    assert np.array_equal(p, rp)
