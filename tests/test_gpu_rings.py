"""The drop-in reconstruct() with ring-offset estimation and the
generalized-Huber data term with noise-variance re-estimation, end to end.

A Shepp-Logan scan gets a known detector offset added to every view (rings
in the image).  With remove_rings = 3 the estimated offset (proj_offset.bin)
must follow the injected one and the final MAP cost must fall back most of
the way to the clean scan's.  With quad_convex = 0 the variance estimate
(variance_estimate.bin) must agree with the reference's reconstruct() on the
same scan.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ringed(scan, amp=0.1, seed=4):
    rng = np.random.default_rng(seed)
    rows, cols = scan["proj_rows"], scan["proj_cols"]
    off = np.zeros((rows, cols))
    cols_hit = rng.choice(cols, cols // 8, replace=False)
    off[:, cols_hit] = amp * rng.standard_normal((rows, cols_hit.size))
    s = dict(scan)
    s["projections"] = (scan["projections"] + off[None]).astype(np.float32)
    return s, off


def _run(scan, tmp_path, name, **kw):
    from phasecon_amd.mbir import reconstruct
    d = tmp_path / name
    d.mkdir()
    cwd = os.getcwd()
    os.chdir(d)
    try:
        obj, equits, cost = reconstruct(scan, convg_thresh=0.1, **kw)
    finally:
        os.chdir(cwd)
    return obj[:, 1:-1], d, cost


@pytest.fixture(scope="module")
def shepp_scan():
    from phasecon_amd.synth import make_scan
    return make_scan("shepp3d", seed=0)


def test_ring_offset_is_recovered(shepp_scan, tmp_path):
    """Shepp-Logan (inside the reconstruction mask: an object outside it
    leaves a residual the offset would absorb).  Mode 3 (windowed zero-sum
    constraints) recovers the injected offset, and with it the MAP cost
    returns most of the way to the clean scan's; mode 2 (sum d = 0 only) can
    trade smooth, radially symmetric residuals between d and the image, so
    only its constraint is checked."""
    from oracle.oracle import offset_windows
    scan, off = _ringed(shepp_scan)
    _, _, c_clean = _run(shepp_scan, tmp_path, "clean")
    _, _, c_plain = _run(scan, tmp_path, "plain")
    _, d3, c_fixed = _run(scan, tmp_path, "rings3", remove_rings=3)
    _, d2, _ = _run(scan, tmp_path, "rings2", remove_rings=2)
    shape = (scan["proj_cols"], scan["proj_rows"])
    est = np.fromfile(d3 / "proj_offset.bin", np.float32).reshape(shape).T
    est2 = np.fromfile(d2 / "proj_offset.bin", np.float32).reshape(shape).T
    true = off - off.mean()
    corr = float(np.sum(est * true) / np.sqrt(np.sum(est ** 2) * np.sum(true ** 2)))
    win = offset_windows(*shape)                      # [M][N_r][N_t]
    viol = np.max(np.abs((win * est.T[None]).sum(axis=(1, 2))))
    print(f"mode 3: offset correlation {corr:.4f}, max |<A_m, d>| {viol:.2g}; final cost "
          f"{c_clean:.8g} clean scan, {c_plain:.8g} with rings, {c_fixed:.8g} with rings removed; "
          f"mode 2: |sum d| {abs(est2.sum()):.2g}", flush=True)
    assert corr > 0.8
    assert viol <= 1e-4 * np.abs(est).sum()
    assert abs(c_fixed - c_clean) < 0.5 * abs(c_plain - c_clean)
    assert abs(est2.sum()) <= 1e-3 * np.abs(est2).sum()


def test_variance_estimate_matches_reference(shepp_scan, tmp_path):
    """quad_convex = 0 (generalized Huber, T = 4, delta = 0.5): the variance
    estimate after the second stage against the reference's own
    reconstruct() on the same scan (tests/golden/variance_shepp3d.json,
    written by make_variance_golden.py; the scan's SHA-256 is checked).
    The estimate is a residual statistic of the image each schedule stops
    at (both at the 0.1 % update threshold, different visiting orders), so
    the bound is 10 %."""
    import json
    sys_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    gold = json.load(open(os.path.join(sys_path, "variance_shepp3d.json")))
    import sys
    sys.path.insert(0, sys_path)
    from make_variance_golden import digest
    assert gold["digest"] == digest(shepp_scan)
    out = {}
    for th in (0.1, 0.02):
        from phasecon_amd.mbir import reconstruct
        d = tmp_path / f"huber{th}"
        d.mkdir()
        cwd = os.getcwd()
        os.chdir(d)
        try:
            reconstruct(shepp_scan, convg_thresh=th, quad_convex=0, huber_T=4.0, huber_delta=0.5)
        finally:
            os.chdir(cwd)
        out[th] = float(np.fromfile(d / "variance_estimate.bin", np.float64)[0])
    print(f"variance estimate {out[0.1]:.6g} (threshold 0.1 %), {out[0.02]:.6g} (0.02 %), "
          f"reference {gold['variance_estimate']:.6g}", flush=True)
    assert abs(out[0.1] - gold["variance_estimate"]) <= 0.1 * gold["variance_estimate"]
