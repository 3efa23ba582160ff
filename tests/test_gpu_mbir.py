"""GPU parity of the MBIR hot path against the FP64 oracle (oracle/).

Tolerances (FP32 kernels vs FP64 restatement): A values and projections
within 1e-6 relative, theta1/theta2 and the functional substitution within
1e-5 relative (the FP32 bar for these terms), FP64 costs within 1e-9 relative,
converged ICD within 1e-5 relative MAP cost and 0.5 % image RMSE.
"""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(tiny_scan):
    from oracle.oracle import Oracle
    from phasecon_amd.mbir import Stage
    st = Stage(tiny_scan)
    o = Oracle(tiny_scan)
    yield st, o
    st.close()


@pytest.fixture(scope="module")
def pair4d(small4d_scan):
    from oracle.oracle import Oracle
    from phasecon_amd.mbir import Stage
    st = Stage(small4d_scan)
    o = Oracle(small4d_scan)
    yield st, o
    st.close()


def _slot_views(st):
    return st.frame_views[: st.frame_ptr[-1]]


@pytest.mark.parametrize("which", ["pair", "pair4d"])
def test_amatrix_matches_oracle(which, request):
    st, o = request.getfixturevalue(which)
    print(f"A build parity on {st.scan['name']}", flush=True)
    st.build_amatrix()
    first, val = st.get_amatrix()
    ofirst, oval = o.amatrix_all()
    sv = _slot_views(st)
    ofirst = ofirst[:, sv]
    oval = oval[:, sv]
    assert np.array_equal(first, ofirst)
    assert rel_err(val, oval) < 1e-6


@pytest.mark.parametrize("which", ["pair", "pair4d"])
def test_forward_projection_matches_oracle(which, request):
    st, o = request.getfixturevalue(which)
    rng = np.random.default_rng(5)
    x = (rng.random(st.shape_x) * 0.02).astype(np.float32)
    ax = st.forward_project(x)
    oax = o.forward_project(x.astype(np.float64))
    assert rel_err(ax, oax) < 1e-5


def test_init_error_matches_oracle(pair):
    st, o = pair
    rng = np.random.default_rng(6)
    x = (rng.random(st.shape_x) * 0.02).astype(np.float32)
    st.set_object(x)
    st.build_amatrix()
    st.init_error()
    e = st.get_error()
    oe = o.init_error(x.astype(np.float64))
    assert rel_err(e, oe) < 1e-5
    assert np.array_equal(st.get_object(), x)


@pytest.mark.parametrize("which", ["pair", "pair4d"])
def test_backprojection_is_adjoint(which, request):
    st, o = request.getfixturevalue(which)
    st.build_amatrix()
    rng = np.random.default_rng(7)
    x = rng.random(st.shape_x).astype(np.float32)
    v = rng.standard_normal(st.shape_e).astype(np.float32)
    ax = st.forward_project(x).astype(np.float64)
    atv = st.backproject(v).astype(np.float64)
    lhs = float(np.sum(ax * v))
    rhs = float(np.sum(x.astype(np.float64) * atv))
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), float(np.sum(np.abs(ax * v))) * 1e-2)


@pytest.mark.parametrize("which", ["pair", "pair4d"])
def test_thetas_match_oracle(which, request):
    st, o = request.getfixturevalue(which)
    rng = np.random.default_rng(8)
    st.build_amatrix()
    e = (rng.standard_normal(st.shape_e) * 0.05).astype(np.float32)
    st.set_error(e)
    g = st.g
    q = np.array([(rng.integers(g.N_time), rng.integers(g.N_z), rng.integers(g.N_y),
                   rng.integers(g.N_x)) for _ in range(64)], np.int32)
    t1, t2 = st.thetas(q)
    o1 = np.zeros(len(q))
    o2 = np.zeros(len(q))
    for i, (t, z, yy, xx) in enumerate(q):
        o1[i], o2[i] = o.thetas(e.astype(np.float64), int(t), int(z), int(yy), int(xx))
    assert rel_err(t2, o2) < 1e-5
    # theta1 is a signed sum: bound its error by the scale of its terms
    assert np.max(np.abs(t1 - o1)) <= 1e-5 * (np.max(np.abs(o1)) + 0.05 * np.max(o2))


def test_substitution_matches_oracle(pair):
    from phasecon_amd.mbir import substitution_batch
    st, o = pair
    rng = np.random.default_rng(9)
    n = 4096
    V = (rng.random(n) * 0.02).astype(np.float32)
    nb = (rng.random((n, 27)) * 0.02).astype(np.float32)
    nb[::5] = V[::5, None]           # all-equal neighbourhoods (delta = 0 branch)
    nbf = (rng.random((n, 27)) > 0.15).astype(np.uint8)
    tn = (rng.random((n, 2)) * 0.02).astype(np.float32)
    tf = (rng.random((n, 2)) > 0.3).astype(np.uint8)
    th2 = (rng.random(n) * 2e4).astype(np.float32)
    th1 = (rng.standard_normal(n) * 20).astype(np.float32)
    out = substitution_batch(st.g, V, th1, th2, nb, nbf, tn, tf)
    ref = np.array([o.substitution(float(V[i]), float(th1[i]), float(th2[i]), nb[i], nbf[i],
                                   tn[i], tf[i]) for i in range(n)])
    err = np.abs(out - ref) / np.maximum(np.abs(ref), np.abs(V) + 1e-4)
    assert err.max() < 1e-5


def test_cost_matches_oracle(pair):
    st, o = pair
    rng = np.random.default_rng(10)
    x = (rng.random(st.shape_x) * 0.02).astype(np.float32)
    st.set_object(x)
    st.build_amatrix()
    st.init_error()
    fresh = st.cost(fresh=True)
    kept = st.cost(fresh=False)
    est = st.cost(estimate=True)
    ox = x.astype(np.float64)
    oc, parts = o.cost(o.init_error(ox), ox)
    assert abs(fresh[0] - oc) <= 1e-9 * abs(oc)
    assert abs(fresh[2] - parts[1]) <= 1e-10 * abs(parts[1])
    assert abs(kept[0] - oc) <= 1e-6 * abs(oc)
    # the stopping estimate: FP32 prior potentials within 1e-6 of the FP64 prior
    assert abs(est[2] - parts[1]) <= 1e-6 * abs(parts[1])
    assert est[1] == kept[1]


@pytest.mark.parametrize("mult", [(1, 1), (2, 2), (4, 4), (8, 8)])
def test_stage_amatrix_and_thetas_match_oracle(mult):
    """A and theta1/theta2 of every multi-resolution stage of a 32-row,
    30-views-per-frame geometry: window widths 3, 5, 9, 17 and 1, 2, 4, 8
    detector rows per voxel."""
    from oracle.oracle import Oracle
    from phasecon_amd.mbir import Stage
    from phasecon_amd.synth import make_scan
    scan = make_scan("cfg4_small", seed=1)
    st = Stage(scan, mult_xy=mult[0], mult_z=mult[1])
    o = Oracle(scan, mult_xy=mult[0], mult_z=mult[1])
    st.build_amatrix()
    first, val = st.get_amatrix()
    ofirst, oval = o.amatrix_all()
    sv = st.frame_views[: st.frame_ptr[-1]]
    assert np.array_equal(first, ofirst[:, sv])
    assert rel_err(val, oval[:, sv]) < 1e-6
    rng = np.random.default_rng(9)
    e = (rng.standard_normal(st.shape_e) * 0.05).astype(np.float32)
    st.set_error(e)
    g = st.g
    q = np.array([(rng.integers(g.N_time), rng.integers(g.N_z), rng.integers(g.N_y),
                   rng.integers(g.N_x)) for _ in range(48)], np.int32)
    t1, t2 = st.thetas(q)
    st.close()
    o1 = np.zeros(len(q))
    o2 = np.zeros(len(q))
    for i, (t, z, yy, xx) in enumerate(q):
        o1[i], o2[i] = o.thetas(e.astype(np.float64), int(t), int(z), int(yy), int(xx))
    print(f"mult {mult}: W {st.W}, theta2 rel {rel_err(t2, o2):.3g}", flush=True)
    assert rel_err(t2, o2) < 1e-5
    assert np.max(np.abs(t1 - o1)) <= 1e-5 * (np.max(np.abs(o1)) + 0.05 * np.max(o2))


@pytest.mark.parametrize("name,n_gpu,tile,alpha,p_t", [
    ("tiny3d", 120, 8, None, None), ("ch4d_small", 120, 8, None, None),
    ("ch4d_small", 120, 4, None, None), ("tall3d", 150, 8, None, None),
    ("ch4d_small", 120, 8, 1.9, None), ("tall3d", 150, 8, 1.9, None),
    ("ch4d_small", 120, 8, 1.97, None), ("ch4d_small", 120, 4, 1.97, None),
    ("tall3d", 150, 8, 1.97, None), ("ch4d_small", 120, 8, 1.97, 1.1),
    ("ch4d_small", 120, 4, 1.8, 1.1)])
def test_icd_converges_to_oracle(name, n_gpu, tile, alpha, p_t, request):
    """tiny3d and ch4d_small run the sweep's pipelined theta path, tall3d (16
    detector rows, 24 views) the non-pipelined one; alpha None is the
    reference's over-relaxation factor 1.5, 1.97 the one the benchmark uses
    at full resolution and 1.8 on its coarse stages (tile 4); p_t 1.1 is the
    4D benchmark's temporal qGGMRF p (the reference has one p, 1.2)."""
    from oracle.oracle import Oracle
    from phasecon_amd.mbir import Stage
    from phasecon_amd.synth import make_scan
    scan = (request.getfixturevalue("tiny_scan") if name == "tiny3d" else
            request.getfixturevalue("small4d_scan") if name == "ch4d_small" else
            make_scan(name, seed=0))
    o = Oracle(scan, p_t=p_t)
    ox, oe, ocost = o.run_icd(300)
    cstar = ocost[-1]
    assert abs(ocost[-1] - ocost[-2]) < 1e-9 * cstar  # oracle converged
    st = Stage(scan, tile=tile, alpha=alpha, p_t=p_t)
    st.build_amatrix()
    st.init_error()
    costs = [st.cost()[0]]
    for it in range(1, n_gpu + 1):
        st.icd(1, it, seed=11)
        costs.append(st.cost()[0])
    costs = np.array(costs)
    final = st.cost(fresh=True)[0]
    x = st.get_object().astype(np.float64)
    print(f"{name}: C*={cstar:.9g} gpu={final:.9g} rel={(final - cstar) / cstar:.3g}", flush=True)
    assert np.all(np.diff(costs) <= 1e-6 * np.abs(costs[1:]))  # monotone up to FP32 noise
    assert (final - cstar) / cstar < 1e-5
    rmse = np.sqrt(np.mean((x - ox) ** 2)) / np.sqrt(np.mean(ox ** 2))
    assert rmse < 5e-3
    st.close()


@pytest.mark.parametrize("tile,name", [(4, "ch4d_small"), (8, "ch4d_small"),
                                       (16, "ch4d_small"), (8, "tall3d")])
def test_sweep_keeps_error_sinogram(tile, name, small4d_scan):
    """After sweeps the maintained e must still be y - A x: every voxel
    change reaches the shared error sinogram exactly once (pipelined theta
    path on ch4d_small, non-pipelined on tall3d)."""
    from phasecon_amd.mbir import Stage
    from phasecon_amd.synth import make_scan
    scan = small4d_scan if name == "ch4d_small" else make_scan(name, seed=0)
    st = Stage(scan, tile=tile)
    st.build_amatrix()
    st.init_error()
    st.icd(6, 1, seed=5)
    kept = st.get_error()
    st.init_error()
    exact = st.get_error()
    scale = np.abs(scan["projections"]).max()
    err = np.abs(kept - exact).max()
    print(f"tile={tile}: max |e_kept - e| = {err:.3g} (scale {scale:.3g})", flush=True)
    assert err <= 2e-5 * scale
    st.close()


def test_reconstruct_dropin_files(tiny_scan, tmp_path):
    from phasecon_amd.mbir import reconstruct
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        obj, equits, cost = reconstruct(tiny_scan, convg_thresh=1.0, log_path=str(tmp_path / "debug.log"))
    finally:
        os.chdir(cwd)
    s = tiny_scan
    assert obj.shape == (1, s["proj_rows"] + 2, s["proj_cols"], s["proj_cols"])
    assert np.all(obj[:, 0] == 0) and np.all(obj[:, -1] == 0)
    f0 = np.fromfile(tmp_path / "object_time_0.bin", np.float32)
    assert np.array_equal(f0, obj[0, 1:-1].ravel())
    costs = np.fromfile(tmp_path / "cost.bin", np.float64)
    assert costs.size >= 2 and np.isfinite(costs).all()
    assert np.fromfile(tmp_path / "status.bin", np.int32).tolist() == [0]
    assert np.fromfile(tmp_path / "variance_estimate.bin", np.float64).tolist() == [1.0]
    assert equits >= 2


def _up_axis(a, axis):
    """2x along axis by the reference's rule for the next stage's initial
    object (XT_ICD_update.c:358-390): ends copied, interior samples 3/4 and
    1/4 of the two nearest coarse samples."""
    a = np.moveaxis(a.astype(np.float64), axis, -1)
    n = a.shape[-1]
    out = np.empty(a.shape[:-1] + (2 * n,))
    out[..., 0], out[..., -1] = a[..., 0], a[..., -1]
    out[..., 1] = (3 * a[..., 0] + a[..., 1]) / 4
    out[..., -2] = (a[..., -2] + 3 * a[..., -1]) / 4
    out[..., 2:-2:2] = (a[..., :-2] + 3 * a[..., 1:-1]) / 4
    out[..., 3:-2:2] = (3 * a[..., 1:-1] + a[..., 2:]) / 4
    return np.moveaxis(out, -1, axis)


@pytest.mark.parametrize("mz", [2, 1])
def test_upsample_object(mz, small4d_scan):
    """mbir_hip_upsample_object: a constant stays constant; a linear ramp is
    reproduced exactly at interior samples (the 3/4, 1/4 weights are linear
    interpolation at the fine sample centres); random data equals the rule
    above within FP32 rounding; masked zeroes exactly the voxels outside the
    fine stage's update mask and leaves the others."""
    from phasecon_amd.mbir import Stage
    coarse = Stage(small4d_scan, mult_xy=2, mult_z=mz, tile=4)
    fine = Stage(small4d_scan, tile=4)
    T, Z, Y, X = coarse.shape_x
    print(f"coarse {coarse.shape_x} -> fine {fine.shape_x}", flush=True)
    coarse.set_object(np.full(coarse.shape_x, 0.37, np.float32))
    fine.upsample_from(coarse, masked=False)
    assert np.allclose(fine.get_object(), 0.37, rtol=1e-7)
    ramp = np.broadcast_to(np.arange(X, dtype=np.float32) * 0.25, coarse.shape_x).copy()
    coarse.set_object(ramp)
    fine.upsample_from(coarse, masked=False)
    got = fine.get_object()
    xf = (np.arange(2 * X) + 0.5) / 2 - 0.5  # fine sample centres in coarse units
    assert np.allclose(got[..., 1:-1], (xf * 0.25)[1:-1], atol=1e-6)
    r = np.random.default_rng(11)
    xc = r.random(coarse.shape_x).astype(np.float32)
    coarse.set_object(xc)
    want = _up_axis(_up_axis(xc, 3), 2)
    if mz == 2:
        want = _up_axis(want, 1)
    fine.upsample_from(coarse, masked=False)
    got = fine.get_object()
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) < 1e-6
    fine.upsample_from(coarse, masked=True)
    gm = fine.get_object()
    m = fine.mask.astype(bool)
    assert np.all(gm[..., ~m] == 0)
    assert np.array_equal(gm[..., m], got[..., m])
    coarse.close()
    fine.close()


def test_tiles_below_four_are_refused(small4d_scan):
    """2x2-column tiles make the sweep diverge on the GPU (a defect not yet
    found; the tiled schedule converges at that size in the CPU simulation),
    so the option refuses them with a message instead of running."""
    from phasecon_amd.mbir import Stage
    with pytest.raises(RuntimeError, match=r"tile must be in \[4, 1024\]"):
        Stage(small4d_scan, tile=2)
