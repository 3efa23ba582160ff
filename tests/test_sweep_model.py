"""The over-relaxed parallel sweep's stability on the benchmark's coarsest
stage, in the oracle's model of the GPU schedule (orc_icd_equit_async).

The 4D benchmark scan's 64x64x2 stage (voxels of 4x4 detector channels and
4 rows, tile 4: about 50 concurrent tiles per frame) is the stage that
diverged on the GPU at alpha 1.97 (bench step with seed offset 4, guard
recovery at its equit 40).  The model runs every tile of a launch in
lock-step, one column per step, and gives each column's theta the error
sinogram without the other tiles' updates of the previous step (the GPU
issues a column's theta loads while the previous column is solved).  In
that model the stage diverges at alpha 1.97 and converges monotonically at
1.8 (the benchmark's over-relaxation of this stage); under the cost-rise
guard of the equit loops (mbir_guard_check) the divergence at 1.97 is
caught, the object of the lowest cost restored, alpha lowered to 1.485, and
the stage then converges.  CPU only.
"""
import numpy as np
import pytest

SEED = 761525  # the benchmark's seed offset 4


@pytest.fixture(scope="module")
def coarse_stage():
    from oracle.oracle import Oracle
    from phasecon_amd.synth import make_scan
    return Oracle(make_scan("ch4d", seed=0), mult_xy=4, mult_z=4)


def _equit(o, x, e, it, alpha):
    # tile 4, one column of lag within a frame, each frame its own cache group
    o.icd_equit_async(x, e, it, 4, alpha, SEED + 7919 * it, 0, 1, -1, 8, 1)
    return o.cost(e, x)[0]


def test_model_diverges_at_197_and_converges_at_18(coarse_stage):
    o = coarse_stage
    runs = {}
    for alpha in (1.97, 1.8):
        x = np.zeros(o.shape_x)
        e = o.init_error(x)
        c = [o.cost(e, x)[0]]
        for it in range(1, 13):
            c.append(_equit(o, x, e, it, alpha))
            if not np.isfinite(c[-1]) or c[-1] > 10 * c[0]:
                break
        runs[alpha] = np.array(c)
        print(f"alpha {alpha}: " + " ".join(f"{v:.4g}" for v in c), flush=True)
    c = runs[1.97]
    assert np.any(np.diff(c) > 0) and c[-1] > c[1]          # rises, and ends above equit 1
    c = runs[1.8]
    assert len(c) == 13 and np.all(np.diff(c) < 0)           # monotone over 12 equits


def test_guard_recovers_the_model_divergence(coarse_stage):
    import ctypes as C
    from phasecon_amd import lib
    L = lib.load()
    o = coarse_stage
    x = np.zeros(o.shape_x)
    e = o.init_error(x)
    g = lib.GuardState()
    L.mbir_guard_init(C.byref(g), o.cost(e, x)[0], 1e-6, 3)
    saved = (x.copy(), e.copy())
    alpha, recovered_at, after = 1.97, None, []
    for it in range(1, 21):
        cost = _equit(o, x, e, it, alpha)
        act = L.mbir_guard_check(C.byref(g), cost, 1)   # the model's e is exact FP64
        assert act != lib.GUARD_FAIL
        if act == lib.GUARD_RECOVER:
            assert recovered_at is None, "one recovery suffices"
            recovered_at = it
            x[...], e[...] = saved
            alpha = L.mbir_guard_alpha(alpha)
            after.append(o.cost(e, x)[0])
            continue
        if act == lib.GUARD_SAVE:
            saved = (x.copy(), e.copy())
        if recovered_at is not None:
            after.append(cost)
    print(f"recovered at equit {recovered_at}, alpha then {alpha:.4g}; costs after: "
          + " ".join(f"{v:.4g}" for v in after), flush=True)
    assert recovered_at is not None and abs(alpha - 1.485) < 1e-12
    assert len(after) >= 8 and np.all(np.diff(after) < 0)


def _cfg4_coarse_scan(frames=8, seed=0):
    """Config 4's geometry at its coarsest stage on the CPU: a 512-channel,
    32-row detector, 28-32 views per frame (K = 8 interlaced), 8 frames, the
    workload's drifting-ellipsoid phantom at the 64x64x4 stage's resolution
    (voxels of 8 x 8 channels and 8 rows) projected by the oracle, Poisson
    counts at I0 = 5000."""
    from oracle.oracle import Oracle
    from phasecon_amd.mbir import interlaced_views
    from phasecon_amd.synth import drifting_ellipsoids, uneven_frames
    views, cols, rows = 30 * frames, 512, 32
    angles, times = interlaced_views(8, views, views)
    per = uneven_frames(views, frames, 28, 32, seed + 77)
    first = np.concatenate([[0], np.cumsum(per)[:-1]])
    recon_times = np.concatenate([times[first], [float(views)]]).astype(np.float32)
    recon_times[0] = 0.0
    scan = dict(name="cfg4_coarse", projections=np.zeros((views, rows, cols), np.float32),
                weights=np.ones((views, rows, cols), np.float32), angles=angles,
                proj_times=times, recon_times=recon_times, proj_rows=rows, proj_cols=cols,
                proj_num=views, recon_num=frames, vox_wid=1.0, rot_center=cols / 2.0,
                sig_s=0.01, sig_t=0.001, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.1, seed=seed)
    o = Oracle(scan, mult_xy=8, mult_z=8)
    line = o.forward_project(drifting_ellipsoids(64, 4, frames, seed).astype(np.float64))
    line = np.ascontiguousarray(line.transpose(0, 2, 1))
    counts = np.maximum(np.random.default_rng(seed + 1000).poisson(5e3 * np.exp(-line)), 1.0)
    scan["projections"] = np.log(5e3 / counts).astype(np.float32)
    scan["weights"] = counts.astype(np.float32)
    return scan


def test_config4_coarse_stage_tile4_diverges_where_tile8_converges():
    """Config 4's 64x64x4 stage (4 of its frames), where the GPU sweep
    diverges with tile 4 at alpha >= 1.5 and converges with tile 8 at 1.9.
    In the model (frames on their own cache groups, tiles seeing each
    other's updates 4 columns late), tile 4 at alpha 1.5 (64 concurrent
    tiles per frame, each column 17 channels x 32 rows of ~30 views)
    diverges within 8 equits, while tile 8 at alpha 1.9 (16 tiles per
    frame) lowers the cost at each of 6 equits; with
    1 or 2 columns of lag tile 4 converges too, so on this geometry the GPU's
    effective staleness exceeds the one column that reproduces the 4D
    benchmark's coarsest stage above."""
    from oracle.oracle import Oracle
    scan = _cfg4_coarse_scan(frames=4)
    runs = {}
    for T, alpha, n in ((4, 1.5, 8), (8, 1.9, 6)):
        o = Oracle(scan, mult_xy=8, mult_z=8)
        x = np.zeros(o.shape_x)
        e = o.init_error(x)
        c = [o.cost(e, x)[0]]
        for it in range(1, n + 1):
            o.icd_equit_async(x, e, it, T, alpha, 761521 + 7919 * it, 0, 4, -1, 8, 1)
            c.append(o.cost(e, x)[0])
            if not np.isfinite(c[-1]) or c[-1] > 10 * c[0]:
                break
        runs[T] = np.array(c)
        print(f"tile {T}, alpha {alpha}: " + " ".join(f"{v:.4g}" for v in c), flush=True)
    c4, c8 = runs[4], runs[8]
    assert np.any(np.diff(c4) > 0) and c4[-1] > c4[0]
    assert len(c8) == 7 and np.all(np.diff(c8) < 0)


def test_two_by_two_tiles_diverge_with_staleness_where_tile4_converges():
    """Tiles of 2x2 columns, which the GPU sweep diverges with (and the tile
    option therefore refuses), on ch4d_small: every tile of a launch in
    lock-step, tiles seeing each other's updates one column late.  Tile 2
    at alpha 1.5 (256 concurrent tiles of 4 columns per frame) diverges
    within 10 equits while tile 4 at the same alpha and lag lowers the cost
    at every equit; without the lag tile 2 converges too (the synchronous
    tiled schedule), so the divergence is the asynchronous staleness of that
    many concurrent columns, not the tile size itself."""
    from oracle.oracle import Oracle
    from phasecon_amd.synth import make_scan
    o = Oracle(make_scan("ch4d_small", seed=0))
    runs = {}
    for T, lag in ((2, 1), (4, 1), (2, 0)):
        x = np.zeros(o.shape_x)
        e = o.init_error(x)
        c = [o.cost(e, x)[0]]
        for it in range(1, 11):
            o.icd_equit_async(x, e, it, T, 1.5, 1000 + it, 0, lag, lag, 1, 0)
            c.append(o.cost(e, x)[0])
            if not np.isfinite(c[-1]) or c[-1] > 10 * c[0]:
                break
        runs[(T, lag)] = np.array(c)
        print(f"tile {T}, lag {lag}: " + " ".join(f"{v:.4g}" for v in c), flush=True)
    c2, c4, c2s = runs[(2, 1)], runs[(4, 1)], runs[(2, 0)]
    assert np.any(np.diff(c2) > 0) and c2[-1] > c2[1]
    assert len(c4) == 11 and np.all(np.diff(c4) < 0)
    assert len(c2s) == 11 and np.all(np.diff(c2s) < 0)
