"""FP16 A with a per-column FP32 scale (the "afp16" option).

Pins: the stored values are bit-exact against a numpy emulation of the
rounding (scale = max |a| of the column in FP32, q = fp16(a / scale) rounded
to nearest even, value = q * scale in FP32); the sweep's error sinogram stays
y - A16 x for the rounded A16 it streams (checked by projecting A16 on the
CPU, FP64); and a run with FP16 A, refreshed every 10 equits with the exact
on-the-fly projector, reaches the FP64 oracle's converged MAP cost within
1e-5 relative and its image within 0.5 % RMSE.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _quantise(val):
    """numpy emulation of the device rounding of A (per column = axis 0)."""
    s = np.abs(val).reshape(val.shape[0], -1).max(axis=1).astype(np.float32)
    s3 = s[:, None, None]
    with np.errstate(invalid="ignore", divide="ignore"):
        q = np.where(s3 > 0, val / s3, np.float32(0)).astype(np.float32).astype(np.float16)
    return (q.astype(np.float32) * s3).astype(np.float32), s


def _project_stored(st, first, val, x):
    """A x in FP64 from the stored column layout: view slot s of column c
    covers bins first[c, s] + k (k < W) of its view, and voxel z the
    detector rows t with t // (N_t / N_z) == z."""
    g = st.g
    ax = np.zeros(st.shape_e, np.float64)  # [N_p][N_r][N_t]
    ncol = g.N_x * g.N_y
    zo = g.N_t // g.N_z
    t_of = np.arange(g.N_t) // zo
    for f in range(g.N_time):
        xc = x[f].reshape(g.N_z, ncol).T.astype(np.float64)[:, t_of]  # [col][N_t]
        for s in range(int(st.frame_ptr[f]), int(st.frame_ptr[f + 1])):
            p = int(st.frame_views[s])
            r0 = first[:, s]
            for k in range(st.W):
                r = r0 + k
                ok = (r0 >= 0) & (r < g.N_r) & (val[:, s, k] != 0)
                np.add.at(ax[p], r[ok], val[ok, s, k, None].astype(np.float64) * xc[ok])
    return ax


@pytest.mark.parametrize("name", ["ch4d_small", "tall3d"])
def test_afp16_values_match_emulated_rounding(name, small4d_scan):
    from phasecon_amd.mbir import Stage
    from phasecon_amd.synth import make_scan
    scan = small4d_scan if name == "ch4d_small" else make_scan(name, seed=0)
    st = Stage(scan)
    st.build_amatrix()
    first, a32 = st.get_amatrix()
    st.option("afp16", 1)
    st.build_amatrix()
    first16, a16 = st.get_amatrix()
    st.close()
    emu, s = _quantise(a32)
    print(f"{name}: {np.mean(a16 != a32):.3f} of the values changed by rounding", flush=True)
    assert np.array_equal(first16, first)
    assert np.array_equal(a16, emu)
    # half an FP16 ulp of the column's largest value at most
    assert np.all(np.abs(a16 - a32) <= np.ldexp(s, -11)[:, None, None])
    assert np.mean(a16 != a32) > 0.5


@pytest.mark.parametrize("name,tile", [("ch4d_small", 8), ("tall3d", 8)])
def test_afp16_sweep_keeps_error_sinogram(name, tile, small4d_scan):
    """After sweeps from x = 0 the maintained e equals y - A16 x, with the
    FP16-rounded A16 the sweep streamed (pipelined theta path on ch4d_small,
    non-pipelined on tall3d)."""
    from phasecon_amd.mbir import Stage
    from phasecon_amd.synth import make_scan
    scan = small4d_scan if name == "ch4d_small" else make_scan(name, seed=0)
    st = Stage(scan, tile=tile, afp16=True)
    st.set_object(np.zeros(st.shape_x, np.float32))
    st.build_amatrix()
    st.init_error()
    st.icd(6, 1, seed=5)
    kept = st.get_error().astype(np.float64)
    x = st.get_object()
    first, a16 = st.get_amatrix()
    expect = st.y.astype(np.float64) - _project_stored(st, first, a16, x)
    st.init_error()  # exact A: differs from y - A16 x by (A - A16) x
    exact = st.get_error().astype(np.float64)
    st.close()
    scale = np.abs(scan["projections"]).max()
    err = np.abs(kept - expect).max()
    gap = np.abs(exact - expect).max()
    print(f"{name}: max |e_kept - (y - A16 x)| = {err:.3g}; |(A - A16) x| up to {gap:.3g} "
          f"(scale {scale:.3g})", flush=True)
    assert err <= 2e-5 * scale


def test_afp16_converges_to_oracle(small4d_scan):
    """The error against the FP64-A oracle run: relative MAP cost (exact A,
    FP64) and image RMSE of a run with FP16 A."""
    from oracle.oracle import Oracle
    from phasecon_amd.mbir import Stage
    o = Oracle(small4d_scan)
    ox, oe, ocost = o.run_icd(300)
    cstar = ocost[-1]
    st = Stage(small4d_scan, tile=8, afp16=True)
    st.build_amatrix()
    st.init_error()
    for it in range(1, 121):
        st.icd(1, it, seed=11)
        if it % 10 == 0:
            st.init_error()
    final = st.cost(fresh=True)[0]
    x = st.get_object().astype(np.float64)
    st.close()
    rel = (final - cstar) / cstar
    rmse = np.sqrt(np.mean((x - ox) ** 2)) / np.sqrt(np.mean(ox ** 2))
    print(f"FP16 A vs FP64-A oracle: C*={cstar:.9g} gpu={final:.9g} rel={rel:.3g} "
          f"rmse={rmse:.3g}", flush=True)
    assert rel < 1e-5
    assert rmse < 5e-3
