"""The sweep's tile autotuner (phasecon_amd/autotune.py): it times every
candidate tile that fits, keeps the fastest, and leaves the stage as it
found it (object unchanged, error sinogram y - A x)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tune_tile_picks_the_fastest_and_restores_the_stage(small4d_scan):
    from phasecon_amd.autotune import tune_tile
    from phasecon_amd.mbir import Stage
    st = Stage(small4d_scan, tile=8, alpha=1.8)
    rng = np.random.default_rng(2)
    x0 = (rng.random(st.shape_x) * 0.02).astype(np.float32)
    st.set_object(x0)
    st.build_amatrix()
    st.init_error()
    e0 = st.get_error()
    print("tuning tiles 4, 8, 16, 32 on ch4d_small", flush=True)
    best, tms = tune_tile(st, candidates=(4, 8, 16, 32))
    assert set(tms) == {4, 8, 16, 32}
    assert all(v > 0 for v in tms.values())
    assert best == min(tms, key=tms.get)
    assert np.array_equal(st.get_object(), x0)
    assert np.max(np.abs(st.get_error() - e0)) <= 1e-6 * np.abs(e0).max()
    # the chosen tile is the stage's tile: equits run and keep e = y - A x
    st.icd(2, 1, seed=3)
    kept = st.get_error()
    st.init_error()
    assert np.max(np.abs(kept - st.get_error())) <= 2e-5 * np.abs(small4d_scan["projections"]).max()
    st.close()
