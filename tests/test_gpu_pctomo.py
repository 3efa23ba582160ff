"""The PCMBIR tomographic ICD on the GPU (mbir_hip_pc_*, the ENABLE_TOMO_RECONS
path of PCMBIR_3D) against the oracle's FP64 restatement (orc_pc_equit,
pinned in test_pctomo_oracle.py).

With a tile as wide as half the grid each 2x2 colour holds one tile, so a
launch runs one wavefront per frame of one parity; those frames share no
views and are not temporal neighbours (the temporal filter is 0), so the
launch is deterministic and the oracle's serial order is the GPU's.  One
equit from a start n0 oracle equits from zero (rounded to FP32; n0 where the
oracle alone amplifies a 1e-7 input perturbation by at most about 50: 1 on
tiny3d at 10, 48 on ch4d_small at 10, 5 at 20) is compared voxel by voxel:
both objects within 2e-5 of the largest |x| of each, both error sinograms
within 2e-5 of the largest |target|.  e = target - k A x from
mbir_hip_pc_init_error against the oracle's; the FP64 cost against
orc_pc_cost within 1e-6 relative.  60 asynchronous GPU equits at tile 8
lower the cost monotonically, to no more than 1 % above the oracle's own
60-equit cost (tile 16) and within 5 % of it, and the maintained error
sinograms stay within 1e-4 of target - k A x recomputed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 2e-5
NCONV = 60
PHASE = (0.02, 1e-3, 0.01, 0.01)
MAG = (0.005, 1e-3, 0.01, 0.01)


def _problem(name, wavenumber=30.0, dob=50.0, alpha=1.5, seed=0, **over):
    from oracle.oracle import Oracle, pc_prior
    from phasecon_amd.mbir import tomo_prior
    from phasecon_amd.synth import make_scan
    scan = make_scan(name, seed=seed, **over)
    o = Oracle(scan)
    Po = pc_prior(PHASE, MAG, 1, o.g.delta_recon, wavenumber, delta_over_beta=dob, alpha=alpha)
    Pg = tomo_prior(PHASE, MAG, 1, o.g.delta_recon, wavenumber, delta_over_beta=dob, alpha=alpha)
    r = np.random.default_rng(seed + 11)
    xd_true = 0.004 + 0.002 * r.random(o.shape_x)
    xb_true = xd_true / dob
    td = wavenumber * o.forward_project(xd_true) + 1e-3 * r.normal(size=o.y.shape)
    tb = wavenumber * o.forward_project(xb_true) + 1e-4 * r.normal(size=o.y.shape)
    td = td.astype(np.float32).astype(np.float64)
    tb = tb.astype(np.float32).astype(np.float64)
    return scan, o, Po, Pg, td, tb


def _start(o, P, td, tb, n0):
    xd = np.zeros(o.shape_x)
    xb = np.zeros(o.shape_x)
    ed, eb = o.pc_init_error(P, td, tb, xd, xb)
    for it in range(1, n0 + 1):
        o.pc_equit(P, xd, xb, ed, eb, it, 16, 5 + it)
    return xd.astype(np.float32).astype(np.float64), xb.astype(np.float32).astype(np.float64)


# (workload, delta/beta, n0, overrides): the last case has 200 voxels per
# column (a lane runs several voxels of a column, the slab 58 KiB of LDS)
@pytest.mark.parametrize("name,dob,n0,over", [
    ("tiny3d", 50.0, 10, {}), ("tiny3d", None, 10, {}), ("ch4d_small", 50.0, 20, {}),
    ("tiny3d", 50.0, 10, {"rows": 200, "cols": 8})])
def test_tomo_equit_matches_oracle(name, dob, n0, over):
    from phasecon_amd.mbir import TomoStage
    scan, o, Po, Pg, td, tb = _problem(name, dob=dob or 50.0, **over)
    if dob is None:
        from oracle.oracle import pc_prior
        from phasecon_amd.mbir import tomo_prior
        Po = pc_prior(PHASE, MAG, 1, o.g.delta_recon, 30.0, alpha=1.5)
        Pg = tomo_prior(PHASE, MAG, 1, o.g.delta_recon, 30.0, alpha=1.5)
    T = max(o.g.N_x, o.g.N_y) // 2
    print(f"{name} {over}: tile {T}, delta/beta {dob}", flush=True)
    xd0, xb0 = _start(o, Po, td, tb, n0)
    st = TomoStage(scan, Pg, tile=T)
    try:
        st.set_objects(xd0, xb0)
        st.init_errors(td, tb)
        ed_g, eb_g = st.get_errors()
        ed_o, eb_o = o.pc_init_error(Po, td, tb, xd0, xb0)
        scale_e = max(np.abs(td).max(), np.abs(tb).max())
        assert np.abs(ed_g - ed_o).max() < 1e-6 * scale_e
        assert np.abs(eb_g - eb_o).max() < 1e-6 * scale_e
        cg = st.tomo_cost()
        co = o.pc_cost(Po, ed_g.astype(np.float64), eb_g.astype(np.float64), xd0, xb0)
        print(f"cost gpu {cg[0]:.10g} oracle {co[0]:.10g}", flush=True)
        assert cg[0] == pytest.approx(co[0], rel=1e-6)
        seed = 4242
        stats = st.equits(1, n0 + 1, seed)
        xd_g, xb_g = st.get_objects()
        ed_g, eb_g = st.get_errors()
    finally:
        st.close()
    xd, xb = xd0.copy(), xb0.copy()
    ed, eb = o.pc_init_error(Po, td, tb, xd, xb)
    so = o.pc_equit(Po, xd, xb, ed, eb, n0 + 1, T, seed)
    dxd = np.abs(xd_g - xd).max() / np.abs(xd).max()
    dxb = np.abs(xb_g - xb).max() / np.abs(xb).max()
    de = max(np.abs(ed_g - ed).max(), np.abs(eb_g - eb).max()) / scale_e
    moved = np.abs(xd - xd0).max() / np.abs(xd).max()
    print(f"delta {dxd:.3g}, beta {dxb:.3g}, e {de:.3g}; the equit moved delta by {moved:.3g}; "
          f"stats gpu {stats} oracle {so}", flush=True)
    assert dxd < TOL and dxb < TOL and de < TOL
    assert moved > 20 * TOL
    assert stats[0] == pytest.approx(so[0], rel=1e-3)


def test_tomo_icd_converges_with_oracle():
    from phasecon_amd.mbir import TomoStage
    scan, o, Po, Pg, td, tb = _problem("ch4d_small", seed=1)
    st = TomoStage(scan, Pg, tile=8)
    try:
        st.set_objects(np.zeros(o.shape_x), np.zeros(o.shape_x))
        st.init_errors(td, tb)
        costs = [st.tomo_cost()[0]]
        for it in range(1, NCONV + 1):
            st.equits(1, it, 100 + it)
            costs.append(st.tomo_cost()[0])
        ed_k, eb_k = st.get_errors()
        xd_g, xb_g = st.get_objects()
        st.init_errors(td, tb)
        ed_f, eb_f = st.get_errors()
    finally:
        st.close()
    costs = np.array(costs)
    xd = np.zeros(o.shape_x)
    xb = np.zeros(o.shape_x)
    ed, eb = o.pc_init_error(Po, td, tb, xd, xb)
    for it in range(1, NCONV + 1):
        o.pc_equit(Po, xd, xb, ed, eb, it, 16, 100 + it)
    c_o = o.pc_cost(Po, ed, eb, xd, xb)[0]
    print("gpu costs " + " ".join(f"{v:.8g}" for v in costs[[0, 1, 5, 10, 20, 40, 60]]) +
          f"; oracle after {NCONV}: {c_o:.8g}", flush=True)
    assert np.all(np.diff(costs) <= 1e-6 * costs[1:])
    assert costs[-1] < 1e-3 * costs[0]
    assert costs[-1] <= c_o * 1.01
    assert abs(costs[-1] - c_o) / c_o < 0.05
    scale_e = max(np.abs(td).max(), np.abs(tb).max())
    drift = max(np.abs(ed_k - ed_f).max(), np.abs(eb_k - eb_f).max()) / scale_e
    print(f"maintained e vs recomputed: {drift:.3g}", flush=True)
    assert drift < 1e-4
