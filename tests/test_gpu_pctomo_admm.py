"""The PCMBIR ADMM head loop with the tomographic step on the GPU
(mbir_hip_pc_admm_*, pcmbir_hip_set_tomo_target, phasecon_amd.pctomo)
against the oracle (oracle/pctomo_oracle.py, pinned in
test_pctomo_admm_oracle.py).

* ADMM state: init (both start modes), the proximal point and the dual
  update against the oracle's arithmetic from the same objects and
  estimates: FP32 state within 1e-6 of the largest |z|, residuals within
  1e-5 relative.
* The coupled simplex pass and the coupled retrieval (ENABLE_TOMO_RECONS
  Cost_NMS and pretcost) against the oracle's four-colour pass, both FP64:
  identical search paths except where a libm-vs-device ulp flips a
  comparison (at most 3 % of pixels); three ADMM iterations' pretcosts within
  1e-3 and z within 1e-3 relative RMS (FP32 transforms).
* Two head iterations of the driver (10 ICD equits at a tile of half the
  grid, whose order the oracle reproduces, then 2 retrieval iterations)
  against the oracle's head loop: residuals within 2e-3 relative, objects
  within 1e-4 relative RMS (measured: 3e-4 and 1e-5), and e = Z - V - k A x
  on the GPU after each iteration within 1e-5 of the largest |z|.
"""
import numpy as np
import pytest

from oracle import pcmbir_oracle as PO
from oracle import pctomo_oracle as PT

pytestmark = pytest.mark.gpu

NMS = dict(phase_sig=2e-2, phase_c=1e-2, mag_sig=2e-2, mag_c=1e-2, delta_over_beta=50.0,
           decorrelate=True)


def _stage(name="ch4d_small", seed=0, wavenumber=30.0):
    from oracle.oracle import Oracle, pc_prior
    from phasecon_amd.mbir import TomoStage, tomo_prior
    from phasecon_amd.synth import make_scan
    scan = make_scan(name, seed=seed)
    o = Oracle(scan)
    args = ((0.02, 1e-3, 0.01, 0.01), (0.005, 1e-3, 0.01, 0.01), 1, o.g.delta_recon, wavenumber)
    Po = pc_prior(*args, delta_over_beta=50.0, mu=0.5)
    Pg = tomo_prior(*args, delta_over_beta=50.0, mu=0.5)
    return o, Po, TomoStage(scan, Pg, tile=8)


@pytest.mark.parametrize("from_object", [False, True])
def test_admm_state_matches_oracle(from_object):
    o, Po, st = _stage()
    try:
        r = np.random.default_rng(3)
        xd = (0.004 + 0.002 * r.random(o.shape_x)).astype(np.float32)
        xb = (xd / 50).astype(np.float32)
        k = Po.wavenumber
        kd = k * o.forward_project(xd.astype(np.float64))
        kb = k * o.forward_project(xb.astype(np.float64))
        zd = kd * (1 + 0.1 * r.normal(size=kd.shape))
        zb = kb * (1 + 0.1 * r.normal(size=kb.shape))
        st.set_objects(xd, xb)
        st.admm_init(zd, zb, from_object)
        sd, sb = PT.admm_init(zd, kd, from_object), PT.admm_init(zb, kb, from_object)
        scale = max(np.abs(zd).max(), np.abs(zb).max())

        def check(where):
            got = [st.admm_state(w) for w in range(3)] + [st.get_errors()]
            for i, key in enumerate(("Z", "V", "P", "e")):
                for obj, s in ((0, sd), (1, sb)):
                    err = np.abs(got[i][obj] - s[key]).max() / scale
                    assert err < 1e-6, (where, key, obj, err)

        check("init")
        ad, ab = st.admm_target()
        assert np.abs(ad - PT.admm_target(sd)).max() < 1e-6 * scale
        assert np.abs(ab - PT.admm_target(sb)).max() < 1e-6 * scale
        for it in range(2):
            znd = kd * (1 + 0.05 * r.normal(size=kd.shape))
            znb = kb * (1 + 0.05 * r.normal(size=kb.shape))
            res = st.admm_update(znd, znb)
            pb, db = PT.admm_update(sb, znb)
            pd, dd = PT.admm_update(sd, znd)
            print(f"update {it}: gpu {res} oracle {[pb, db, pd, dd]}", flush=True)
            assert np.allclose(res, [pb, db, pd, dd], rtol=1e-5)
            check(f"update {it}")
    finally:
        st.close()


def test_coupled_nms_pass_matches_oracle():
    from phasecon_amd.pcmbir import PhaseRetrieval
    views, rows, cols = 2, 8, 16
    r = np.random.default_rng(41)
    zr = r.random((views, rows, cols)) * 0.05 + 0.02
    zi = r.random((views, rows, cols)) * 0.5 + 0.2
    b = PO.transmittance(zr * 1.2, zi * 0.9).astype(np.complex64).astype(np.complex128)
    a = zr * 1.1 + 0.01 + 1j * (zi * 0.95)
    mu = 0.5
    ctx = PhaseRetrieval(views, rows, cols, 1.0, 1.0, PO.wavelength(17.0), 1e5, None)
    try:
        ctx.set_nms_prior(**NMS, nu=0.5)
        ctx.set_tomo_target(mu, a)
        gr, gi, gits = ctx.nms_pass(zr, zi, b)
    finally:
        ctx.close()
    pr = PO.NmsPrior(2e-2, 1e-2, 2e-2, 1e-2, 50.0, True)
    oits = 0
    for k in range(views):
        orr, oii, it = PO.nms_pass(zr[k], zi[k], b[k], 0.5, pr, order="colour", tomo=(mu, a[k]))
        oits += it
        bad = (np.abs(gr[k] - orr) > 1e-8) | (np.abs(gi[k] - oii) > 1e-8)
        print(f"view {k}: {bad.sum()} of {bad.size} pixels differ", flush=True)
        assert bad.mean() <= 0.03
    assert abs(gits - oits) <= 0.03 * oits


def test_coupled_estimate_matches_oracle():
    from phasecon_amd.pcmbir import PhaseRetrieval
    views, rows, cols = 2, 16, 16
    r = np.random.default_rng(52)
    lam = PO.wavelength(17.0)
    geo = (1.0, 1.0, lam, 1e5, None)
    ztr = r.random((views, rows, cols)) * 0.05 + 0.02
    zti = r.random((views, rows, cols)) * 0.5 + 0.2
    y = np.abs(PO.fresnel(PO.transmittance(ztr, zti), *geo)) + 0j
    y = y.astype(np.complex64).astype(np.complex128)
    D = np.ones_like(y)
    Lam = (1.0 + r.random(y.shape)).astype(np.float32).astype(np.float64)
    zr, zi = ztr * 0.9, zti * 1.1
    w = PO.transmittance(zr, zi).astype(np.complex64).astype(np.complex128)
    a = ztr + 1j * zti
    mu, nu = 0.5, 0.5
    ctx = PhaseRetrieval(views, rows, cols, *geo)
    try:
        ctx.set_problem(y, D, Lam, np.zeros_like(w), w, None, nu)
        _, om = ctx.state()
        ctx.set_nms_prior(**NMS, nu=nu)
        ctx.set_tomo_target(mu, a)
        g = ctx.estimate(zr, zi, 3, 0.0, 3, 0.0)
    finally:
        ctx.close()
    pr = PO.NmsPrior(2e-2, 1e-2, 2e-2, 1e-2, 50.0, True)
    for k in range(views):
        o = PO.estimate_complex_projection(y[k], om[k], D[k], zr[k], zi[k], Lam[k], w[k], pr, geo,
                                           nu, 3, 0.0, 3, 0.0, order="colour", tomo=(mu, a[k]))
        zrms = float(np.sqrt(np.mean(np.abs((g["zr"][k] - o["zr"]) + 1j * (g["zi"][k] - o["zi"]))
                                     ** 2) / np.mean(np.abs(o["zr"] + 1j * o["zi"]) ** 2)))
        print(f"view {k}: cost {g['cost_init'][k]:.8g}->{g['cost_final'][k]:.8g} vs "
              f"{o['costs'][0]:.8g}->{o['costs'][-1]:.8g}; z rms {zrms:.3g}", flush=True)
        assert g["cost_init"][k] == pytest.approx(o["costs"][0], rel=1e-5)
        assert np.allclose(g["costs"][k], o["costs"][1:], rtol=1e-3)
        assert zrms < 1e-3


def test_head_loop_matches_oracle():
    from phasecon_amd.mbir import tomo_prior
    from phasecon_amd.pctomo import TomoADMM
    from test_pctomo_admm_oracle import _prior, tomo_problem
    p = tomo_problem()
    o, Po = p["o"], p["Pt"]
    T = max(o.g.N_x, o.g.N_y) // 2
    Pg = tomo_prior((0.02, 1e-3, 0.01, 0.01), (0.005, 1e-3, 0.01, 0.01), 1, o.g.delta_recon,
                    Po.wavenumber, delta_over_beta=50.0, mu=Po.mu, alpha=Po.alpha)
    nu, seed = 0.5, 31
    drv = TomoADMM(p["scan"], Pg, p["meas"], p["D"], p["Lam"], p["geo"], nu, NMS, tile=T)
    hl = PT.HeadLoop(o, Po, p["meas"], p["D"], p["Lam"], _prior(), p["geo"], nu,
                     np.zeros(o.shape_x), np.zeros(o.shape_x), p["z0d"], p["z0b"])
    scale = max(np.abs(p["zd"]).max(), np.abs(p["zb"]).max())
    try:
        drv.start(np.zeros(o.shape_x), np.zeros(o.shape_x), p["z0d"], p["z0b"])
        for head in range(2):
            out = drv.head_iteration(10, 2, 0.0, 2, 0.0, seed=seed)
            hl.icd(10, 1, T, [seed + 10 * head + i for i in range(10)])
            zd, zb, _ = hl.retrieval(2, 0.0, 2, 0.0)
            res_o = hl.update(zd, zb)
            xd, xb = drv.objects()
            dx = float(np.sqrt(np.mean((xd - hl.xd) ** 2) / np.mean(hl.xd ** 2)))
            k = Po.wavenumber
            Z, V = drv.st.admm_state(0), drv.st.admm_state(1)
            ed, eb = drv.st.get_errors()
            inv = max(np.abs(ed - (Z[0] - V[0] - k * o.forward_project(xd.astype(np.float64))))
                      .max(), np.abs(eb - (Z[1] - V[1] - k * o.forward_project(
                          xb.astype(np.float64)))).max()) / scale
            print(f"head {head + 1}: gpu residuals {out['residuals']} oracle {res_o}; "
                  f"ICD cost {out['icd_cost']}; delta rms diff {dx:.3g}; e invariant {inv:.3g}",
                  flush=True)
            assert np.allclose(out["residuals"], res_o, rtol=2e-3)
            assert dx < 1e-4
            assert inv < 1e-5
    finally:
        drv.close()
