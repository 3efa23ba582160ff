"""GPU parity of the PCMBIR phase-retrieval kernels (libpcmbir4d.so, through
its C-ABI) against the FP64 numpy oracle (oracle/pcmbir_oracle.py).

Tolerances (FP32 kernels against an FP64 oracle): Fresnel transforms within
5e-6 relative RMS (FP32 radix-2 FFTs of up to 2^20 points); steepest-descent
iterates within 1e-4 relative RMS after 5 iterations; costs within 1e-5.
"""
import numpy as np
import pytest

from oracle import pcmbir_oracle as P

pytestmark = pytest.mark.gpu

LAM = P.wavelength(17.0)


def rnd(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def rel_rms(a, b):
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2) / np.mean(np.abs(b) ** 2)))


@pytest.mark.parametrize("views,rows,cols,dr,dc,windowed", [
    (3, 16, 32, 1.0, 1.0, False), (2, 64, 64, 0.7, 1.3, True), (2, 256, 256, 1.0, 1.0, True),
    (1, 512, 512, 1.0, 1.0, True), (2, 2, 1024, 1.0, 0.5, False)])
def test_fresnel_matches_oracle(views, rows, cols, dr, dc, windowed):
    from phasecon_amd.pcmbir import PhaseRetrieval
    dist = 1e5
    win = P.window(rows, cols, dr, dc, 2.0 / np.sqrt(LAM * dist)) if windowed else None
    ctx = PhaseRetrieval(views, rows, cols, dr, dc, LAM, dist, win)
    x = rnd((views, rows, cols), rows + cols)
    for herm in (False, True):
        got = ctx.fresnel(x, herm)
        want = P.fresnel(x, dr, dc, LAM, dist, win=win, herm=herm)
        err = rel_rms(got, want)
        print(f"{views}x{rows}x{cols} herm={herm}: rel rms {err:.3g}", flush=True)
        assert err < 5e-6
    ctx.close()


def phase_problems(views, rows, cols, seed, windowed):
    r = np.random.default_rng(seed)
    win = P.window(rows, cols, 1.0, 1.0, 2.0 / np.sqrt(LAM * 1e5)) if windowed else None
    y = np.abs(rnd((views, rows, cols), seed)) + 0.5
    D = np.exp(-0.1 * r.random((views, rows, cols)) - 0.3j * r.random((views, rows, cols)))
    Lam = 0.5 + r.random((views, rows, cols))
    v = rnd((views, rows, cols), seed + 1) * 0.1 + 1.0
    w = rnd((views, rows, cols), seed + 2) * 0.2 + 1.0
    pps = [P.PhaseProblem(y[k], D[k], Lam[k], v[k], 0.7, 1.0, 1.0, LAM, 1e5, win)
           for k in range(views)]
    om = np.stack([pp.phase_projection(w[k]) for k, pp in enumerate(pps)])
    return pps, win, y, D, Lam, v, w, om


@pytest.mark.parametrize("windowed", [False, True])
def test_sd_iterations_match_oracle(windowed):
    from phasecon_amd.pcmbir import PhaseRetrieval
    views, rows, cols = 3, 32, 64
    pps, win, y, D, Lam, v, w, om = phase_problems(views, rows, cols, 20, windowed)
    ctx = PhaseRetrieval(views, rows, cols, 1.0, 1.0, LAM, 1e5, win)
    ctx.set_problem(y, D, Lam, v, w, om, 0.7)
    cost0 = ctx.cost()
    for k, pp in enumerate(pps):
        assert cost0[k] == pytest.approx(pp.cost(om[k], w[k]), rel=1e-5)
    iters, upd = ctx.sd_run(5, 0.0)
    gw, gom = ctx.state()
    assert list(iters) == [5] * views
    for k, pp in enumerate(pps):
        ow, oom, n, u = pp.sd_run(om[k], w[k], 5, 0.0)
        assert n == 5
        print(f"view {k}: w rel rms {rel_rms(gw[k], ow):.3g}, Omega {rel_rms(gom[k], oom):.3g}, "
              f"upd {upd[k]:.6g} vs {u:.6g}", flush=True)
        assert rel_rms(gw[k], ow) < 1e-4
        assert rel_rms(gom[k], oom) < 1e-4
        assert upd[k] == pytest.approx(u, rel=1e-3)
    ctx.close()


def test_sd_stopping_rule_per_view():
    """A view stops after the first iteration j > 1 below the threshold
    (XT_CmplxProjEst.c:165-169): with an unreachable-high threshold every view
    runs exactly 3 iterations; with a zero threshold all max_iters."""
    from phasecon_amd.pcmbir import PhaseRetrieval
    pps, win, y, D, Lam, v, w, om = phase_problems(2, 16, 16, 30, False)
    ctx = PhaseRetrieval(2, 16, 16, 1.0, 1.0, LAM, 1e5, win)
    ctx.set_problem(y, D, Lam, v, w, om, 0.7)
    iters, _ = ctx.sd_run(10, 1e30)
    assert list(iters) == [3, 3]
    assert pps[0].sd_run(om[0], w[0], 10, 1e30)[2] == 3
    ctx.close()


PRIOR = dict(phase_sig=2e-2, phase_c=1e-2, mag_sig=2e-2, mag_c=1e-2, delta_over_beta=50.0)


def test_nms_pass_matches_oracle():
    """The GPU's four-colour pass against the oracle's four-colour pass, both
    FP64: identical search paths except where a libm-vs-device-library ulp
    flips a comparison (allowed on at most 3 % of pixels)."""
    from phasecon_amd.pcmbir import PhaseRetrieval
    views, rows, cols = 2, 8, 16
    r = np.random.default_rng(40)
    zr = r.random((views, rows, cols)) * 0.05 + 0.02
    zi = r.random((views, rows, cols)) * 0.5 + 0.2
    b = P.transmittance(zr * 1.2, zi * 0.9).astype(np.complex64).astype(np.complex128)
    ctx = PhaseRetrieval(views, rows, cols, 1.0, 1.0, LAM, 1e5, None)
    ctx.set_nms_prior(**PRIOR, decorrelate=True, nu=0.5)
    gr, gi, gits = ctx.nms_pass(zr, zi, b)
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    oits = 0
    for k in range(views):
        orr, oii, it = P.nms_pass(zr[k], zi[k], b[k], 0.5, pr, order="colour")
        oits += it
        bad = (np.abs(gr[k] - orr) > 1e-8) | (np.abs(gi[k] - oii) > 1e-8)
        print(f"view {k}: {bad.sum()} of {bad.size} pixels differ", flush=True)
        assert bad.mean() <= 0.03
    assert abs(gits - oits) <= 0.03 * oits
    ctx.close()


def test_estimate_matches_oracle():
    """Three ADMM iterations (NMS pass, 3 steepest-descent steps, dual
    update) on two views: pretcosts within 1e-3, z within 1e-3 relative RMS,
    residual histories within 2e-2 (FP32 transforms against FP64)."""
    from phasecon_amd.pcmbir import PhaseRetrieval
    views, rows, cols = 2, 16, 16
    pps, win, y, D, Lam, v, w, om = phase_problems(views, rows, cols, 50, True)
    r = np.random.default_rng(51)
    zr = r.random((views, rows, cols)) * 0.05 + 0.02
    zi = r.random((views, rows, cols)) * 0.5 + 0.2
    y = np.abs(y).astype(np.complex128)
    ctx = PhaseRetrieval(views, rows, cols, 1.0, 1.0, LAM, 1e5, win)
    ctx.set_problem(y, D, Lam, v, w, om, 0.5)
    ctx.set_nms_prior(**PRIOR, decorrelate=True, nu=0.5)
    g = ctx.estimate(zr, zi, 3, 0.0, 3, 0.0)
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    geo = (1.0, 1.0, LAM, 1e5, win)
    for k in range(views):
        o = P.estimate_complex_projection(y[k], om[k], D[k], zr[k], zi[k], Lam[k], w[k], pr, geo,
                                          0.5, 3, 0.0, 3, 0.0, order="colour")
        print(f"view {k}: cost {g['cost_init'][k]:.8g}->{g['cost_final'][k]:.8g} vs "
              f"{o['costs'][0]:.8g}->{o['costs'][-1]:.8g}; z rms "
              f"{rel_rms(g['zr'][k] + 1j * g['zi'][k], o['zr'] + 1j * o['zi']):.3g}", flush=True)
        assert g["iters"][k] == o["iters"] == 3
        assert g["cost_init"][k] == pytest.approx(o["costs"][0], rel=1e-5)
        assert g["cost_final"][k] == pytest.approx(o["costs"][-1], rel=1e-3)
        assert rel_rms(g["zr"][k] + 1j * g["zi"][k], o["zr"] + 1j * o["zi"]) < 1e-3
        assert np.allclose(g["primal"][k], o["primal"], rtol=2e-2)
        assert np.allclose(g["dual"][k], o["dual"], rtol=2e-2)
    ctx.close()
