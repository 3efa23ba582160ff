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


@pytest.mark.parametrize("fft_reg", [1, 0])
@pytest.mark.parametrize("views,rows,cols,dr,dc,windowed", [
    (3, 16, 32, 1.0, 1.0, False), (2, 64, 64, 0.7, 1.3, True), (2, 256, 256, 1.0, 1.0, True),
    (1, 512, 512, 1.0, 1.0, True), (2, 2, 1024, 1.0, 0.5, False), (2, 1024, 8, 1.0, 1.0, True),
    (1, 128, 4, 1.0, 1.0, False), (2, 32, 128, 0.5, 1.0, True)])
def test_fresnel_matches_oracle(views, rows, cols, dr, dc, windowed, fft_reg):
    """Both FFT paths: the register-resident Stockham passes (radix 8, then
    2 or 4; every length from 8 to 1024 along rows or columns among the
    shapes) and the in-LDS radix-2 passes."""
    from phasecon_amd.pcmbir import PhaseRetrieval
    dist = 1e5
    win = P.window(rows, cols, dr, dc, 2.0 / np.sqrt(LAM * dist)) if windowed else None
    ctx = PhaseRetrieval(views, rows, cols, dr, dc, LAM, dist, win)
    ctx.option("fft_reg", fft_reg)
    x = rnd((views, rows, cols), rows + cols)
    for herm in (False, True):
        got = ctx.fresnel(x, herm)
        want = P.fresnel(x, dr, dc, LAM, dist, win=win, herm=herm)
        err = rel_rms(got, want)
        print(f"{views}x{rows}x{cols} herm={herm} fft_reg={fft_reg}: rel rms {err:.3g}",
              flush=True)
        assert err < 5e-6
    ctx.close()


@pytest.mark.parametrize("views,rows,cols", [(2, 512, 64), (2, 1024, 32), (3, 8, 128)])
def test_fresnel_column_workgroup_widths(views, rows, cols):
    """The register column pass at every workgroup width the option allows
    (64 to 1024 threads: 1 to 64 columns per workgroup, at most 64 KiB of
    LDS) gives the oracle's transform; widths that are not a power of two in
    [64, 1024] are refused."""
    from phasecon_amd.pcmbir import PhaseRetrieval
    dist = 1e5
    win = P.window(rows, cols, 1.0, 1.0, 2.0 / np.sqrt(LAM * dist))
    ctx = PhaseRetrieval(views, rows, cols, 1.0, 1.0, LAM, dist, win)
    x = rnd((views, rows, cols), 7 * rows + cols)
    want = P.fresnel(x, 1.0, 1.0, LAM, dist, win=win)
    for threads in (64, 128, 256, 512, 1024):
        ctx.option("fft_col_threads", threads)
        err = rel_rms(ctx.fresnel(x), want)
        print(f"{views}x{rows}x{cols} col threads {threads}: rel rms {err:.3g}", flush=True)
        assert err < 5e-6
    for bad in (0, 32, 96, 2048, 512.5):
        with pytest.raises(RuntimeError, match="fft_col_threads"):
            ctx.option("fft_col_threads", bad)
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
        assert np.allclose(g["costs"][k], o["costs"][1:], rtol=1e-3)
        assert rel_rms(g["zr"][k] + 1j * g["zi"][k], o["zr"] + 1j * o["zi"]) < 1e-3
        assert np.allclose(g["primal"][k], o["primal"], rtol=2e-2)
        assert np.allclose(g["dual"][k], o["dual"], rtol=2e-2)
    ctx.close()


def test_omega_from_w_is_the_phase_projection():
    """set_problem with omega=None: Omega = the phase projection of w, as
    initErrorSinogam sets it (PCMBIR_3D/XT_ICD_update.c:788-789)."""
    from phasecon_amd.pcmbir import PhaseRetrieval
    pps, win, y, D, Lam, v, w, om = phase_problems(2, 16, 32, 60, True)
    ctx = PhaseRetrieval(2, 16, 32, 1.0, 1.0, LAM, 1e5, win)
    ctx.set_problem(y, D, Lam, v, w, None, 0.7)
    _, gom = ctx.state()
    for k in range(2):
        err = rel_rms(gom[k], om[k])
        print(f"view {k}: Omega rel rms {err:.3g}", flush=True)
        assert err < 1e-5
    ctx.close()


@pytest.mark.parametrize("views,rows,cols,dr,dc", [(3, 16, 32, 1.0, 1.0), (2, 256, 64, 0.7, 1.3)])
def test_paganin_matches_oracle(views, rows, cols, dr, dc):
    """FP32 transforms against the FP64 oracle: projections within 1e-5
    relative RMS (thicknesses of order 1)."""
    from phasecon_amd.pcmbir import PhaseRetrieval
    r = np.random.default_rng(rows + cols)
    meas = (0.5 + 0.4 * r.random((views, rows, cols))).astype(complex)
    D = (0.8 + 0.2 * r.random((views, rows, cols))) * np.exp(0.3j * r.random((views, rows, cols)))
    k = 2 * np.pi / LAM
    ctx = PhaseRetrieval(views, rows, cols, dr, dc, LAM, 1e5, None)
    zr, zi = ctx.paganin(meas, D, 50.0, k)
    for v in range(views):
        orr, oii = P.paganin(meas[v], D[v], dr, dc, LAM, 1e5, 50.0)
        er, ei = rel_rms(zr[v], orr), rel_rms(zi[v], oii)
        print(f"view {v}: z_real rel rms {er:.3g}, z_imag {ei:.3g}", flush=True)
        assert er < 1e-5 and ei < 1e-5
    ctx.close()


# ------------------------------------------------ the drop-in (include/pcmbir4d.h)
DROP = dict(vox_wid=1.0, rot_center=8.0, mag_sig_s=2e-2, mag_sig_t=1.0, mag_c_s=1e-2,
            mag_c_t=1e-2, phase_sig_s=2e-2, phase_sig_t=1.0, phase_c_s=1e-2, phase_c_t=1e-2,
            convg_thresh=1.0, obj2det_dist=1e5, light_energy=17.0, pag_regparam=50.0)


def dropin_scan(views, rows, cols, seed):
    """Measurements [view][t][r] and brights [t][r] of a weak phase object."""
    r = np.random.default_rng(seed)
    br = 2500.0 * (0.9 + 0.2 * r.random((rows, cols)))
    trans = np.exp(-0.05 - 0.05 * r.random((views, rows, cols)))
    return br[None] * trans, br


def rd(name, shape):
    return np.fromfile(name + ".bin", np.float64).reshape(shape)


def test_dropin_paganin_files_match_oracle(tmp_path, monkeypatch):
    """recon_type 1: pag_*_ret.bin hold the oracle's Paganin retrieval of
    the normalised amplitudes sqrt(meas) / sqrt(mean bright) with D =
    sqrt(bright) / sqrt(mean bright) (PCMBIR_3D/XT_Init.c:362-389), in the
    sinogram layout [view][r][t]; the auxiliary files have the reference's
    shapes."""
    from phasecon_amd import pcmbir
    monkeypatch.chdir(tmp_path)
    views, rows, cols = 3, 8, 16
    meas, br = dropin_scan(views, rows, cols, 70)
    rc = pcmbir.phcontomo_reconstruct(meas, br, np.zeros(views), np.arange(views * 1.0),
                                      [0.0, float(views)], **DROP, recon_type=1,
                                      log_path=str(tmp_path / "log"))
    assert rc == 0, (tmp_path / "log").read_text()
    bavg = np.sqrt(br.mean())
    y = (np.sqrt(meas) / bavg).transpose(0, 2, 1).astype(np.float32).astype(complex)
    D = np.broadcast_to((np.sqrt(br) / bavg).T.astype(np.float32), y.shape).astype(complex)
    zr, zi = rd("pag_mag_ret", (views, cols, rows)), rd("pag_phase_ret", (views, cols, rows))
    for v in range(views):
        orr, oii = P.paganin(y[v], D[v], 1.0, 1.0, LAM, 1e5, 50.0)
        assert rel_rms(zr[v], orr) < 1e-5 and rel_rms(zi[v], oii) < 1e-5
    aux = rd("mag_tomo_aux", (views, cols, rows, 4))
    assert np.array_equal(aux[..., 1], zr) and not aux[..., [0, 2, 3]].any()
    assert rd("mag_object_time_0", (rows, cols, cols)).shape == (rows, cols, cols)
    assert np.fromfile("status.bin", np.int32).tolist() == [0]
    assert np.fromfile("pret_primal_proj_2.bin", np.float64).tolist() == [0.0]


def test_dropin_phase_retrieval_files_match_the_c_abi(tmp_path, monkeypatch):
    """recon_type 2 after recon_type 1: the files equal an independent run of
    the C-ABI from the same start -- z = the Paganin estimate, w = exp(-z),
    Omega its phase projection, nu = 0.5, 100 ADMM / 100 SD iterations at the
    reference's thresholds (PCMBIR_3D/XT_Init.c:494-521) -- and the TomoAux
    slots 0, 2, 3 follow initErrorSinogam (XT_ICD_update.c:741-776)."""
    from phasecon_amd import pcmbir
    from phasecon_amd.pcmbir import PhaseRetrieval
    monkeypatch.chdir(tmp_path)
    views, rows, cols = 2, 8, 16
    meas, br = dropin_scan(views, rows, cols, 80)
    args = (meas, br, np.zeros(views), np.arange(views * 1.0), [0.0, float(views)])
    assert pcmbir.phcontomo_reconstruct(*args, **DROP, recon_type=1) == 0
    zp_r = rd("pag_mag_ret", (views, cols, rows))
    zp_i = rd("pag_phase_ret", (views, cols, rows))
    print("drop-in recon_type 2", flush=True)
    assert pcmbir.phcontomo_reconstruct(*args, **DROP, recon_type=2,
                                        log_path=str(tmp_path / "log2")) == 0
    # the drop-in's inputs: float32 arrays, normalised in FP64
    m32, b32 = meas.astype(np.float32).astype(np.float64), br.astype(np.float32).astype(np.float64)
    bavg = np.sqrt(b32.mean())
    y = (np.sqrt(m32) / bavg).transpose(0, 2, 1)
    D = np.broadcast_to((np.sqrt(b32) / bavg).T, y.shape)
    win = P.window(cols, rows, 1.0, 1.0, 2.0 / np.sqrt(LAM * 1e5))
    ctx = PhaseRetrieval(views, cols, rows, 1.0, 1.0, LAM, 1e5, win)
    e = np.exp(-zp_r)
    ctx.set_problem(y, D, 1.0, np.zeros(y.shape), e * np.cos(-zp_i) + 1j * e * np.sin(-zp_i),
                    None, 0.5)
    # the drop-in's prior parameters arrive as C floats
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    ctx.set_nms_prior(f32(2e-2), f32(1e-2), f32(2e-2), f32(1e-2), 50.0, True, p=1.2, nu=0.5)
    print("direct C-ABI run", flush=True)
    g = ctx.estimate(zp_r, zp_i, 100, 1e-7, 100, 0.0005)
    w, om = ctx.state()
    vd = ctx.dual()
    ctx.close()
    aux_r = rd("mag_tomo_aux", (views, cols, rows, 4))
    aux_i = rd("phase_tomo_aux", (views, cols, rows, 4))
    assert np.allclose(aux_r[..., 1], g["zr"], rtol=0, atol=1e-9)
    assert np.allclose(aux_i[..., 1], g["zi"], rtol=0, atol=1e-9)
    for aux, zp in ((aux_r, zp_r), (aux_i, zp_i)):
        avg = zp.mean()
        up = np.concatenate([zp[:, :1], zp[:, :-1]], axis=1)
        dn = np.concatenate([zp[:, 1:], 2 * zp[:, -1:]], axis=1)
        assert np.allclose(aux[..., 2], 0.01 * avg + up, rtol=1e-12)
        assert np.allclose(aux[..., 3], 0.02 * avg + dn, rtol=1e-12)
        assert np.allclose(aux[..., 0], (zp + aux[..., 2]) / 2, rtol=1e-12)
    for name, want in (("mag_pret_aux", w.real), ("omega_imag", om.imag),
                       ("mag_pret_dual", vd.real), ("phase_pret_dual", vd.imag)):
        got = rd(name, want.shape)
        print(f"{name}: max |diff| {np.max(np.abs(got - want)):.3g}, "
              f"max |value| {np.max(np.abs(want)):.3g}", flush=True)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
    for v in range(views):
        n = g["iters"][v]
        cost = np.fromfile(f"pret_cost_proj_{v}.bin", np.float64)
        assert cost.size == n + 1
        want = np.r_[g["cost_init"][v], g["costs"][v, :n]]
        print(f"view {v}: {n} iterations, pretcost {cost[0]:.8g} -> {cost[-1]:.8g}", flush=True)
        assert np.allclose(cost, want, rtol=1e-9)
        dual = np.fromfile(f"pret_dual_proj_{v}.bin", np.float64)
        assert np.allclose(dual, np.r_[0.0, g["dual"][v, :n]], rtol=1e-9)
    assert np.all(rd("mag_object_time_0", (rows, cols, cols)) == 1e-15)
