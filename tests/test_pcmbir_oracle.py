"""Pins the numpy PCMBIR oracle (oracle/pcmbir_oracle.py) by closed forms.

FFTW is absent, so no reference build pins the Fresnel transform; the checks
below fix what the mathematics fixes, each chosen so that a plausible mistake
(a sign of the phase, a wrong frequency index, a missing 1/N, the window on
the adjoint, a dropped term of the gradient) fails one of them.  CPU only.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import pcmbir_oracle as P  # noqa: E402

GEO = dict(d_rows=1.0, d_cols=1.0, lam=P.wavelength(17.0), dist=1e5)


def rnd(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def test_wavelength_uses_reference_constant():
    # Q10: the reference's Planck constant is 10x below CODATA
    assert P.wavelength(17.0) == pytest.approx(4.135668e-18 * 299792458e6 / 17.0 / 10, rel=1e-15)


def test_freqs_match_fft_grid():
    u, v = P.freqs(8, 16, 0.5, 2.0)
    assert np.allclose(u[:, 0], np.fft.fftfreq(8, 0.5))
    assert np.allclose(v[0], np.fft.fftfreq(16, 2.0))


def test_unit_window_is_unitary_and_adjoint_inverts():
    x = rnd((16, 32), 1)
    y = P.fresnel(x, **GEO)
    assert np.linalg.norm(y) == pytest.approx(np.linalg.norm(x), rel=1e-12)
    back = P.fresnel(y, **GEO, herm=True)
    assert np.allclose(back, x, atol=1e-12)


def test_adjoint_identity_without_window():
    x, z = rnd((16, 16), 2), rnd((16, 16), 3)
    lhs = np.vdot(P.fresnel(x, **GEO), z)
    rhs = np.vdot(x, P.fresnel(z, **GEO, herm=True))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_window_breaks_adjointness_as_the_reference_does():
    # Q11: H applies the window, H^H does not
    win = P.window(16, 16, 1.0, 1.0, 0.2)
    x, z = rnd((16, 16), 4), rnd((16, 16), 5)
    lhs = np.vdot(P.fresnel(x, **GEO, win=win), z)
    rhs = np.vdot(x, P.fresnel(z, **GEO, herm=True))
    assert abs(lhs - rhs) > 1e-3 * abs(lhs)


def test_zero_distance_is_identity():
    x = rnd((8, 8), 6)
    assert np.allclose(P.fresnel(x, 1.0, 1.0, GEO["lam"], 0.0), x, atol=1e-13)


@pytest.mark.parametrize("k0,l0", [(1, 0), (3, 5), (-2, 7)])
def test_plane_wave_eigenvalue(k0, l0):
    rows, cols, dr, dc = 16, 32, 0.7, 1.3
    r = np.arange(rows)[:, None]
    t = np.arange(cols)[None, :]
    x = np.exp(2j * np.pi * (k0 * r / rows + l0 * t / cols))
    u0, v0 = k0 / (rows * dr), l0 / (cols * dc)
    sigma = 0.3
    win = P.window(rows, cols, dr, dc, sigma)
    y = P.fresnel(x, dr, dc, GEO["lam"], GEO["dist"], win=win)
    phase = -np.pi * GEO["lam"] * GEO["dist"] * (u0 * u0 + v0 * v0)
    wv = np.exp(-(u0 * u0 + v0 * v0) / (2 * sigma * sigma))
    assert np.allclose(y, x * np.exp(1j * phase) * wv, atol=1e-12)


def problem(seed, rows=16, cols=16, windowed=False):
    """A view's problem; with the reference's window (windowed) H^H is not
    the adjoint of H (Q11), so gradient and step are exact only without it."""
    r = np.random.default_rng(seed)
    win = (P.window(rows, cols, 1.0, 1.0, 2.0 / np.sqrt(GEO["lam"] * GEO["dist"]))
           if windowed else None)
    y = np.abs(rnd((rows, cols), seed)) + 0.5
    D = np.exp(-0.1 * r.random((rows, cols)) - 1j * 0.3 * r.random((rows, cols)))
    Lam = 0.5 + r.random((rows, cols))
    v = rnd((rows, cols), seed + 1) * 0.1 + 1.0
    return P.PhaseProblem(y, D, Lam, v, 0.7, 1.0, 1.0, GEO["lam"], GEO["dist"], win)


def test_gradient_is_the_cost_derivative():
    pp = problem(10)
    w = rnd((16, 16), 11) * 0.2 + 1.0
    om = pp.phase_projection(w)
    g = pp.gradient(om, w)
    h = 1e-6
    for (i, j) in [(0, 0), (3, 7), (15, 2)]:
        for part, unit in ((g.real, 1.0), (g.imag, 1j)):
            wp, wm = w.copy(), w.copy()
            wp[i, j] += h * unit
            wm[i, j] -= h * unit
            fd = (pp.cost(om, wp) - pp.cost(om, wm)) / (2 * h)
            assert fd == pytest.approx(part[i, j], rel=1e-5, abs=1e-7)


def test_windowed_gradient_differs_from_the_cost_derivative():
    # Q11 consequence: with the window the reference's gradient is inexact
    pp = problem(10, windowed=True)
    w = rnd((16, 16), 11) * 0.2 + 1.0
    om = pp.phase_projection(w)
    g = pp.gradient(om, w)
    h = 1e-6
    wp, wm = w.copy(), w.copy()
    wp[3, 7] += h
    wm[3, 7] -= h
    fd = (pp.cost(om, wp) - pp.cost(om, wm)) / (2 * h)
    assert abs(fd - g.real[3, 7]) > 1e-4 * abs(fd)


def test_stepsize_is_the_exact_line_minimiser():
    pp = problem(12)
    w = rnd((16, 16), 13) * 0.2 + 1.0
    om = pp.phase_projection(w)
    g = pp.gradient(om, w)
    a = pp.stepsize(om, g)
    c0 = pp.cost(om, w - a * g)
    for s in (0.9, 1.1):
        assert pp.cost(om, w - s * a * g) > c0
    # the cost is quadratic along g: three points fix the parabola's vertex
    c = [pp.cost(om, w - t * a * g) for t in (0.0, 1.0, 2.0)]
    vertex = 1.0 - (c[2] - c[0]) / (2 * (c[2] - 2 * c[1] + c[0]))
    assert vertex == pytest.approx(1.0, abs=1e-9)


def test_phase_projection_is_unit_modulus_and_optimal():
    pp = problem(14)
    w = rnd((16, 16), 15) * 0.2 + 1.0
    om = pp.phase_projection(w)
    assert np.allclose(np.abs(om), 1.0, atol=1e-12)
    # Omega maximises Re(conj(y) Omega H D w) pointwise: any other phase costs more
    c0 = pp.cost(om, w)
    assert pp.cost(om * np.exp(0.05j), w) > c0


def test_sd_decreases_the_cost():
    pp = problem(16)  # unit window: exact gradient and step
    w = rnd((16, 16), 17) * 0.2 + 1.0
    om = pp.phase_projection(w)
    c_prev = pp.cost(om, w)
    for _ in range(5):
        w, om, upd = pp.sd_iter(om, w)
        c = pp.cost(om, w)
        assert c <= c_prev * (1 + 1e-12)
        c_prev = c


# ------------------------------------------------------------- NMS (a13)
PRIOR = dict(phase_sig=2e-2, phase_c=1e-2, mag_sig=2e-2, mag_c=1e-2, delta_over_beta=50.0)


def test_prior_filter_is_the_normalised_inverse_distance_slice():
    f = P.prior_filter_2d()
    s = 6 + 12 / np.sqrt(2) + 8 / np.sqrt(3)
    assert f[1, 1] == 0
    assert f[0, 1] == pytest.approx(1 / s) and f[0, 0] == pytest.approx(1 / np.sqrt(2) / s)
    assert np.allclose(f, f.T) and np.allclose(f, f[::-1, ::-1])


def test_qggmrf_value_special_cases():
    pr = P.NmsPrior(**PRIOR, decorrelate=False)
    assert pr.value(0.0, 0) == 0.0
    # p = q: the potential is the quadratic d^2 / (sigma^2 (c + 1))
    pq = P.NmsPrior(**PRIOR, decorrelate=False, p=2.0)
    assert pq.value(0.3, 1) == pytest.approx(0.09 / (2e-2 ** 2 * (1e-2 + 1)), rel=1e-14)


def test_cost_without_neighbours_is_the_data_term():
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    zr, zi = 0.3, 1.1
    b = 0.4 - 0.2j
    fl = np.zeros((3, 3), bool)
    c = P.cost_nms(zr, zi, b.real, b.imag, 0.5, np.zeros((3, 3)), np.zeros((3, 3)), fl, pr)
    assert c == pytest.approx(0.5 / 2 * abs(np.exp(-(zr + 1j * zi)) - b) ** 2, rel=1e-14)


def test_cost_of_equal_neighbours_has_no_prior_term():
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    fl = np.ones((3, 3), bool)
    nr, ni = np.full((3, 3), 0.3), np.full((3, 3), 1.1)
    c = P.cost_nms(0.3, 1.1, 0.0, 0.0, 0.5, nr, ni, fl, pr)
    assert c == pytest.approx(0.5 / 2 * np.exp(-0.6), rel=1e-14)


def test_cost_single_neighbour_decorrelated_terms():
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    fl = np.zeros((3, 3), bool)
    fl[1, 2] = True
    nr, ni = np.zeros((3, 3)), np.zeros((3, 3))
    nr[1, 2], ni[1, 2] = 0.1, 0.5
    zr, zi = 0.3, 1.1
    data = P.cost_nms(zr, zi, 0, 0, 0.5, nr, ni, np.zeros((3, 3), bool), pr)
    c = P.cost_nms(zr, zi, 0, 0, 0.5, nr, ni, fl, pr)
    db, dd = zr - 0.1, zi - 0.5
    k = PRIOR["delta_over_beta"]
    want = pr.filt[1, 2] * (pr.value(dd + k * db, 0) + pr.value(dd - k * db, 1))
    assert c - data == pytest.approx(want, rel=1e-12)


def test_nms_finds_the_known_minimum():
    pr = P.NmsPrior(**PRIOR, decorrelate=False)
    zs = 0.35 + 0.8j
    b = np.exp(-zs)
    fl = np.zeros((3, 3), bool)
    zr, zi, it = P.nms_search(0.3, 0.7, b.real, b.imag, 0.5, np.zeros((3, 3)),
                              np.zeros((3, 3)), fl, pr, thresh=1e-9, max_iter=400)
    assert abs(zr - zs.real) < 1e-4 and abs(zi - zs.imag) < 1e-4


def test_nms_pass_never_raises_a_pixel_cost():
    # the reference checks this after every search (XT_CmplxProjEst.c:162-163)
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    r = np.random.default_rng(3)
    zr, zi = r.random((6, 6)) * 0.05 + 0.02, r.random((6, 6)) * 0.5 + 0.2
    b = P.transmittance(zr * 1.2, zi * 0.8)
    for (j, k) in [(0, 0), (2, 3), (5, 5)]:
        nr, ni, fl = P.neighbours(zr, zi, j, k)
        c0 = P.cost_nms(zr[j, k], zi[j, k], b[j, k].real, b[j, k].imag, 0.5, nr, ni, fl, pr)
        a, bb, _ = P.nms_search(zr[j, k], zi[j, k], b[j, k].real, b[j, k].imag, 0.5, nr, ni, fl, pr)
        assert P.cost_nms(a, bb, b[j, k].real, b[j, k].imag, 0.5, nr, ni, fl, pr) <= c0


def test_product_geometry_helpers_match_the_oracle():
    from phasecon_amd.pcmbir import wavelength, fresnel_window
    assert wavelength(17.0) == P.wavelength(17.0)
    assert np.allclose(fresnel_window(16, 32, 0.7, 1.3, 0.2), P.window(16, 32, 0.7, 1.3, 0.2),
                       rtol=1e-15, atol=0)


# ------------------------------------------------- composed quantities (a13)
def test_transmittance_is_the_complex_exponential():
    r = np.random.default_rng(5)
    zr, zi = r.random((4, 5)), r.standard_normal((4, 5)) * 3
    assert np.allclose(P.transmittance(zr, zi), np.exp(-(zr + 1j * zi)), rtol=1e-14, atol=0)


def test_pretcost_at_zero_distance_by_brute_force():
    """At zero distance H is the identity, so the data term is the per-pixel
    sum of Lam (|y| - |D| exp(-zr))^2 / 2; the prior over the four forward
    pairs equals half the sum over every ordered pair of 8-neighbours (the
    filter is symmetric and the potential even), enumerated independently."""
    rows, cols = 5, 6
    r = np.random.default_rng(6)
    zr, zi = r.random((rows, cols)) * 0.1, r.random((rows, cols))
    meas = (r.random((rows, cols)) + 0.5).astype(complex)
    D = np.exp(-0.2 * r.random((rows, cols)) - 1j * r.random((rows, cols)))
    Lam = 0.5 + r.random((rows, cols))
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    got = P.pret_cost(meas, D, Lam, zr, zi, pr, (1.0, 1.0, P.wavelength(17.0), 0.0, None))
    data = np.sum(Lam * (meas.real - np.abs(D) * np.exp(-zr)) ** 2) / 2
    T = pr.T
    prior = 0.0
    for j in range(rows):
        for k in range(cols):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    a, b = j + dj, k + dk
                    if (dj, dk) == (0, 0) or not (0 <= a < rows and 0 <= b < cols):
                        continue
                    db, dd = zr[j, k] - zr[a, b], zi[j, k] - zi[a, b]
                    prior += pr.filt[1 + dj, 1 + dk] * (
                        pr.value(T[0, 0] * dd + T[0, 1] * db, 0)
                        + pr.value(T[1, 0] * dd + T[1, 1] * db, 1)) / 2
    assert got == pytest.approx(data + prior, rel=1e-12)


def test_phase_cost_minimiser_at_zero_distance():
    """With H the identity and |Omega| = 1 the phase-retrieval cost is
    separable and minimised by w* = (Lam conj(Omega D) y + nu v) /
    (Lam |D|^2 + nu), derived by hand: the gradient vanishes there and every
    perturbation raises the cost."""
    rows, cols, nu = 4, 8, 0.7
    r = np.random.default_rng(7)
    y = rnd((rows, cols), 8)
    D = np.exp(-0.1 * r.random((rows, cols)) - 0.3j * r.random((rows, cols)))
    Lam = 0.5 + r.random((rows, cols))
    v = rnd((rows, cols), 9)
    om = np.exp(6j * r.random((rows, cols)))
    pp = P.PhaseProblem(y, D, Lam, v, nu, 1.0, 1.0, P.wavelength(17.0), 0.0, None)
    ws = (Lam * np.conj(om * D) * y + nu * v) / (Lam * np.abs(D) ** 2 + nu)
    assert np.max(np.abs(pp.gradient(om, ws))) < 1e-12 * np.max(np.abs(y))
    c0 = pp.cost(om, ws)
    for s in range(5):
        assert pp.cost(om, ws + 1e-3 * rnd((rows, cols), 100 + s)) > c0


# ----------------------------------------------------- Paganin (next row)
def test_paganin_uniform_slab_gives_its_projection():
    """A uniform slab of thickness T transmits exp(-2 k beta T) at every
    pixel; the filter passes the mean unchanged, so the retrieved projection
    is (k beta T, k delta T) -- the slab's own, by Beer-Lambert."""
    lam, dob = P.wavelength(17.0), 50.0
    k = 2 * np.pi / lam
    beta, T = 1e-9, 2000.0
    inten = np.exp(-2 * k * beta * T)
    D = np.full((8, 16), 0.9 + 0.1j)
    meas = np.sqrt(inten) * np.abs(D) * np.ones((8, 16))
    zr, zi = P.paganin(meas.astype(complex), D, 1.0, 1.0, lam, 1e5, dob)
    assert np.allclose(zr, k * beta * T, rtol=1e-12)
    assert np.allclose(zi, k * dob * beta * T, rtol=1e-12)


@pytest.mark.parametrize("k0,l0", [(1, 0), (2, 3), (0, 5)])
def test_paganin_filter_is_the_published_form(k0, l0):
    """For an intensity 1 + eps cos(2 pi (u0 r + v0 t)) the filtered
    intensity is 1 + eps f cos(...) with Paganin's filter
    f = 1 / (1 + pi lambda z (delta/beta) |u0|^2) (Paganin et al., J.
    Microsc. 206 (2002) 33-40), so the thickness is -log of that exactly."""
    rows, cols, dr, dc = 16, 32, 0.7, 1.3
    lam, dist, dob, eps = P.wavelength(17.0), 1e5, 30.0, 0.2
    r = np.arange(rows)[:, None] * dr
    t = np.arange(cols)[None, :] * dc
    u0, v0 = k0 / (rows * dr), l0 / (cols * dc)
    ph = 2 * np.pi * (u0 * r + v0 * t)
    inten = 1 + eps * np.cos(ph)
    D = np.ones((rows, cols), complex)
    zr, zi = P.paganin(np.sqrt(inten).astype(complex), D, dr, dc, lam, dist, dob)
    f = 1 / (1 + np.pi * lam * dist * dob * (u0 ** 2 + v0 ** 2))
    thick = -np.log(1 + eps * f * np.cos(ph))
    assert np.allclose(zr, thick / 2, rtol=0, atol=1e-13)
    assert np.allclose(zi, thick * dob / 2, rtol=0, atol=1e-11)


def _admm_problem(n=8, seed=0, uniform=False):
    """A noise-free view: y = |H exp(-z*)| (D = 1, Lambda = 1)."""
    r = np.random.default_rng(seed)
    if uniform:
        zr_t, zi_t = np.full((n, n), 0.025), np.full((n, n), 0.35)
    else:
        zr_t, zi_t = 0.02 + 0.01 * r.random((n, n)), 0.3 + 0.1 * r.random((n, n))
    geo = (1.0, 1.0, P.wavelength(17.0), 1e5, None)
    y = np.abs(P.fresnel(P.transmittance(zr_t, zi_t), *geo[:4])).astype(complex)
    return zr_t, zi_t, y, np.ones((n, n), complex), np.ones((n, n)), geo


def test_admm_solution_is_a_fixed_point():
    """estimate_complex_projection started at the exact solution of a
    noise-free uniform view (where the prior is flat) stays there: z, w and
    the dual v unchanged, primal residual 0, and the pretcost 0."""
    zr_t, zi_t, y, D, Lam, geo = _admm_problem(uniform=True)
    pr = P.NmsPrior(**PRIOR, decorrelate=True)
    w0 = P.transmittance(zr_t, zi_t)
    om0 = P.PhaseProblem(y, D, Lam, None, 0.5, *geo).phase_projection(w0)
    with np.errstate(invalid="ignore"):   # dual residual 0 / |v| = 0 / 0, as in the reference
        res = P.estimate_complex_projection(y, om0, D, zr_t.copy(), zi_t.copy(), Lam, w0, pr, geo,
                                            0.5, 3, 0.0, 5, 0.0)
    assert np.array_equal(res["zr"], zr_t) and np.array_equal(res["zi"], zi_t)
    assert np.max(np.abs(res["w"] - w0)) < 1e-12
    assert np.max(np.abs(res["v"])) < 1e-12
    assert max(res["primal"]) < 1e-12
    assert res["costs"][-1] < 1e-20


def test_admm_dual_is_the_scaled_gradient_of_the_data_term():
    """Scaled ADMM for min_w f(w) + g(z) s.t. w = exp(-z), with the z-step on
    b = w - v, the w-step to the target exp(-z) + v and v += exp(-z) - w:
    when the w-step is solved exactly, its optimality condition
    grad f(w) + nu (w - exp(-z) - v_old) = 0 and the dual update give
    grad f(w) = nu v after every iteration, f(w) = |y - Omega H D w|^2_Lam / 2
    at the phase projection Omega of w.  A flipped dual sign gives
    grad f = -nu v (relative error 2), and the steps in another order break
    the identity.  The residuals fall on this noise-free view."""
    zr_t, zi_t, y, D, Lam, geo = _admm_problem(seed=0)
    r = np.random.default_rng(1)
    zr0 = zr_t + 0.01 * r.standard_normal(zr_t.shape)
    zi0 = zi_t + 0.05 * r.standard_normal(zi_t.shape)
    pr = P.NmsPrior(phase_sig=1.0, phase_c=0.01, mag_sig=1.0, mag_c=0.01, delta_over_beta=50.0,
                    decorrelate=False)
    nu = 0.5
    w0 = P.transmittance(zr0, zi0)
    pp = P.PhaseProblem(y, D, Lam, None, nu, *geo)
    res = P.estimate_complex_projection(y, pp.phase_projection(w0), D, zr0, zi0, Lam, w0, pr, geo,
                                        nu, 20, 0.0, 50, 0.0,
                                        nms_kw=dict(thresh=1e-12, max_iter=400))
    w, v = res["w"], res["v"]
    om = pp.phase_projection(w)
    gf = pp.herm_omega_hd(om, (pp.omega_hd(om, w) - y) * Lam)
    assert np.linalg.norm(gf - nu * v) < 1e-9 * np.linalg.norm(nu * v)
    prim = np.array(res["primal"])
    assert prim[-1] < 0.01 * prim[0]
    assert res["dual"][-1] < 0.01 * res["dual"][0]
    assert res["costs"][-1] < 0.1 * res["costs"][0]


def test_admm_z_step_targets_w_minus_v():
    """Without a prior (sigmas 1e6) the z-step's NMS cost is
    nu/2 |exp(-z) - b|^2, minimised where exp(-z) = b; after the loop the last
    z must therefore satisfy exp(-z) = w_prev - v_prev, the state that
    iteration started from (XT_CmplxProjEst.c:121-122: b = w - v)."""
    zr_t, zi_t, y, D, Lam, geo = _admm_problem(seed=0)
    r = np.random.default_rng(1)
    zr0 = zr_t + 0.01 * r.standard_normal(zr_t.shape)
    zi0 = zi_t + 0.05 * r.standard_normal(zi_t.shape)
    pr = P.NmsPrior(phase_sig=1e6, phase_c=0.01, mag_sig=1e6, mag_c=0.01, delta_over_beta=50.0,
                    decorrelate=False)
    w0 = P.transmittance(zr0, zi0)
    pp = P.PhaseProblem(y, D, Lam, None, 0.5, *geo)
    res = P.estimate_complex_projection(y, pp.phase_projection(w0), D, zr0, zi0, Lam, w0, pr, geo,
                                        0.5, 3, 0.0, 50, 0.0,
                                        nms_kw=dict(thresh=1e-14, max_iter=1000))
    b = res["w_prev"] - res["v_prev"]
    assert np.abs(res["v_prev"]).max() > 1e-3          # the dual is not trivially zero
    assert np.abs(P.transmittance(res["zr"], res["zi"]) - b).max() < 1e-8
