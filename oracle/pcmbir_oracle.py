"""FP64 numpy restatement of the PCMBIR phase-retrieval path (TEST
INFRASTRUCTURE ONLY: imported by tests/ to test against, never by the product).

Follows, per function, the reference file:line it cites (paths relative to the
reference's Source_Code/PCMBIR_3D/).  The reference computes its FFTs with
FFTW3, which is not present here; FFTW's conventions (forward kernel
exp(-2 pi i jk/N), backward exp(+2 pi i jk/N), both unnormalised) are those of
numpy.fft.fft2 and N * numpy.fft.ifft2, and the reference divides by N once in
the spectral multiply, so H = ifft2(fft2(x) * factor) below.  The FFTW results
themselves are not pinned by any reference fixture; these functions are pinned
instead by closed forms in tests/test_pcmbir_oracle.py (unitarity and
H^H H = I with a unit window, the identity at zero distance, the plane-wave
eigenvalue, the gradient by finite differences of the cost, the step size as
the exact line minimiser, the cost's hand-derived minimiser at zero distance,
the NMS cost's special cases and a known minimum, pretcost by brute force at
zero distance, the transmittance against numpy's complex exponential, the
Paganin filter against its published form 1 / (1 + pi lambda z (delta/beta)
|u|^2) and a uniform slab's thickness).

estimate_complex_projection (XT_CmplxProjEst.c:105-229) is pinned by what
the ADMM splitting fixes: the exact solution of a noise-free uniform view is
a fixed point; after every iteration the scaled dual equals the data term's
gradient over nu (grad f(w) = nu v, which a flipped dual sign, another
w-step target or another step order breaks); without a prior the z-step
lands on exp(-z) = w - v; and the residuals fall on a noise-free view.  Its
two residual normalisations (the RMS of exp(-z) - w over the pixels, and the
change of w relative to |v|) are pinned only by their reading of the
reference.  nms_pass(order="colour") is the GPU's four-colour visiting
order and has no reference counterpart; order="raster" is the reference's.

Arrays are per view [rows][cols] = [N_r][N_t], complex128 (real/imag pairs
of the reference's *_real / *_imag arrays).
"""
from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).tiny  # EPSILON_ERROR = DBL_MIN, XT_Constants.h:48
PLANCKS_CONSTANT = 4.135668e-19  # XT_Constants.h:84 (keV s; 10x below CODATA)
LIGHT_SPEED = 299792458e6        # XT_Constants.h:85 (um/s)


def wavelength(light_energy_kev):
    """XT_Init.c:258."""
    return PLANCKS_CONSTANT * LIGHT_SPEED / light_energy_kev


def freqs(rows, cols, d_rows, d_cols):
    """Spatial frequencies of the FFT grid (XT_FresnelTran.c:33-41): index i
    maps to i/(rows d) below rows/2 and -(rows-i)/(rows d) from rows/2 on."""
    i = np.arange(rows, dtype=np.float64)
    j = np.arange(cols, dtype=np.float64)
    u = np.where(i >= rows // 2, -(rows - i), i) / (rows * d_rows)
    v = np.where(j >= cols // 2, -(cols - j), j) / (cols * d_cols)
    return u[:, None], v[None, :]


def window(rows, cols, d_rows, d_cols, sigma):
    """Gaussian frequency window, create_FresnelTranWindow (XT_Init.c:210-231)."""
    u, v = freqs(rows, cols, d_rows, d_cols)
    return np.exp(-(u * u + v * v) / (2 * sigma * sigma))


def fresnel_factor(rows, cols, d_rows, d_cols, lam, dist, win=None, herm=False):
    """Spectral factor of H (XT_FresnelTran.c:33-50: exp(i phi) W with
    phi = -pi lam z (u^2 + v^2)) or of H^H (XT_FresnelTran.c:104-118:
    exp(-i phi), no window -- Q11)."""
    u, v = freqs(rows, cols, d_rows, d_cols)
    phi = -(u * u + v * v) * np.pi * lam * dist
    if herm:
        return np.exp(-1j * phi)
    w = np.ones((rows, cols)) if win is None else win
    return (np.cos(phi) * w) + 1j * (np.sin(phi) * w)


def fresnel(x, d_rows, d_cols, lam, dist, win=None, herm=False):
    """y = H x or H^H x over the last two axes (compute_FresnelTran,
    XT_FresnelTran.c:12-78; compute_HermFresnelTran, :82-137)."""
    rows, cols = x.shape[-2:]
    f = fresnel_factor(rows, cols, d_rows, d_cols, lam, dist, win, herm)
    return np.fft.ifft2(np.fft.fft2(x) * f)


EXPECTED_COUNT = 2500.0  # EXPECTED_COUNT_MEASUREMENT (PCMBIR_3D/XT_Constants.h:89)


def simulate(mag, phase, angles, proj_rows, proj_cols, vox_wid, rot_center, obj2det_dist,
             light_energy):
    """Noise-free measurements of the phantom simulator (ForwardProject,
    PCMBIR_3D/XT_genSinogram.c:146-264): the phantoms' projections
    (oracle.phantom_project), the exit field sqrt(I0) exp(-p_mag - i
    p_phase), Fresnel propagation with the simulator's window (sigma = 1 /
    sqrt(lambda z / 2), PCMBIR_3D/XT_Init.c:682-686), the intensity averaged
    over each detector pixel's samples.  Returns [n_p][proj_rows][proj_cols]."""
    from . import oracle as O
    lam = wavelength(light_energy)
    k = 2 * np.pi / lam
    n_z, n_xy = np.shape(mag)[0], np.shape(mag)[-1]
    pr, pi = O.phantom_project(mag, phase, angles, proj_rows, proj_cols, vox_wid, rot_center, k)
    d_r, d_t = vox_wid * proj_cols / n_xy, vox_wid * proj_rows / n_z
    field = np.sqrt(EXPECTED_COUNT) * np.exp(-pr) * (np.cos(-pi) + 1j * np.sin(-pi))
    field = np.swapaxes(field, 1, 2)                      # [view][r][t], the transform's layout
    win = window(n_xy, n_z, d_r, d_t, 1.0 / np.sqrt(0.5 * lam * obj2det_dist))
    y = fresnel(field, d_r, d_t, lam, obj2det_dist, win)
    inten = np.abs(y) ** 2                                # [view][r][t]
    rs, ts = n_xy // proj_cols, n_z // proj_rows
    m = inten.reshape(len(angles), proj_cols, rs, proj_rows, ts).mean(axis=(2, 4))
    return np.swapaxes(m, 1, 2)


def paganin(meas, D, d_rows, d_cols, lam, dist, delta_over_beta):
    """Single-material Paganin retrieval of one view
    (paganins_1mat_phase_retrieval, XT_Paganin.c:77-134): intensity
    meas^2 / |D|^2 (real part of the measurement amplitude), FFT, times
    1 / ((z (delta/beta) / (2 k) (u^2 + v^2) 4 pi^2 + 1) N), inverse FFT,
    thickness -log|Re|, and (z_real, z_imag) = (thick / 2,
    thick (delta/beta) / 2); k = 2 pi / lam."""
    rows, cols = meas.shape
    k = 2 * np.pi / lam
    reg = delta_over_beta / (2 * k)
    inten = meas.real ** 2 / (D.real ** 2 + D.imag ** 2)
    u, v = freqs(rows, cols, d_rows, d_cols)
    tran = dist * reg * (u * u + v * v) * 4 * np.pi * np.pi + 1
    f = np.fft.ifft2(np.fft.fft2(inten) / tran)
    thick = -np.log(np.abs(f.real))
    return thick / 2, thick * delta_over_beta / 2


class PhaseProblem:
    """One view's phase-retrieval problem (XT_PhaseRet.c): measurements y,
    object-independent factor D, weights Lambda (real), proximal target v,
    penalty nu, Fresnel geometry."""

    def __init__(self, y, D, Lam, v, nu, d_rows, d_cols, lam, dist, win):
        self.y, self.D, self.Lam, self.v, self.nu = y, D, Lam, v, nu
        self.geo = (d_rows, d_cols, lam, dist, win)

    def H(self, x):
        return fresnel(x, *self.geo)

    def HH(self, x):
        return fresnel(x, *self.geo, herm=True)

    def phase_projection(self, w):
        """Omega = (y / (H D w)) / |y / (H D w)| (XT_PhaseRet.c:10-33, with
        cmplx_div's DBL_MIN guard, XT_CmplxArith.c)."""
        hdw = self.H(self.D * w)
        q = self.y * np.conj(hdw) / (np.abs(hdw) ** 2 + EPS)
        return q / (np.abs(q) + EPS)

    def omega_hd(self, om, x):
        """Omega H D x (mult_OmegaHD, XT_PhaseRet.c:59-79)."""
        return self.H(self.D * x) * om

    def herm_omega_hd(self, om, x):
        """(Omega H D)^H x (mult_Herm_OmegaHD, XT_PhaseRet.c:35-57)."""
        return self.HH(np.conj(om) * x) * np.conj(self.D)

    def gradient(self, om, w):
        """g = (OHD)^H Lam (OHD) w - (OHD)^H Lam y + nu (w - v)
        (compute_gradient, XT_PhaseRet.c:81-131)."""
        b1 = self.herm_omega_hd(om, self.omega_hd(om, w) * self.Lam)
        b2 = self.herm_omega_hd(om, self.Lam * self.y)
        return b1 - b2 + self.nu * (w - self.v)

    def stepsize(self, om, g):
        """alpha = |g|^2 / (g^H (OHD)^H Lam (OHD) g + nu |g|^2)
        (compute_stepsize, XT_PhaseRet.c:133-170; the real part of the
        Hermitian form, as the reference accumulates it)."""
        num = np.sum(np.abs(g) ** 2)
        b = self.herm_omega_hd(om, self.omega_hd(om, g) * self.Lam)
        den = np.sum(g.real * b.real + g.imag * b.imag) + self.nu * num
        return num / (den + EPS)

    def cost(self, om, w):
        """|y - OHD w|^2_Lam / 2 + nu |w - v|^2 / 2 (compute_cost,
        XT_PhaseRet.c:172-205)."""
        r = self.y - self.omega_hd(om, w)
        return (np.sum(np.abs(r) ** 2 * self.Lam) / 2
                + self.nu * np.sum(np.abs(w - self.v) ** 2) / 2)

    def sd_iter(self, om, w):
        """One steepest-descent iteration (steepest_descent_iter,
        XT_PhaseRet.c:207-270): gradient, exact step, update of w, then the
        phase projection.  Returns (w, Omega, 100 sum|alpha g| / sum|w|).
        The reference's cost evaluation before the gradient only feeds debug
        output and is left out."""
        g = self.gradient(om, w)
        alpha = self.stepsize(om, g)
        w = w - alpha * g
        upd = np.sum(np.abs(alpha * g)) / np.sum(np.abs(w))
        return w, self.phase_projection(w), upd * 100

    def sd_run(self, om, w, iters, thresh):
        """The steepest-descent loop of estimate_complex_projection
        (XT_CmplxProjEst.c:165-169): stop after an iteration j > 1 whose
        update is below thresh.  Returns (w, Omega, iterations run, last
        update)."""
        upd = 0.0
        j = 0
        for j in range(iters):
            w, om, upd = self.sd_iter(om, w)
            if upd < thresh and j > 1:
                j += 1
                break
        else:
            j = iters
        return w, om, j, upd


# --------------------------------------------------------------- a13: NMS
def prior_filter_2d():
    """The 3x3 in-plane slice Spatial_Filter[1][.][.] of the 26-neighbour
    1/distance filter normalised over all 26 neighbours (initFilter,
    XT_Init.c:145-177; the time filter is 0)."""
    d = np.array([[np.sqrt((i - 1.0) ** 2 + (j - 1.0) ** 2) for j in range(3)] for i in range(3)])
    s = 6 * 1.0 + 12 / np.sqrt(2.0) + 8 / np.sqrt(3.0)
    f = np.zeros((3, 3))
    f[d > 0] = 1.0 / d[d > 0] / s
    return f


class NmsPrior:
    """qGGMRF parameters of the projection-domain prior of Cost_NMS
    (XT_Prior.c:7-15 values, XT_Init.c:482-490 sigmas, compute_DecorrObjTransform
    XT_Init.c:234-246), q = 2, p = MRF_P = 1.2."""

    def __init__(self, phase_sig, phase_c, mag_sig, mag_c, delta_over_beta, decorrelate, p=1.2):
        self.p = p
        self.ph = (phase_c, phase_sig ** 2, phase_sig ** (2 - p))
        self.mg = (mag_c, mag_sig ** 2, mag_sig ** (2 - p))
        self.T = (np.array([[1.0, delta_over_beta], [1.0, -delta_over_beta]]) if decorrelate
                  else np.eye(2))
        self.filt = prior_filter_2d()

    def value(self, d, which):
        c, sq, sqp = self.ph if which == 0 else self.mg
        ad = abs(d)
        return (ad * ad / sq) / (c + ad ** (2 - self.p) / sqp)


def cost_nms(zr, zi, br, bi, nu, zr_nb, zi_nb, flags, pr, tomo=None):
    """Cost_NMS (XT_Search.c:8-43): nu/2 |exp(-z) - b|^2 plus, without
    ENABLE_TOMO_RECONS (Q12), the decorrelated qGGMRF terms of the 3x3
    neighbours, or with it (tomo = (mu, ar, ai), XT_Search.c:18-19) the
    proximal term mu/2 |z - a|^2 towards the tomographic projection."""
    e = np.exp(-zr)
    cv, sv = e * np.cos(-zi), e * np.sin(-zi)
    cost = ((cv - br) * (cv - br) + (sv - bi) * (sv - bi)) * nu / 2
    if tomo is not None:
        mu, ar, ai = tomo
        return cost + ((zr - ar) * (zr - ar) + (zi - ai) * (zi - ai)) * mu / 2
    T = pr.T
    for i in range(3):
        for j in range(3):
            if flags[i][j]:
                db = zr - zr_nb[i][j]
                dd = zi - zi_nb[i][j]
                cost += pr.filt[i, j] * pr.value(T[0, 0] * dd + T[0, 1] * db, 0)
                cost += pr.filt[i, j] * pr.value(T[1, 0] * dd + T[1, 1] * db, 1)
    return cost


NMS_DEFAULTS = dict(rho=1.0, chi=2.0, gamma=0.5, sigma=0.5, thresh=0.005, max_iter=50)


def nms_search(zr1, zi1, br, bi, nu, zr_nb, zi_nb, flags, pr, rho=1.0, chi=2.0, gamma=0.5,
               sigma=0.5, thresh=0.005, max_iter=50, tomo=None):
    """Nelder_Mead_Simplex_2DSearch (XT_Search.c:100-215, including its
    simplex initialisation (2 zr, zi), (zr, 2 zi)): returns the best vertex
    and the iteration count."""
    f = lambda a, b: cost_nms(a, b, br, bi, nu, zr_nb, zi_nb, flags, pr, tomo)  # noqa: E731
    zr = [0.0, zr1, 2 * zr1, zr1]
    zi = [0.0, zi1, zi1, 2 * zi1]
    c = [0.0, f(zr[1], zi[1]), f(zr[2], zi[2]), f(zr[3], zi[3])]

    def swap(a, b):
        zr[a], zr[b] = zr[b], zr[a]
        zi[a], zi[b] = zi[b], zi[a]
        c[a], c[b] = c[b], c[a]

    def sort():  # NMS_Sort
        if c[3] < c[2]:
            swap(2, 3)
        if c[2] < c[1]:
            swap(1, 2)
        if c[3] < c[2]:
            swap(2, 3)

    def accept(r, i_, cost):  # NMS_Accept
        zr[3], zi[3], c[3] = r, i_, cost
        if c[3] < c[2]:
            swap(2, 3)
        if c[2] < c[1]:
            swap(1, 2)

    def shrink():  # NMS_Shrink
        zr[2] = zr[1] + sigma * (zr[2] - zr[1])
        zi[2] = zi[1] + sigma * (zi[2] - zi[1])
        c[2] = f(zr[2], zi[2])
        zr[3] = zr[1] + sigma * (zr[3] - zr[1])
        zi[3] = zi[1] + sigma * (zi[3] - zi[1])
        c[3] = f(zr[3], zi[3])
        sort()

    sort()
    zr[0] = (zr[1] + zr[2]) / 2.0
    zi[0] = (zi[1] + zi[2]) / 2.0
    it = 0
    for it in range(max_iter):
        ocr, oci = list(zr), list(zi)
        zr_ref = zr[0] + rho * (zr[0] - zr[3])
        zi_ref = zi[0] + rho * (zi[0] - zi[3])
        c_ref = f(zr_ref, zi_ref)
        if c_ref < c[2] and c_ref >= c[1]:
            accept(zr_ref, zi_ref, c_ref)
        elif c_ref < c[1]:
            zr_e = zr[0] + chi * (zr_ref - zr[0])
            zi_e = zi[0] + chi * (zi_ref - zi[0])
            c_e = f(zr_e, zi_e)
            if c_e < c_ref:
                accept(zr_e, zi_e, c_e)
            else:
                accept(zr_ref, zi_ref, c_ref)
        elif c_ref >= c[2] and c_ref < c[3]:
            zr_o = zr[0] + gamma * (zr_ref - zr[0])
            zi_o = zi[0] + gamma * (zi_ref - zi[0])
            c_o = f(zr_o, zi_o)
            if c_o <= c_ref:
                accept(zr_o, zi_o, c_o)
            else:
                shrink()
        elif c_ref >= c[3]:
            zr_i = zr[0] - gamma * (zr[0] - zr[3])
            zi_i = zi[0] - gamma * (zi[0] - zi[3])
            c_i = f(zr_i, zi_i)
            if c_i < c[3]:
                accept(zr_i, zi_i, c_i)
            else:
                shrink()
        zr[0] = (zr[1] + zr[2]) / 2.0
        zi[0] = (zi[1] + zi[2]) / 2.0
        num = sum(abs(ocr[k] - zr[k]) + abs(oci[k] - zi[k]) for k in range(4))
        den = sum(abs(ocr[k]) + abs(oci[k]) for k in range(4))
        t = num / den * 100
        if t < thresh and it > 1:
            break
    else:
        it = max_iter
    return zr[1], zi[1], it


def neighbours(zr, zi, j, k):
    """3x3 neighbourhood of pixel (j, k) with in-bounds flags
    (XT_CmplxProjEst.c:124-160)."""
    rows, cols = zr.shape
    nr = np.zeros((3, 3))
    ni = np.zeros((3, 3))
    fl = np.zeros((3, 3), bool)
    for m in range(3):
        for n in range(3):
            a, b = j + m - 1, k + n - 1
            if 0 <= a < rows and 0 <= b < cols:
                nr[m, n], ni[m, n], fl[m, n] = zr[a, b], zi[a, b], True
    return nr, ni, fl


def nms_pass(zr, zi, b, nu, pr, order="raster", tomo=None, **kw):
    """One NMS pass over a view's pixels with b = w - v: the reference's raster
    order (XT_CmplxProjEst.c:118-170, each pixel sees its already-updated
    neighbours), or the four-colour order of the GPU kernel ((j%2, k%2) =
    (0,0), (0,1), (1,0), (1,1), pixels of a colour independent).  tomo =
    (mu, a), a complex [rows][cols] (real: magnitude, imag: phase): the
    tomographic coupling of cost_nms.  Returns the new (zr, zi) and the summed
    iteration counts."""
    zr, zi = zr.copy(), zi.copy()
    rows, cols = zr.shape
    if order == "raster":
        pix = [(j, k) for j in range(rows) for k in range(cols)]
    else:
        pix = [(j, k) for cj in (0, 1) for ck in (0, 1)
               for j in range(cj, rows, 2) for k in range(ck, cols, 2)]
    total = 0
    for j, k in pix:
        nr, ni, fl = neighbours(zr, zi, j, k)
        tk = None if tomo is None else (tomo[0], tomo[1][j, k].real, tomo[1][j, k].imag)
        zr[j, k], zi[j, k], it = nms_search(zr[j, k], zi[j, k], b[j, k].real, b[j, k].imag, nu,
                                            nr, ni, fl, pr, tomo=tk, **kw)
        total += it
    return zr, zi, total


def transmittance(zr, zi):
    """exp(-z) as (exp(-zr) cos(-zi), exp(-zr) sin(-zi)) (XT_CmplxProjEst.c:175-178)."""
    e = np.exp(-zr)
    return e * np.cos(-zi) + 1j * e * np.sin(-zi)


def pret_cost(meas, D, Lam, zr, zi, pr, geo, tomo=None):
    """compute_pretcost (XT_CmplxProjEst.c:15-103): sum Lam (|y| -
    |H(D exp(-z))|)^2 / 2 over the real part of the measurements, plus
    without ENABLE_TOMO_RECONS the decorrelated prior over the right, down,
    down-right and down-left pairs, or with it (tomo = (mu, a),
    XT_CmplxProjEst.c:38-41) sum mu/2 |z - a|^2."""
    h = fresnel(transmittance(zr, zi) * D, *geo)
    fwd = np.sum((meas.real - np.abs(h)) ** 2 * Lam / 2)
    if tomo is not None:
        mu, a = tomo
        return fwd + np.sum(mu * ((zr - a.real) ** 2 + (zi - a.imag) ** 2) / 2)
    rows, cols = zr.shape
    T, f = pr.T, pr.filt
    prior = 0.0
    for j in range(rows):
        for k in range(cols):
            pairs = []
            if k + 1 < cols:
                pairs.append((j, k + 1, f[1, 2]))
            if j + 1 < rows:
                pairs.append((j + 1, k, f[2, 1]))
                if k + 1 < cols:
                    pairs.append((j + 1, k + 1, f[2, 2]))
                if k - 1 >= 0:
                    pairs.append((j + 1, k - 1, f[2, 0]))
            for (a, b, wt) in pairs:
                db = zr[j, k] - zr[a, b]
                dd = zi[j, k] - zi[a, b]
                prior += wt * pr.value(T[0, 0] * dd + T[0, 1] * db, 0)
                prior += wt * pr.value(T[1, 0] * dd + T[1, 1] * db, 1)
    return fwd + prior


def estimate_complex_projection(meas, om, D, zr, zi, Lam, w, pr, geo, nu, pret_iters,
                                pret_thresh, sd_iters, sd_thresh, order="raster", nms_kw=None,
                                tomo=None):
    """One view's ADMM loop (estimate_complex_projection,
    XT_CmplxProjEst.c:105-229): NMS pass on z with b = w - v, steepest
    descent on w towards exp(-z) + v, dual update v += exp(-z) - w,
    residuals, pretcost, stop on the dual residual.  tomo = (mu, a): the
    ENABLE_TOMO_RECONS coupling of the NMS and of the pretcost to the
    tomographic projection a (None: Q12, the shipped build).  Returns a dict
    of the final state and the per-iteration histories."""
    nms_kw = nms_kw or {}
    v = np.zeros_like(w)
    pp = PhaseProblem(meas, D, Lam, None, nu, *geo)
    costs = [pret_cost(meas, D, Lam, zr, zi, pr, geo, tomo)]
    prim, dual, nms_its, sd_its = [], [], [], []
    it = 0
    w_prev, v_prev = w, v
    for it in range(pret_iters):
        w_prev, v_prev = w, v
        zr, zi, nit = nms_pass(zr, zi, w - v, nu, pr, order=order, tomo=tomo, **nms_kw)
        nms_its.append(nit)
        wold = w.copy()
        pp.v = transmittance(zr, zi) + v
        w, om, n, _ = pp.sd_run(om, w, sd_iters, sd_thresh)
        sd_its.append(n)
        t = transmittance(zr, zi) - w
        v = v + t
        pr_res = np.sqrt(np.sum(np.abs(t) ** 2) / t.size)
        du_res = np.sqrt(np.sum(np.abs(w - wold) ** 2) / np.sum(np.abs(v) ** 2))
        prim.append(pr_res)
        dual.append(du_res)
        costs.append(pret_cost(meas, D, Lam, zr, zi, pr, geo, tomo))
        if du_res < pret_thresh and it > 1:
            it += 1
            break
    else:
        it = pret_iters
    return dict(zr=zr, zi=zi, w=w, v=v, om=om, iters=it, costs=costs, primal=prim, dual=dual,
                nms_iters=nms_its, sd_iters=sd_its, w_prev=w_prev, v_prev=v_prev)
