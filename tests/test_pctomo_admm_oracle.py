"""Pins of the oracle's PCMBIR ADMM head loop with the tomographic step
(oracle/pctomo_oracle.py; ENABLE_TOMO_RECONS is compiled out of the shipped
reference, so no reference build runs it).  CPU only.

* The dual update is the scaled-dual ADMM step of the constraint z = k A x:
  V' = V + k A x - z, after which the error sinogram is again Z - V - k A x,
  the primal residual the RMS of z - k A x and the dual residual
  |z - P| / |V'| (a flipped dual sign, a missing V or the wrong previous
  estimate breaks one of these).
* The coupled simplex search (Cost_NMS with ENABLE_TOMO_RECONS) with nu = 0
  is a quadratic with its minimum at a, and with b = exp(-a) both terms are
  minimised at a: the search lands there.  The coupled pretcost adds
  mu/2 |z - a|^2, by brute force over the pixels.
* Two head iterations on a small scan whose measurements come from known
  objects: the ICD keeps e = Z - V - k A x, the primal residual falls and
  the objects approach the truth.
"""
import numpy as np
import pytest

from oracle import pcmbir_oracle as PO
from oracle import pctomo_oracle as PT


def test_dual_update_is_the_scaled_admm_step():
    r = np.random.default_rng(0)
    shape = (3, 8, 4)
    z0, kax = r.normal(size=shape), r.normal(size=shape)
    for from_object in (False, True):
        s = PT.admm_init(z0, kax, from_object)
        assert np.array_equal(s["P"], z0)
        assert np.allclose(s["e"], s["Z"] - s["V"] - kax)
        if from_object:
            assert np.allclose(s["e"], 0.0)
        for _ in range(3):
            a = PT.admm_target(s)
            assert np.allclose(a, kax + s["V"])
            V, P = s["V"].copy(), s["P"].copy()
            znew = kax + 0.3 * r.normal(size=shape)
            primal, dual = PT.admm_update(s, znew)
            assert np.allclose(s["V"], V + kax - znew)
            assert np.allclose(s["e"], s["Z"] - s["V"] - kax)
            assert np.array_equal(s["Z"], znew) and np.array_equal(s["P"], znew)
            assert primal == pytest.approx(np.sqrt(np.mean((znew - kax) ** 2)), rel=1e-12)
            assert dual == pytest.approx(np.linalg.norm(znew - P) / np.linalg.norm(s["V"]),
                                         rel=1e-12)


def _prior():
    return PO.NmsPrior(2e-2, 1e-2, 2e-2, 1e-2, 50.0, True)


def test_coupled_search_lands_on_the_tomographic_projection():
    r = np.random.default_rng(1)
    pr = _prior()
    nb = np.zeros((3, 3))
    fl = np.ones((3, 3), bool)
    for nu in (0.0, 0.5):
        for _ in range(10):
            ar, ai = 0.02 + 0.05 * r.random(), 0.2 + 0.5 * r.random()
            b = PO.transmittance(ar, ai)
            zr, zi, it = PO.nms_search(1.3 * ar, 0.8 * ai, b.real, b.imag, nu, nb, nb, fl, pr,
                                       tomo=(0.5, ar, ai))
            assert abs(zr - ar) < 2e-3 * ar and abs(zi - ai) < 2e-3 * ai, (nu, ar, ai, zr, zi)
            assert it > 2


def test_coupled_pretcost_by_brute_force():
    r = np.random.default_rng(2)
    rows, cols = 8, 16
    lam = PO.wavelength(17.0)
    geo = (1.0, 1.0, lam, 1e5, None)
    zr, zi = 0.02 + 0.05 * r.random((rows, cols)), 0.2 + 0.5 * r.random((rows, cols))
    a = (zr + 0.01 * r.normal(size=zr.shape)) + 1j * (zi + 0.03 * r.normal(size=zi.shape))
    meas = np.abs(PO.fresnel(PO.transmittance(zr, zi), *geo)) + 0j
    D = np.ones((rows, cols), complex)
    Lam = 1.0 + r.random((rows, cols))
    mu = 0.7
    fwd = PO.pret_cost(meas, D, Lam, zr, zi, _prior(), geo, tomo=(mu, np.zeros_like(a)))
    fwd -= sum(mu / 2 * (zr[j, k] ** 2 + zi[j, k] ** 2) for j in range(rows)
               for k in range(cols))
    got = PO.pret_cost(meas, D, Lam, zr, zi, _prior(), geo, tomo=(mu, a))
    want = fwd + sum(mu / 2 * ((zr[j, k] - a[j, k].real) ** 2 + (zi[j, k] - a[j, k].imag) ** 2)
                     for j in range(rows) for k in range(cols))
    assert got == pytest.approx(want, rel=1e-12)
    assert abs(fwd) < 1e-12  # the measurements are exact at z


def tomo_problem(name="tiny3d", seed=0, wavenumber=30.0, dob=50.0, mu=0.5, alpha=1.5):
    """Known objects, their projections z = k A x, noise-free amplitude
    measurements |H exp(-z)| (D = 1, unit weights), and the start: objects
    zero, z the truth blurred by 20 %."""
    from oracle.oracle import Oracle, pc_prior
    from phasecon_amd.synth import make_scan
    scan = make_scan(name, seed=seed)
    o = Oracle(scan)
    Pt = pc_prior((0.02, 1e-3, 0.01, 0.01), (0.005, 1e-3, 0.01, 0.01), 1, o.g.delta_recon,
                  wavenumber, delta_over_beta=dob, mu=mu, alpha=alpha)
    r = np.random.default_rng(seed + 7)
    xd = 0.004 + 0.002 * r.random(o.shape_x)
    xb = xd / dob
    zd = wavenumber * o.forward_project(xd)
    zb = wavenumber * o.forward_project(xb)
    lam = PO.wavelength(17.0)
    geo = (1.0, 1.0, lam, 2e4, None)
    meas = np.abs(PO.fresnel(PO.transmittance(zb, zd), *geo)) + 0j
    D = np.ones_like(meas)
    Lam = np.ones(meas.shape)
    return dict(scan=scan, o=o, Pt=Pt, xd=xd, xb=xb, zd=zd, zb=zb, meas=meas, D=D, Lam=Lam,
                geo=geo, z0d=0.8 * zd, z0b=0.8 * zb)


def test_head_loop_keeps_errors_and_converges():
    p = tomo_problem()
    o, Pt = p["o"], p["Pt"]
    hl = PT.HeadLoop(o, Pt, p["meas"], p["D"], p["Lam"], _prior(), p["geo"], 0.5,
                     np.zeros(o.shape_x), np.zeros(o.shape_x), p["z0d"], p["z0b"])
    T = max(o.g.N_x, o.g.N_y) // 2
    errs, prim = [], []
    for head in range(2):
        hl.icd(10, 1, T, [31 + 10 * head + i for i in range(10)])
        k = Pt.wavenumber
        kd, kb = k * o.forward_project(hl.xd), k * o.forward_project(hl.xb)
        assert np.abs(hl.sd["e"] - (hl.sd["Z"] - hl.sd["V"] - kd)).max() < 1e-10
        assert np.abs(hl.sb["e"] - (hl.sb["Z"] - hl.sb["V"] - kb)).max() < 1e-10
        zd, zb, out = hl.retrieval(2, 0.0, 2, 0.0)
        res = hl.update(zd, zb)
        prim.append(res[[0, 2]])
        errs.append(np.abs(hl.xd - p["xd"]).max() / p["xd"].max())
        print(f"head {head + 1}: residuals {res}, delta error {errs[-1]:.3g}", flush=True)
    assert np.all(prim[1] < prim[0])
    assert errs[1] < errs[0] < 0.5
