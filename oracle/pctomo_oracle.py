"""FP64 restatement of PCMBIR's ADMM head loop with the tomographic
reconstruction switched on (ENABLE_TOMO_RECONS, compiled out of the shipped
reference, Q12) -- TEST INFRASTRUCTURE ONLY: imported by tests/ to test
against, never by the product.

Paths are relative to the reference's Source_Code/PCMBIR_3D/.  The head loop
(XT_ICD_update.c:1037-1151) alternates
  1. ICD equits of the two objects (delta, beta) on e = Z - V - k A x
     (Oracle.pc_equit, the coupled ICD of XT_VoxUpdate.c:23-80 and
     XT_Prior.c:137-239);
  2. the proximal point a = Z - e = k A x + V of every view
     (XT_ICD_update.c:1089-1093);
  3. the phase retrieval of every view with the tomographic coupling
     mu/2 |z - a|^2 in its simplex search and pretcost
     (pcmbir_oracle.estimate_complex_projection with tomo);
  4. the dual update e1 = z - a, V' = -e1, e' = e1 + V - V', the primal and
     dual residuals, P = Z = z (XT_ICD_update.c:1095-1140).
Sinograms are per object [N_p][N_r][N_t]; the phase retrieval's z has the
magnitude (beta) as its real and the phase (delta) as its imaginary part.

Pinned in tests/test_pctomo_admm_oracle.py by what ADMM fixes rather than by
re-reading the reference: after every update e = Z - V - k A x again, V' = V
+ k A x - z (the scaled dual of the constraint z = k A x: a flipped sign or a
missing V breaks both), the primal residual is the RMS of z - k A x and the
dual residual |z - P| / |V'|; the coupled simplex search with nu = 0 lands on
a; the pretcost's coupling term by brute force.
"""
from __future__ import annotations

import numpy as np

from oracle import pcmbir_oracle as PO


def admm_init(z, kax, from_object):
    """initErrorSinogam with ENABLE_TOMO_RECONS for one object
    (XT_ICD_update.c:698-760): P = z; Z = z, or k A x with from_object
    (recon_type 2, :733-735); V = 0; e = Z - V - k A x (:757-758)."""
    P = np.array(z, np.float64)
    Z = np.array(kax if from_object else z, np.float64)
    V = np.zeros_like(Z)
    e = Z - V - kax
    return dict(Z=Z, V=V, P=P, e=e)


def admm_target(s):
    """a = Z - e (XT_ICD_update.c:1089-1093: -(e - Z))."""
    return s["Z"] - s["e"]


def admm_update(s, znew):
    """XT_ICD_update.c:1095-1140 for one object, with a = Z - e recomputed
    (unchanged between the target and the update): returns (primal, dual)."""
    a = admm_target(s)
    e1 = znew - a
    vold = s["V"]
    vnew = -e1
    enew = e1 + vold - vnew
    primal = np.sqrt(np.sum((vnew + enew) ** 2) / enew.size)
    dual = np.sqrt(np.sum((znew - s["P"]) ** 2) / np.sum(vnew ** 2))
    s.update(V=vnew, e=enew, P=np.array(znew, np.float64), Z=np.array(znew, np.float64))
    return primal, dual


class HeadLoop:
    """The head loop on one scan: o an oracle.Oracle of the scan's geometry,
    Pt its oracle.PcPrior, views the per-view phase problem (meas, D, Lam as
    [N_p][N_r][N_t] complex / real arrays), nms the pcmbir_oracle.NmsPrior
    (unused by the coupled search but kept for its shape), geo the Fresnel
    geometry (d_rows, d_cols, lam, dist, window), nu the retrieval's
    penalty."""

    def __init__(self, o, Pt, meas, D, Lam, nms, geo, nu, xd, xb, zd, zb, from_object=False):
        self.o, self.Pt = o, Pt
        self.meas, self.D, self.Lam = meas, D, Lam
        self.nms, self.geo, self.nu = nms, geo, nu
        self.xd = np.array(xd, np.float64)
        self.xb = np.array(xb, np.float64)
        k = Pt.wavenumber
        kd = k * o.forward_project(self.xd)
        kb = k * o.forward_project(self.xb)
        self.sd = admm_init(zd, kd, from_object)
        self.sb = admm_init(zb, kb, from_object)
        # w = exp(-Z) and Omega its phase projection (XT_ICD_update.c:759-764)
        self.w = PO.transmittance(self.sb["Z"], self.sd["Z"])
        self.om = np.empty_like(self.w)
        for v in range(self.w.shape[0]):
            pp = PO.PhaseProblem(meas[v], D[v], Lam[v], None, nu, *geo)
            self.om[v] = pp.phase_projection(self.w[v])

    def icd(self, n, first_iter, T, seeds):
        """n equits of the coupled ICD on the maintained error sinograms."""
        for i in range(n):
            self.o.pc_equit(self.Pt, self.xd, self.xb, self.sd["e"], self.sb["e"],
                            first_iter + i, T, seeds[i])

    def retrieval(self, pret_iters, pret_thresh, sd_iters, sd_thresh, order="colour"):
        """The coupled phase retrieval of every view towards a; returns the
        new estimates (z_delta, z_beta) and the per-view results."""
        ad, ab = admm_target(self.sd), admm_target(self.sb)
        zd = np.empty_like(ad)
        zb = np.empty_like(ab)
        out = []
        for v in range(ad.shape[0]):
            r = PO.estimate_complex_projection(
                self.meas[v], self.om[v], self.D[v], self.sb["Z"][v], self.sd["Z"][v],
                self.Lam[v], self.w[v], self.nms, self.geo, self.nu, pret_iters, pret_thresh,
                sd_iters, sd_thresh, order=order, tomo=(self.Pt.mu, ab[v] + 1j * ad[v]))
            zb[v], zd[v] = r["zr"], r["zi"]
            self.w[v], self.om[v] = r["w"], r["om"]
            out.append(r)
        return zd, zb, out

    def update(self, zd, zb):
        """Dual update of both objects: [beta primal, beta dual, delta
        primal, delta dual], the reference's order."""
        pb, db = admm_update(self.sb, zb)
        pd, dd = admm_update(self.sd, zd)
        return np.array([pb, db, pd, dd])
