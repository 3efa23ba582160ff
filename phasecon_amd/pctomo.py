"""PCMBIR with its tomographic reconstruction switched on (the
ENABLE_TOMO_RECONS head loop of PCMBIR_3D/XT_ICD_update.c:1020-1151, compiled
out of the shipped reference) on the GPU.

The head loop alternates the coupled (delta, beta) ICD of the tomographic
mode (mbir.TomoStage, libmbir4d.so) on e = Z - V - k A x, the proximal point
a = k A x + V, the phase retrieval of every view coupled to a
(pcmbir.PhaseRetrieval with set_tomo_target, libpcmbir4d.so), and the dual
update with its residuals.  Both libraries keep their state on the device;
the per-view projections cross between them as FP64 host arrays once per
head iteration.  Sinograms are [N_p][N_r][N_t] (the phase retrieval's
[n_views][rows][cols]); z's real part is the magnitude (beta), its
imaginary part the phase (delta).
"""
from __future__ import annotations

import numpy as np

from .mbir import TomoStage
from .pcmbir import PhaseRetrieval


class TomoADMM:
    """scan: the geometry of the tomographic grid (a synth-style scan dict;
    its projections and weights are not used); prior: mbir.tomo_prior(...)
    (its mu is the ADMM's); meas, D, Lam: the phase problem of every view
    [N_p][N_r][N_t] (measurements, illumination factor, weights); geo =
    (d_rows, d_cols, wavelength, distance, window); nu: the retrieval's
    penalty; nms: the projection-domain prior of pcmbir_hip_set_nms_prior
    (phase_sig, phase_c, mag_sig, mag_c, delta_over_beta, decorrelate), kept
    for the uncoupled parts of the retrieval's set-up.  tile: the sweep's
    tile edge; 8 fits columns of up to 256 voxels in LDS and stays stable on
    the PCMBIR workload's 256^3 object, where tile 4 diverges."""

    def __init__(self, scan, prior, meas, D, Lam, geo, nu, nms, tile=8):
        self.prior = prior
        self.st = TomoStage(scan, prior, tile=tile)
        self.tile = tile
        n_p, n_r, n_t = self.st.shape_e
        d_rows, d_cols, lam, dist, win = geo
        self.pr = PhaseRetrieval(n_p, n_r, n_t, d_rows, d_cols, lam, dist, win)
        self.pr.set_nms_prior(**nms, nu=nu)
        self.meas = np.asarray(meas).reshape(n_p, n_r, n_t)
        self.D = np.asarray(D).reshape(n_p, n_r, n_t)
        self.Lam = np.broadcast_to(Lam, (n_p, n_r, n_t))
        self.nu = nu
        self.rise_tol = 1e-6  # relative ICD cost rise over a head iteration that stops the run
        self.zd = self.zb = None
        self.equits = 0  # ICD equits run, over all head iterations

    def start(self, xd, xb, zd, zb, from_object=False):
        """Objects and the projection estimate z (e.g. Paganin's), as
        initObject / initErrorSinogam leave them (XT_ICD_update.c:698-764):
        Z = z, or k A x of the objects with from_object (recon_type 2); V =
        0; w = exp(-Z), Omega its phase projection, the retrieval's dual 0."""
        self.st.set_objects(xd, xb)
        self.st.admm_init(zd, zb, from_object)
        if from_object:
            zd, zb = (a.astype(np.float64) for a in self.st.admm_state(0))
        self.zd = np.array(zd, np.float64).reshape(self.st.shape_e)
        self.zb = np.array(zb, np.float64).reshape(self.st.shape_e)
        e = np.exp(-self.zb)
        w = e * np.cos(-self.zd) + 1j * e * np.sin(-self.zd)
        self.pr.set_problem(self.meas, self.D, self.Lam, np.zeros_like(w), w, None, self.nu)
        self.equits = 0

    def head_iteration(self, n_icd, pret_iters, pret_thresh, sd_iters, sd_thresh,
                       seed=761521):
        """One head iteration: n_icd ICD equits numbered from 1, as each head
        iteration restarts Iter (XT_ICD_update.c:1050; zero skipping starts
        at Iter 2), the j-th equit of the run seeded seed + j; the coupled
        phase retrieval; the dual update.  Returns the ICD costs before and
        after, the retrieval's per-view results and the residuals [beta
        primal, beta dual, delta primal, delta dual]."""
        c0 = self.st.tomo_cost()[0]
        for i in range(n_icd):
            self.st.equits(1, i + 1, seed + self.equits)
            self.equits += 1
        c1 = self.st.tomo_cost()[0]
        if not c1 <= c0 * (1.0 + self.rise_tol):
            # the reference stops on a cost increase (XT_ICD_update.c:1171);
            # here a rise means the parallel sweep overshot (too many columns
            # updated at once on one sinogram: take a larger tile)
            raise RuntimeError(f"tomographic ICD cost rose from {c0:.10g} to {c1:.10g} over "
                               f"{n_icd} equits (tile {self.tile})")
        ad, ab = self.st.admm_target()
        self.pr.set_tomo_target(self.prior.mu, ab + 1j * ad)
        out = self.pr.estimate(self.zb, self.zd, pret_iters, pret_thresh, sd_iters, sd_thresh)
        self.zb, self.zd = out["zr"], out["zi"]
        res = self.st.admm_update(self.zd, self.zb)
        return dict(icd_cost=(c0, c1), retrieval=out, residuals=res)

    def objects(self):
        return self.st.get_objects()

    def close(self):
        self.st.close()
        self.pr.close()
