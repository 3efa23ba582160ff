"""ctypes binding of phasecon_amd/libpcmbir4d.so (include/pcmbir_hip.h): the
Fresnel transform and the batched steepest-descent phase retrieval on the GPU.
There is no fallback: a missing library raises."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpcmbir4d.so")

_SIGS = {
    "pcmbir_hip_last_error": (C.c_char_p, []),
    "pcmbir_hip_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_void_p]),
    "pcmbir_hip_destroy": (None, [C.c_void_p]),
    "pcmbir_hip_fresnel": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pcmbir_hip_set_phase_problem": (C.c_int, [C.c_void_p] + [C.c_void_p] * 6 + [C.c_double]),
    "pcmbir_hip_sd_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_void_p, C.c_void_p]),
    "pcmbir_hip_get_phase_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pcmbir_hip_phase_cost": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pcmbir_hip_fft_time": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pcmbir_hip_set_nms_prior": (C.c_int, [C.c_void_p] + [C.c_double] * 5 + [C.c_int32] +
                                 [C.c_double] * 2),
    "pcmbir_hip_nms_pass": (C.c_int, [C.c_void_p] * 5),
    "pcmbir_hip_estimate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                      C.c_int32, C.c_double] + [C.c_void_p] * 6),
    "pcmbir_hip_get_dual": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pcmbir_hip_set_tomo_target": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
    "pcmbir_hip_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "pcmbir_hip_paganin": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                     C.c_void_p, C.c_void_p]),
}

_LIB = None

PLANCKS_CONSTANT = 4.135668e-19  # keV s, the reference's value (PCMBIR_3D/XT_Constants.h:84)
LIGHT_SPEED = 299792458e6        # um/s (PCMBIR_3D/XT_Constants.h:85)


def wavelength(light_energy_kev):
    """X-ray wavelength in um for the photon energy in keV as the reference
    computes it (PCMBIR_3D/XT_Init.c:258), including its Planck constant."""
    return PLANCKS_CONSTANT * LIGHT_SPEED / light_energy_kev


def fresnel_window(rows, cols, d_rows, d_cols, sigma):
    """Gaussian frequency window exp(-(u^2+v^2)/(2 sigma^2)) on the FFT grid
    (create_FresnelTranWindow, PCMBIR_3D/XT_Init.c:210-231)."""
    i = np.arange(rows, dtype=np.float64)
    j = np.arange(cols, dtype=np.float64)
    u = np.where(i >= rows // 2, -(rows - i), i) / (rows * d_rows)
    v = np.where(j >= cols // 2, -(cols - j), j) / (cols * d_cols)
    return np.exp(-(u[:, None] ** 2 + v[None, :] ** 2) / (2 * sigma * sigma))


def exported_names():
    return list(_SIGS)


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m phasecon_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {load().pcmbir_hip_last_error().decode()}")


def _c64(a, shape):
    """complex array -> contiguous interleaved float32 pairs"""
    a = np.asarray(a).reshape(shape)
    out = np.empty(shape + (2,), np.float32)
    out[..., 0] = a.real
    out[..., 1] = a.imag
    return out


def _cplx(f):
    return f[..., 0].astype(np.float64) + 1j * f[..., 1].astype(np.float64)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class PhaseRetrieval:
    """GPU context for n_views views of rows x cols (powers of two)."""

    def __init__(self, n_views, rows, cols, d_rows, d_cols, lam, dist, window=None):
        self.L = load()
        self.shape = (n_views, rows, cols)
        self.h = C.c_void_p()
        win = None if window is None else np.ascontiguousarray(window, np.float32)
        self._win = win
        _check(self.L.pcmbir_hip_create(C.byref(self.h), n_views, rows, cols, d_rows, d_cols,
                                        lam, dist, None if win is None else _p(win)),
               "pcmbir_hip_create")

    def close(self):
        if self.h:
            self.L.pcmbir_hip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def option(self, key, value):
        _check(self.L.pcmbir_hip_set_option(self.h, key.encode(), float(value)),
               "pcmbir_hip_set_option")

    def fresnel(self, x, herm=False):
        xi = _c64(x, self.shape)
        y = np.empty_like(xi)
        _check(self.L.pcmbir_hip_fresnel(self.h, int(herm), _p(xi), _p(y)), "pcmbir_hip_fresnel")
        return _cplx(y)

    def set_problem(self, y, D, lam, v, w, omega, nu):
        """omega=None: Omega becomes the phase projection of w."""
        self._keep = [_c64(a, self.shape) for a in (y, D, v, w)]
        oo = None if omega is None else _c64(omega, self.shape)
        lm = np.ascontiguousarray(np.broadcast_to(lam, self.shape), np.float32)
        self._keep += [lm, oo]
        yy, dd, vv, ww = self._keep[:4]
        _check(self.L.pcmbir_hip_set_phase_problem(self.h, _p(yy), _p(dd), _p(lm), _p(vv),
                                                   _p(ww), None if oo is None else _p(oo),
                                                   float(nu)),
               "pcmbir_hip_set_phase_problem")

    def sd_run(self, max_iters, thresh):
        it = np.zeros(self.shape[0], np.int32)
        upd = np.zeros(self.shape[0], np.float64)
        _check(self.L.pcmbir_hip_sd_run(self.h, int(max_iters), float(thresh), _p(it), _p(upd)),
               "pcmbir_hip_sd_run")
        return it, upd

    def state(self):
        w = np.empty(self.shape + (2,), np.float32)
        om = np.empty_like(w)
        _check(self.L.pcmbir_hip_get_phase_state(self.h, _p(w), _p(om)), "get_phase_state")
        return _cplx(w), _cplx(om)

    def cost(self):
        out = np.zeros(self.shape[0], np.float64)
        _check(self.L.pcmbir_hip_phase_cost(self.h, _p(out)), "pcmbir_hip_phase_cost")
        return out

    def fft_time(self):
        ms = C.c_double()
        n = C.c_int64()
        _check(self.L.pcmbir_hip_fft_time(self.h, C.byref(ms), C.byref(n)), "pcmbir_hip_fft_time")
        return ms.value, n.value

    def set_nms_prior(self, phase_sig, phase_c, mag_sig, mag_c, delta_over_beta, decorrelate,
                      p=1.2, nu=0.5):
        _check(self.L.pcmbir_hip_set_nms_prior(self.h, phase_sig, phase_c, mag_sig, mag_c,
                                               delta_over_beta, int(decorrelate), p, nu),
               "pcmbir_hip_set_nms_prior")

    def set_tomo_target(self, mu, a):
        """The ENABLE_TOMO_RECONS coupling of the NMS and the pretcost:
        mu/2 |z - a|^2 with a complex [n_views][rows][cols] (real: magnitude,
        imag: phase) in place of the projection-domain prior; a = None
        switches it off."""
        if a is None:
            _check(self.L.pcmbir_hip_set_tomo_target(self.h, 0.0, None, None),
                   "pcmbir_hip_set_tomo_target")
            return
        a = np.asarray(a).reshape(self.shape)
        self._tomo = (np.ascontiguousarray(a.real, np.float64),
                      np.ascontiguousarray(a.imag, np.float64))
        _check(self.L.pcmbir_hip_set_tomo_target(self.h, float(mu), _p(self._tomo[0]),
                                                 _p(self._tomo[1])),
               "pcmbir_hip_set_tomo_target")

    def nms_pass(self, zr, zi, b):
        zr = np.ascontiguousarray(zr, np.float64).reshape(self.shape).copy()
        zi = np.ascontiguousarray(zi, np.float64).reshape(self.shape).copy()
        bb = _c64(b, self.shape)
        tot = C.c_int64()
        _check(self.L.pcmbir_hip_nms_pass(self.h, _p(zr), _p(zi), _p(bb), C.byref(tot)),
               "pcmbir_hip_nms_pass")
        return zr, zi, tot.value

    def estimate(self, zr, zi, pret_iters, pret_thresh, sd_iters, sd_thresh):
        n = self.shape[0]
        zr = np.ascontiguousarray(zr, np.float64).reshape(self.shape).copy()
        zi = np.ascontiguousarray(zi, np.float64).reshape(self.shape).copy()
        it = np.zeros(n, np.int32)
        c0 = np.zeros(n)
        c1 = np.zeros(n)
        pr = np.zeros((n, max(pret_iters, 1)))
        du = np.zeros((n, max(pret_iters, 1)))
        cs = np.zeros((n, max(pret_iters, 1)))
        _check(self.L.pcmbir_hip_estimate(self.h, _p(zr), _p(zi), int(pret_iters),
                                          float(pret_thresh), int(sd_iters), float(sd_thresh),
                                          _p(it), _p(c0), _p(c1), _p(pr), _p(du), _p(cs)),
               "pcmbir_hip_estimate")
        return dict(zr=zr, zi=zi, iters=it, cost_init=c0, cost_final=c1, primal=pr, dual=du,
                    costs=cs)

    def dual(self):
        v = np.empty(self.shape + (2,), np.float32)
        _check(self.L.pcmbir_hip_get_dual(self.h, _p(v)), "pcmbir_hip_get_dual")
        return _cplx(v)

    def paganin(self, y, D, delta_over_beta, wavenumber):
        """Single-material Paganin retrieval: returns (z_real, z_imag), FP64."""
        yy, dd = _c64(y, self.shape), _c64(D, self.shape)
        zr = np.empty(self.shape, np.float64)
        zi = np.empty(self.shape, np.float64)
        _check(self.L.pcmbir_hip_paganin(self.h, _p(yy), _p(dd), float(delta_over_beta),
                                         float(wavenumber), _p(zr), _p(zi)),
               "pcmbir_hip_paganin")
        return zr, zi


_DROPIN = {
    "phcontomo_reconstruct": (C.c_int, [C.c_void_p, C.c_void_p] + [C.c_void_p] * 5 +
                              [C.c_int32] * 4 + [C.c_float] * 14 + [C.c_uint8, C.c_uint8,
                                                                    C.c_void_p]),
    "phcontomo_forward_project": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p] +
                                  [C.c_int32] * 3 + [C.c_float] * 5 + [C.c_void_p]),
    "phcontomo_forward_project_ex": (C.c_int, [C.c_void_p] * 4 + [C.c_int32] * 2 +
                                     [C.c_void_p] + [C.c_int32] * 3 + [C.c_float] * 4 +
                                     [C.c_uint64, C.c_void_p]),
}


def dropin_names():
    return list(_DROPIN)


def phcontomo_reconstruct(measurements, brights, angles, proj_times, recon_times, vox_wid,
                          rot_center, mag_sig_s, mag_sig_t, mag_c_s, mag_c_t, phase_sig_s,
                          phase_sig_t, phase_c_s, phase_c_t, convg_thresh, obj2det_dist,
                          light_energy, pag_regparam, restart=0, recon_type=1, log_path=None):
    """Calls the drop-in phcontomo_reconstruct (include/pcmbir4d.h) with the
    reference's arguments; measurements [proj_num][proj_rows][proj_cols],
    brights [proj_rows][proj_cols].  Results are the CWD files; returns the
    return code."""
    L = load()
    for name, (res, args) in _DROPIN.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    meas = np.ascontiguousarray(measurements, np.float32)
    br = np.ascontiguousarray(brights, np.float32)
    keep = [meas, br] + [np.ascontiguousarray(a, np.float32)
                         for a in (angles, proj_times, recon_times)]
    proj_num, proj_rows, proj_cols = meas.shape
    mag = C.c_void_p()
    ph = C.c_void_p()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen((log_path or "/dev/null").encode(), b"w")
    f = C.c_float
    try:
        rc = L.phcontomo_reconstruct(
            C.byref(mag), C.byref(ph), *[_p(a) for a in keep], proj_rows, proj_cols, proj_num,
            len(keep[4]) - 1, f(vox_wid), f(rot_center), f(mag_sig_s), f(mag_sig_t), f(mag_c_s),
            f(mag_c_t), f(phase_sig_s), f(phase_sig_t), f(phase_c_s), f(phase_c_t),
            f(convg_thresh), f(obj2det_dist), f(light_energy), f(pag_regparam),
            C.c_uint8(restart), C.c_uint8(recon_type), C.c_void_p(fp))
    finally:
        libc.fclose(C.c_void_p(fp))
    if rc == 0 and (mag.value or ph.value):
        raise RuntimeError("phcontomo_reconstruct returned object pointers; expected NULL")
    return rc


def _dropin():
    L = load()
    for name, (res, args) in _DROPIN.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def forward_project(mag, phase, angles, proj_rows, proj_cols, vox_wid, rot_center, obj2det_dist,
                    light_energy, noise_seed=0):
    """The phantom simulator on given phantoms [n_z][n_xy][n_xy]
    (phcontomo_forward_project_ex): returns (measurements [views][proj_rows]
    [proj_cols], brights [proj_rows][proj_cols]).  Writes SimMagProjs.bin,
    SimPhaseProjs.bin, mag_phantom.bin and phase_phantom.bin in the CWD, as
    the reference's simulator does.  noise_seed 0: no noise."""
    L = _dropin()
    mag = np.ascontiguousarray(mag, np.float32)
    phase = np.ascontiguousarray(phase, np.float32)
    ang = np.ascontiguousarray(angles, np.float32)
    n_z, n_xy = mag.shape[0], mag.shape[-1]
    meas = np.zeros((ang.size, proj_rows, proj_cols), np.float32)
    br = np.zeros((proj_rows, proj_cols), np.float32)
    f = C.c_float
    rc = L.phcontomo_forward_project_ex(_p(meas), _p(br), _p(mag), _p(phase), n_xy, n_z, _p(ang),
                                        proj_rows, proj_cols, ang.size, f(vox_wid),
                                        f(rot_center), f(obj2det_dist), f(light_energy),
                                        int(noise_seed), None)
    if rc != 0:
        raise RuntimeError("phcontomo_forward_project_ex failed")
    return meas, br


def phcontomo_forward_project(angles, proj_rows, proj_cols, vox_wid, rot_center, obj2det_dist,
                              light_energy, pag_regparam=0.0, log_path=None):
    """The drop-in phcontomo_forward_project (include/pcmbir4d.h) with the
    reference's arguments (phantoms from the files named by
    PHCONTOMO_MAG_PHANTOM / PHCONTOMO_PHASE_PHANTOM); returns (rc,
    measurements, brights)."""
    L = _dropin()
    ang = np.ascontiguousarray(angles, np.float32)
    meas = np.zeros((ang.size, proj_rows, proj_cols), np.float32)
    br = np.zeros((proj_rows, proj_cols), np.float32)
    pm, pb = C.c_void_p(meas.ctypes.data), C.c_void_p(br.ctypes.data)
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen((log_path or "/dev/null").encode(), b"w")
    f = C.c_float
    try:
        rc = L.phcontomo_forward_project(C.byref(pm), C.byref(pb), _p(ang), proj_rows, proj_cols,
                                         ang.size, f(vox_wid), f(rot_center), f(obj2det_dist),
                                         f(light_energy), f(pag_regparam), C.c_void_p(fp))
    finally:
        libc.fclose(C.c_void_p(fp))
    return rc, meas, br
