"""Python mirror of the MBIR engine's interface.

``reconstruct`` calls the native drop-in (include/mbir4d.h, via
reconstruct_ex for the extra options) with the reference's argument list;
``Stage`` drives one multi-resolution stage through the HIP C-ABI
(include/mbir_hip.h) for tests, the benchmark and the warm start.

Array conventions follow the reference: projections/weights in file layout
[view][row][col]; objects [frame][z][y][x]; sinogram-shaped stage arrays in
the internal [view][col r][row t] layout (MBIR_4D/XT_Init.c:294-301).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _lib

NB = 65


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def multires_schedule(proj_rows, proj_cols):
    L = _lib.load()
    mxy = np.zeros(8, np.int32)
    mz = np.zeros(8, np.int32)
    n = L.mbir_multires_schedule(C.c_int32(proj_rows), C.c_int32(proj_cols), _p(mxy), _p(mz))
    return [(int(mxy[i]), int(mz[i])) for i in range(n)]


def view_map(proj_times, recon_times):
    L = _lib.load()
    t = np.ascontiguousarray(proj_times, np.float32)
    r = np.ascontiguousarray(recon_times, np.float32)
    ptr = np.zeros(len(r), np.int32)
    views = np.zeros(max(len(t), 1), np.int32)
    n = L.mbir_view_map(C.c_int32(len(t)), _p(t), C.c_int32(len(r) - 1), _p(r), _p(ptr), _p(views))
    if n < 0:
        raise ValueError("projection or reconstruction times decrease")
    return ptr, views[:n].copy()


def interlaced_views(K, n_theta, n_views, proj_start=0, min_acq_time=0.0, rot_speed=0.0):
    """Time-interlaced view angles and times (mbir_interlaced_views, the
    APS driver's gen_interlaced_views_0_to_Inf)."""
    L = _lib.load()
    views = np.zeros(n_views, np.float32)
    times = np.zeros(n_views, np.float32)
    n = L.mbir_interlaced_views(_p(views), _p(times), int(K), int(n_theta), int(proj_start),
                                int(n_views), float(min_acq_time), float(rot_speed))
    if n < 0:
        raise ValueError("bad interlaced-view arguments")
    return views, times


def phantom_project(mag, phase, angles, proj_rows, proj_cols, vox_wid, rot_center, wavenumber):
    """The phase-contrast simulator's phantom projections on the GPU
    (mbir_hip_phantom_project): (real, imag) [views][n_z][n_xy]."""
    L = _lib.load()
    mag = np.ascontiguousarray(mag, np.float32)
    phase = np.ascontiguousarray(phase, np.float32)
    ang = np.ascontiguousarray(angles, np.float32)
    n_z, n_xy = mag.shape[0], mag.shape[-1]
    pr = np.zeros((ang.size, n_z, n_xy), np.float32)
    pi = np.zeros_like(pr)
    _lib.check(L.mbir_hip_phantom_project(_p(mag), _p(phase), n_xy, n_z, ang.size, _p(ang),
                                          proj_rows, proj_cols, float(vox_wid), float(rot_center),
                                          float(wavenumber), _p(pr), _p(pi)), "phantom_project")
    return pr, pi


def std_downsample(proj, weights, ratio, bh_coef=0.0):
    """Projections and weights [views][rows][cols] from normalised data
    [views][rows*ratio][cols*ratio] (mbir_hip_std_downsample, the
    standard-format driver's read_ProjWeightData)."""
    L = _lib.load()
    proj = np.ascontiguousarray(proj, np.float32)
    weights = np.ascontiguousarray(weights, np.float32)
    n, R, Cc = proj.shape
    if R % ratio or Cc % ratio or weights.shape != proj.shape:
        raise ValueError("shapes must agree and divide by ratio")
    rows, cols = R // ratio, Cc // ratio
    out_p = np.zeros((n, rows, cols), np.float32)
    out_w = np.zeros((n, rows, cols), np.float32)
    _lib.check(L.mbir_hip_std_downsample(_p(proj), _p(weights), n, rows, cols, int(ratio),
                                         float(bh_coef), _p(out_p), _p(out_w)), "std_downsample")
    return out_p, out_w


def aps_normalize(proj, white, dark, ratio, bh_coef=0.5):
    """Projections and weights [views][rows][cols] from raw uint16 counts
    (mbir_hip_aps_normalize, the APS driver's read_data flat-field step):
    proj [views][rows*ratio][cols*ratio], white / dark frames of the same
    frame shape."""
    L = _lib.load()
    proj = np.ascontiguousarray(proj, np.uint16)
    white = np.ascontiguousarray(white, np.uint16)
    dark = np.ascontiguousarray(dark, np.uint16)
    n, R, Cc = proj.shape
    if R % ratio or Cc % ratio or white.shape[1:] != (R, Cc) or dark.shape[1:] != (R, Cc):
        raise ValueError("frame shapes must agree and divide by ratio")
    rows, cols = R // ratio, Cc // ratio
    out_p = np.zeros((n, rows, cols), np.float32)
    out_w = np.zeros((n, rows, cols), np.float32)
    _lib.check(L.mbir_hip_aps_normalize(_p(proj), _p(white), _p(dark), n, white.shape[0],
                                        dark.shape[0], rows, cols, int(ratio), float(bh_coef),
                                        _p(out_p), _p(out_w)), "aps_normalize")
    return out_p, out_w


class Stage:
    """One multi-resolution stage of one scan, resident on the GPU."""

    def __init__(self, scan, mult_xy=1, mult_z=1, quad_convex=1, huber_delta=0.5,
                 huber_T=4.0, p_s=None, p_t=None, tile=16, alpha=None, var_est=1.0,
                 afp16=False, full_mask=False):
        L = self.L = _lib.load()
        s = self.scan = scan
        N_p, N_t, N_r = s["proj_num"], s["proj_rows"], s["proj_cols"]
        self.frame_ptr, self.frame_views = view_map(s["proj_times"], s["recon_times"])
        self.frame_views = np.ascontiguousarray(np.resize(self.frame_views, max(len(self.frame_views), 1)))
        g = self.g = _lib.HipGeom()
        f = C.c_float
        rc = L.mbir_stage_geometry(
            C.byref(g), C.c_int32(N_t), C.c_int32(N_r), C.c_int32(N_p), C.c_int32(s["recon_num"]),
            f(s["vox_wid"]), f(s["rot_center"]), f(s["sig_s"]), f(s["sig_t"]), f(s["c_s"]),
            f(s["c_t"]), C.c_int32(mult_xy), C.c_int32(mult_z), C.c_int32(quad_convex),
            f(huber_delta), f(huber_T), _p(self.frame_ptr),
            C.c_double(s["p_s"] if p_s is None else p_s),
            C.c_double(s["p_t"] if p_t is None else p_t))
        if rc:
            raise ValueError("mbir_stage_geometry rejected the geometry")
        g.var_est = var_est
        self.angles = np.ascontiguousarray(s["angles"], np.float32)
        self.cos = np.zeros(N_p, np.float64)
        self.sin = np.zeros(N_p, np.float64)
        L.mbir_sincos(C.c_int32(N_p), _p(self.angles), _p(self.cos), _p(self.sin))
        self.H = np.zeros((N_p, NB), np.float32)
        L.mbir_detector_response(C.byref(g), _p(self.angles), _p(self.H))
        self.y = self._internal(s["projections"])
        self.w = (self._internal(s["weights"]) / np.float32(var_est)).astype(np.float32)
        self.mask = np.zeros((g.N_y, g.N_x), np.uint8)
        # full_mask: every column updated, as PCMBIR's radius_obj = vox_wid
        # proj_cols (PCMBIR_3D/XT_Init.c:303) circumscribes the grid
        if full_mask:
            self.mask[:] = 1
        else:
            L.mbir_update_mask(C.byref(g), C.c_double(float(s["vox_wid"]) * N_r / 2),
                               _p(self.mask))
        h = C.c_void_p()
        _lib.check(L.mbir_hip_create(C.byref(h), C.byref(g), _p(self.H), _p(self.cos),
                                     _p(self.sin), _p(self.frame_ptr), _p(self.frame_views),
                                     _p(self.y), _p(self.w), None, _p(self.mask)),
                   "mbir_hip_create")
        self.h = h
        self.option("tile", tile)
        self.alpha = float(g.alpha) if alpha is None else float(alpha)
        self.alpha0 = self.alpha  # the configured value (the guard may lower alpha)
        if alpha is not None:
            self.option("alpha", alpha)
        self.afp16 = bool(afp16)
        if afp16:
            self.option("afp16", 1)

    def _internal(self, a):
        s = self.scan
        N_p, N_t, N_r = s["proj_num"], s["proj_rows"], s["proj_cols"]
        a = np.ascontiguousarray(a, np.float32).reshape(N_p, N_t, N_r)
        out = np.zeros((N_p, N_r, N_t), np.float32)
        self.L.mbir_sino_to_internal(C.c_int32(N_p), C.c_int32(N_t), C.c_int32(N_r), _p(a), _p(out))
        return out

    # -- sizes
    @property
    def shape_x(self):
        g = self.g
        return (g.N_time, g.N_z, g.N_y, g.N_x)

    @property
    def shape_e(self):
        g = self.g
        return (g.N_p, g.N_r, g.N_t)

    @property
    def W(self):
        return 2 * self.g.mult_xy + 1

    # -- calls
    def option(self, key, value):
        _lib.check(self.L.mbir_hip_set_option(self.h, key.encode(), float(value)), "set_option")
        if key == "alpha":
            self.alpha = float(value)

    def set_object(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(self.shape_x)
        _lib.check(self.L.mbir_hip_set_object(self.h, _p(x)), "set_object")

    def get_object(self):
        x = np.zeros(self.shape_x, np.float32)
        _lib.check(self.L.mbir_hip_get_object(self.h, _p(x)), "get_object")
        return x

    def upsample_from(self, coarse, masked=True):
        """This stage's object := the coarse stage's, up-sampled on the
        device (mbir_hip_upsample_object); masked zeroes voxels outside this
        stage's update mask."""
        _lib.check(self.L.mbir_hip_upsample_object(self.h, coarse.h, int(bool(masked))),
                   "mbir_hip_upsample_object")

    def build_amatrix(self):
        _lib.check(self.L.mbir_hip_build_amatrix(self.h), "build_amatrix")

    def get_amatrix(self):
        g = self.g
        n = g.N_x * g.N_y
        ns = int(self.frame_ptr[-1])
        first = np.zeros((n, max(ns, 1)), np.int32)
        val = np.zeros((n, max(ns, 1), self.W), np.float32)
        _lib.check(self.L.mbir_hip_get_amatrix(self.h, _p(first), _p(val)), "get_amatrix")
        return first[:, :ns], val[:, :ns]

    def bp_normalized(self):
        """Per voxel sum(w a (y - d)) / sum(w a^2) (mbir_hip_bp_normalized),
        [N_time][N_z][N_y][N_x]; needs build_amatrix()."""
        out = np.zeros(self.shape_x, np.float32)
        _lib.check(self.L.mbir_hip_bp_normalized(self.h, _p(out)), "bp_normalized")
        return out

    def fbp(self):
        """Filtered back-projection (mbir_hip_fbp), [N_time][N_z][N_y][N_x],
        scaled so an object uniform inside the mask keeps its value; needs
        build_amatrix()."""
        out = np.zeros(self.shape_x, np.float32)
        _lib.check(self.L.mbir_hip_fbp(self.h, _p(out)), "fbp")
        return out

    def init_error(self):
        _lib.check(self.L.mbir_hip_init_error(self.h), "init_error")

    def get_error(self):
        e = np.zeros(self.shape_e, np.float32)
        _lib.check(self.L.mbir_hip_get_error(self.h, _p(e)), "get_error")
        return e

    def set_error(self, e):
        e = np.ascontiguousarray(e, np.float32).reshape(self.shape_e)
        _lib.check(self.L.mbir_hip_set_error(self.h, _p(e)), "set_error")

    def forward_project(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(self.shape_x)
        ax = np.zeros(self.shape_e, np.float32)
        _lib.check(self.L.mbir_hip_forward_project(self.h, _p(x), _p(ax)), "forward_project")
        return ax

    def backproject(self, v):
        v = np.ascontiguousarray(v, np.float32).reshape(self.shape_e)
        bt = np.zeros(self.shape_x, np.float32)
        _lib.check(self.L.mbir_hip_backproject(self.h, _p(v), _p(bt)), "backproject")
        return bt

    def icd(self, n, first_iter, seed=761521):
        st = np.zeros(4, np.float64)
        _lib.check(self.L.mbir_hip_icd_equits(self.h, n, first_iter, seed, _p(st)), "icd_equits")
        return st

    def icd_graph(self, n, first_iter, seed=761521):
        """n equits through the captured HIP graph (mbir_hip_icd_graph);
        asynchronous."""
        _lib.check(self.L.mbir_hip_icd_graph(self.h, int(n), int(first_iter), int(seed)),
                   "icd_graph")

    def nhicd(self, passes, fraction, iteration, seed=761521, stats=False):
        """Non-homogeneous passes over the top `fraction` of each tile's
        columns by filtered update magnitude (mbir_hip_nhicd); needs
        option("nhicd", 1).  Asynchronous unless stats, which returns the
        stats of icd()."""
        st = np.zeros(4, np.float64) if stats else None
        _lib.check(self.L.mbir_hip_nhicd(self.h, int(passes), float(fraction), int(iteration),
                                         int(seed), _p(st) if stats else None), "nhicd")
        return st

    def magnitudes(self, filtered=False):
        """The per-column |dx| map [N_time][N_y][N_x] or its VSC filtering."""
        out = np.zeros((self.shape_x[0],) + self.shape_x[2:], np.float32)
        _lib.check(self.L.mbir_hip_get_magnitudes(self.h, int(bool(filtered)), _p(out)),
                   "get_magnitudes")
        return out

    def offset_update(self, mode):
        """One ring-offset update (mbir_hip_offset_update), mode 1, 2 or 3."""
        _lib.check(self.L.mbir_hip_offset_update(self.h, int(mode)), "offset_update")

    def get_offset(self):
        d = np.zeros((self.g.N_r, self.g.N_t), np.float32)
        _lib.check(self.L.mbir_hip_get_offset(self.h, _p(d)), "get_offset")
        return d

    def variance_update(self):
        """The noise-variance re-estimate; returns the new variance."""
        v = C.c_double()
        _lib.check(self.L.mbir_hip_variance_update(self.h, C.byref(v)), "variance_update")
        return v.value

    def get_weight(self):
        w = np.zeros(self.shape_e, np.float32)
        _lib.check(self.L.mbir_hip_get_weight(self.h, _p(w)), "get_weight")
        return w

    def cost(self, fresh=False, estimate=False):
        """[total, data, prior] FP64 MAP cost; fresh: e recomputed from x;
        estimate: maintained e with FP32 prior potentials (stopping tests)."""
        out = np.zeros(3, np.float64)
        mode = 2 if estimate else int(bool(fresh))
        _lib.check(self.L.mbir_hip_cost(self.h, mode, _p(out)), "cost")
        return out

    def thetas(self, tzyx):
        q = np.ascontiguousarray(tzyx, np.int32).reshape(-1, 4)
        t1 = np.zeros(len(q), np.float32)
        t2 = np.zeros(len(q), np.float32)
        _lib.check(self.L.mbir_hip_voxel_thetas(self.h, len(q), _p(q), _p(t1), _p(t2)), "voxel_thetas")
        return t1, t2

    def kernel_time(self, k):
        ms = C.c_double()
        n = C.c_int64()
        _lib.check(self.L.mbir_hip_kernel_time(self.h, k, C.byref(ms), C.byref(n)), "kernel_time")
        return ms.value, n.value

    def snapshot(self, restore=False):
        """Saves the object on the device, or restores the saved one."""
        _lib.check(self.L.mbir_hip_object_snapshot(self.h, int(bool(restore))), "object_snapshot")

    def xcd_check(self):
        """(work items run, items that ran off their XCD) since option
        xcd_check was set."""
        items, off = C.c_int64(), C.c_int64()
        _lib.check(self.L.mbir_hip_xcd_check(self.h, C.byref(items), C.byref(off)), "xcd_check")
        return items.value, off.value

    def reset_kernel_time(self):
        _lib.check(self.L.mbir_hip_reset_kernel_time(self.h), "reset_kernel_time")

    def synchronize(self):
        _lib.check(self.L.mbir_hip_synchronize(self.h), "synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.L.mbir_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Guard:
    """The engine's cost-rise guard (mbir_guard_check, include/mbir_engine.h)
    over one Stage: call check() with each cost from the maintained error
    sinogram.  It saves the object at each new lowest cost, confirms a rise
    beyond tol from a recomputed error sinogram, and on a confirmed rise
    restores the saved object, recomputes e and lowers the stage's alpha;
    failed is set once max_recoveries are spent."""

    def __init__(self, stage, cost0, tol=1e-6, max_recoveries=3):
        self.L = _lib.load()
        self.st = stage
        self.s = _lib.GuardState()
        self.L.mbir_guard_init(C.byref(self.s), float(cost0), float(tol), int(max_recoveries))
        self.failed = False
        self.events = []
        stage.snapshot()

    @property
    def best(self):
        return self.s.best

    def check(self, cost, where=None):
        """Returns (action, cost): the MBIR_GUARD_ code and the cost the
        stage now has (the saved object's after a recovery)."""
        act = self.L.mbir_guard_check(C.byref(self.s), float(cost), 0)
        if act == _lib.GUARD_CONFIRM:
            cost = self.st.cost(fresh=True)[0]
            act = self.L.mbir_guard_check(C.byref(self.s), float(cost), 1)
        if act == _lib.GUARD_SAVE:
            self.st.snapshot()
        elif act == _lib.GUARD_RECOVER:
            a2 = self.L.mbir_guard_alpha(self.st.alpha)
            self.events.append({"where": where, "cost": float(cost), "best": self.s.best,
                                "alpha": self.st.alpha, "new_alpha": a2})
            self.st.snapshot(restore=True)
            self.st.init_error()
            self.st.option("alpha", a2)
            cost = self.s.best
        elif act == _lib.GUARD_FAIL:
            self.failed = True
        return act, cost


def tomo_prior(phase, mag, mult_xy, delta_recon, wavenumber, delta_over_beta=None, mu=1.0,
               alpha=1.5, p=1.2, q=2.0):
    """The PCMBIR tomographic prior of one stage as mbir_hip_pc_enable takes
    it (PCMBIR_3D/XT_Init.c:144-176, 234-246, 482-502).  phase / mag =
    (sig_s, sig_t, c_s, c_t) of the phase (delta) and magnitude (beta)
    priors; spatial sigmas scaled by mult_xy, temporal by delta_recon, each
    raised to q and q - p; the 26-neighbour filter 1/distance normalised to
    sum 1 and the temporal filter 0 (initFilter); DecorrTran
    [[1, d/b], [1, -d/b]] for a delta_over_beta (recon_type 2), else the
    identity."""
    P = _lib.PcPrior()
    P.mu, P.wavenumber, P.alpha, P.p, P.q = mu, wavenumber, alpha, p, q
    T = (1.0, 0.0, 0.0, 1.0) if delta_over_beta is None else \
        (1.0, delta_over_beta, 1.0, -delta_over_beta)
    for i, v in enumerate(T):
        P.T[i] = v
    for k, (ss, st, cs, ct) in enumerate((phase, mag)):
        P.c_s[k], P.c_t[k] = cs, ct
        P.sig_s_q[k], P.sig_s_qp[k] = (ss * mult_xy) ** q, (ss * mult_xy) ** (q - p)
        P.sig_t_q[k], P.sig_t_qp[k] = (st * delta_recon) ** q, (st * delta_recon) ** (q - p)
    o = np.indices((3, 3, 3)).reshape(3, 27) - 1.0
    dist = np.sqrt((o ** 2).sum(0))
    filt = np.divide(1.0, dist, out=np.zeros(27), where=dist > 0)
    filt /= filt.sum()
    for k in range(27):
        P.filt[k] = filt[k]
    P.filt_t = 0.0
    return P


class TomoStage(Stage):
    """A Stage switched to the PCMBIR tomographic ICD (mbir_hip_pc_*): two
    objects, delta and beta, fitted to two target sinograms (the ADMM's
    phase and magnitude projections plus duals) through A times the
    wavenumber.  The scan supplies the geometry; its projections and
    weights are not used."""

    def __init__(self, scan, prior, mult_xy=1, mult_z=1, tile=8):
        super().__init__(scan, mult_xy=mult_xy, mult_z=mult_z, tile=tile, alpha=prior.alpha,
                         full_mask=True)
        self.prior = prior
        _lib.check(self.L.mbir_hip_pc_enable(self.h, C.byref(prior)), "mbir_hip_pc_enable")
        self.build_amatrix()

    def set_objects(self, delta, beta):
        d = np.ascontiguousarray(delta, np.float32).reshape(self.shape_x)
        b = np.ascontiguousarray(beta, np.float32).reshape(self.shape_x)
        _lib.check(self.L.mbir_hip_pc_set_objects(self.h, _p(d), _p(b)), "pc_set_objects")

    def get_objects(self):
        d = np.zeros(self.shape_x, np.float32)
        b = np.zeros(self.shape_x, np.float32)
        _lib.check(self.L.mbir_hip_pc_get_objects(self.h, _p(d), _p(b)), "pc_get_objects")
        return d, b

    def init_errors(self, target_delta, target_beta):
        """e = target - k A x for both objects; targets in the internal
        sinogram layout [N_p][N_r][N_t]."""
        td = np.ascontiguousarray(target_delta, np.float32).reshape(self.shape_e)
        tb = np.ascontiguousarray(target_beta, np.float32).reshape(self.shape_e)
        _lib.check(self.L.mbir_hip_pc_init_error(self.h, _p(td), _p(tb)), "pc_init_error")

    def get_errors(self):
        ed = np.zeros(self.shape_e, np.float32)
        eb = np.zeros(self.shape_e, np.float32)
        _lib.check(self.L.mbir_hip_pc_get_error(self.h, _p(ed), _p(eb)), "pc_get_error")
        return ed, eb

    def equits(self, n, first_iter, seed=761521):
        """n equits; returns [sum |dx|, voxels updated, voxels skipped]."""
        st = np.zeros(3, np.float64)
        _lib.check(self.L.mbir_hip_pc_equits(self.h, int(n), int(first_iter), seed & 0xFFFFFFFF,
                                             _p(st)), "pc_equits")
        return st

    def tomo_cost(self):
        """FP64 [total, data, prior]."""
        out = np.zeros(3, np.float64)
        _lib.check(self.L.mbir_hip_pc_cost(self.h, _p(out)), "pc_cost")
        return out

    # -- the ADMM head loop's state (mbir_hip_pc_admm_*), sinograms in the
    #    internal layout [N_p][N_r][N_t]
    def admm_init(self, z_delta, z_beta, from_object=False):
        zd = np.ascontiguousarray(z_delta, np.float64).reshape(self.shape_e)
        zb = np.ascontiguousarray(z_beta, np.float64).reshape(self.shape_e)
        _lib.check(self.L.mbir_hip_pc_admm_init(self.h, _p(zd), _p(zb), int(bool(from_object))),
                   "pc_admm_init")

    def admm_target(self):
        """The phase retrieval's proximal point a = k A x + V, (delta, beta)."""
        ad = np.zeros(self.shape_e, np.float64)
        ab = np.zeros(self.shape_e, np.float64)
        _lib.check(self.L.mbir_hip_pc_admm_target(self.h, _p(ad), _p(ab)), "pc_admm_target")
        return ad, ab

    def admm_update(self, z_delta, z_beta):
        """Dual update with the retrieval's new estimate; returns [beta
        primal, beta dual, delta primal, delta dual] residuals."""
        zd = np.ascontiguousarray(z_delta, np.float64).reshape(self.shape_e)
        zb = np.ascontiguousarray(z_beta, np.float64).reshape(self.shape_e)
        res = np.zeros(4, np.float64)
        _lib.check(self.L.mbir_hip_pc_admm_update(self.h, _p(zd), _p(zb), _p(res)),
                   "pc_admm_update")
        return res

    def admm_state(self, which):
        """which 0: Z, 1: V, 2: P; returns (delta, beta) FP32."""
        d = np.zeros(self.shape_e, np.float32)
        b = np.zeros(self.shape_e, np.float32)
        _lib.check(self.L.mbir_hip_pc_admm_state(self.h, int(which), _p(d), _p(b)),
                   "pc_admm_state")
        return d, b


def substitution_batch(g, V, th1, th2, nb, nbf, tn, tf):
    """Batched functional substitution on the GPU (FP32)."""
    L = _lib.load()
    n = len(V)
    arr = [np.ascontiguousarray(a, np.float32) for a in (V, th1, th2, nb, tn)]
    nbf = np.ascontiguousarray(nbf, np.uint8)
    tf = np.ascontiguousarray(tf, np.uint8)
    out = np.zeros(n, np.float32)
    _lib.check(L.mbir_hip_substitution_batch(C.byref(g), C.c_int32(n), _p(arr[0]), _p(arr[1]),
                                             _p(arr[2]), _p(arr[3]), _p(nbf), _p(arr[4]),
                                             _p(tf), _p(out)), "substitution_batch")
    return out


def default_options():
    o = _lib.Options()
    _lib.load().mbir_default_options(C.byref(o))
    return o


def reconstruct(scan, convg_thresh=1.0, remove_rings=0, quad_convex=1, huber_delta=0.5,
                huber_T=4.0, restart=0, options=None, log_path=None):
    """Runs the native reconstruct_ex(); returns (object with halo slices,
    equits of the last stage, final cost)."""
    L = _lib.load()
    s = scan
    keep = [np.ascontiguousarray(s[k], np.float32)
            for k in ("projections", "weights", "angles", "proj_times", "recon_times")]
    obj = C.POINTER(C.c_float)()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen((log_path or "/dev/null").encode(), b"w")
    opt = options or default_options()
    eq = C.c_int32()
    cost = C.c_double()
    f = C.c_float
    try:
        rc = L.reconstruct_ex(C.byref(obj), *[_p(a) for a in keep], C.c_int32(s["proj_rows"]),
                              C.c_int32(s["proj_cols"]), C.c_int32(s["proj_num"]),
                              C.c_int32(s["recon_num"]), f(s["vox_wid"]), f(s["rot_center"]),
                              f(s["sig_s"]), f(s["sig_t"]), f(s["c_s"]), f(s["c_t"]),
                              f(convg_thresh), C.c_int32(remove_rings), C.c_int32(quad_convex),
                              f(huber_delta), f(huber_T), C.c_uint8(restart), C.c_void_p(fp),
                              C.byref(opt), C.byref(eq), C.byref(cost))
    finally:
        libc.fclose(C.c_void_p(fp))
    if rc != 0:
        raise RuntimeError(f"reconstruct returned {rc}: {_lib.last_error()}")
    shape = (s["recon_num"], s["proj_rows"] + 2, s["proj_cols"], s["proj_cols"])
    out = np.ctypeslib.as_array(obj, shape=(int(np.prod(shape)),)).reshape(shape).copy()
    libc.free.argtypes = [C.c_void_p]
    libc.free(C.cast(obj, C.c_void_p))
    return out, eq.value, cost.value
