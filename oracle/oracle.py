"""ctypes access to the test-only oracle libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

* ``Oracle`` wraps oracle/liboracle_mbir.so, the FP64 C restatement of the
  reference MBIR_4D hot path (oracle/mbir_oracle.c).
* ``Reference`` wraps oracle/_ref/libmbir4d_ref.so, the reference's own C
  sources built by ``make -C oracle ref`` (present only where it was built).
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_mbir.so")
REF_SO = os.path.join(HERE, "_ref", "libmbir4d_ref.so")
VIEWS_REF_SO = os.path.join(HERE, "_ref", "libviews_ref.so")
NB = 65  # detector response table entries per view

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class Geom(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("N_r", "N_t", "N_p")] + \
        [(n, C.c_double) for n in ("delta_r", "delta_t", "R0", "RMax", "T0", "TMax",
                                   "OffsetR", "OffsetT")] + \
        [(n, C.c_int32) for n in ("N_x", "N_y", "N_z", "N_time", "mult_xy", "mult_z", "W")] + \
        [(n, C.c_double) for n in ("x0", "y0", "z0", "delta_xy", "delta_z", "radius_obj",
                                   "delta_recon", "p_s", "p_t", "q", "c_s", "c_t",
                                   "sig_s_q", "sig_s_qp", "sig_t_q", "sig_t_qp")] + \
        [("filt", C.c_double * 27), ("filt_t", C.c_double), ("alpha", C.c_double),
         ("huber_T", C.c_double), ("huber_delta", C.c_double), ("var_est", C.c_double)]


class PcPrior(C.Structure):
    """mirror of orc_pc_prior (oracle/mbir_oracle.h)"""
    _fields_ = [("mu", C.c_double), ("wavenumber", C.c_double), ("T", C.c_double * 4)] + \
        [(n, C.c_double * 2) for n in ("c_s", "sq_s", "sqp_s", "c_t", "sq_t", "sqp_t")] + \
        [("p", C.c_double), ("q", C.c_double), ("filt", C.c_double * 27), ("filt_t", C.c_double),
         ("alpha", C.c_double)]


def pc_prior(phase, mag, mult_xy, delta_recon, wavenumber, delta_over_beta=None, mu=1.0,
             alpha=1.5, p=1.2, q=2.0):
    """The PCMBIR tomographic prior and data weights of one stage
    (PCMBIR_3D/XT_Init.c:144-176, 234-246, 482-502): phase / mag = (sig_s,
    sig_t, c_s, c_t); sigmas scaled by mult_xy (spatial) and delta_recon
    (temporal); the 26-neighbour filter 1/distance normalised to sum 1 and
    the temporal filter 0 (initFilter); DecorrTran [[1, d/b], [1, -d/b]]
    with a delta_over_beta (recon_type 2), else the identity."""
    P = PcPrior()
    P.mu, P.wavenumber, P.p, P.q, P.alpha = mu, wavenumber, p, q, alpha
    T = [1.0, 0.0, 0.0, 1.0] if delta_over_beta is None else \
        [1.0, delta_over_beta, 1.0, -delta_over_beta]
    for i, v in enumerate(T):
        P.T[i] = v
    for k, (ss, st, cs, ct) in enumerate((phase, mag)):
        P.c_s[k], P.c_t[k] = cs, ct
        P.sq_s[k], P.sqp_s[k] = (ss * mult_xy) ** q, (ss * mult_xy) ** (q - p)
        P.sq_t[k], P.sqp_t[k] = (st * delta_recon) ** q, (st * delta_recon) ** (q - p)
    d = np.array([[[np.sqrt((i - 1.0) ** 2 + (j - 1.0) ** 2 + (k - 1.0) ** 2)
                    for k in range(3)] for j in range(3)] for i in range(3)]).ravel()
    f = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    f /= f.sum()
    for k in range(27):
        P.filt[k] = f[k]
    P.filt_t = 0.0
    return P


class AsyncOpts(C.Structure):
    _fields_ = [("T", C.c_int32), ("alpha", C.c_double), ("seed", C.c_uint32),
                ("iter", C.c_int32), ("call_index", C.c_int32), ("lag_same", C.c_int32),
                ("lag_cross", C.c_int32), ("groups", C.c_int32), ("placement", C.c_int32)]


def _lib(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (make -C oracle)")
    return C.CDLL(path)


def sino_to_internal(a, N_p, N_t, N_r):
    """file layout [p][t][r] -> internal [p][r][t] (XT_Init.c:294-301)."""
    return np.ascontiguousarray(np.asarray(a).reshape(N_p, N_t, N_r).transpose(0, 2, 1))


class Oracle:
    """The FP64 restatement for one stage of one scan."""

    def __init__(self, scan, mult_xy=1, mult_z=1, quad_convex=1, huber_delta=0.5,
                 huber_T=4.0, p_s=None, p_t=None):
        L = self.L = _lib(ORACLE_SO)
        L.orc_random2.restype = C.c_double
        L.orc_cost.restype = C.c_double
        L.orc_substitution.restype = C.c_double
        L.orc_qggmrf_weight.restype = C.c_double
        L.orc_qggmrf_value.restype = C.c_double
        L.orc_amatrix_column.restype = C.c_int32
        L.orc_setup.restype = C.c_int32
        L.orc_view_map.restype = C.c_int32
        s = self.scan = scan
        N_p, N_t = s["proj_num"], s["proj_rows"]
        self.frame_ptr = np.zeros(s["recon_num"] + 1, np.int32)
        self.frame_views = np.zeros(max(N_p, 1), np.int32)
        L.orc_view_map(C.c_int32(N_p), s["proj_times"].ctypes.data_as(C.c_void_p),
                       C.c_int32(s["recon_num"]), s["recon_times"].ctypes.data_as(C.c_void_p),
                       self.frame_ptr.ctypes.data_as(C.c_void_p),
                       self.frame_views.ctypes.data_as(C.c_void_p))
        g = self.g = Geom()
        rc = L.orc_setup(C.byref(g), N_t, s["proj_cols"], N_p, s["recon_num"],
                         C.c_float(s["vox_wid"]), C.c_float(s["rot_center"]),
                         C.c_float(s["sig_s"]), C.c_float(s["sig_t"]),
                         C.c_float(s["c_s"]), C.c_float(s["c_t"]), mult_xy, mult_z,
                         quad_convex, C.c_float(huber_delta), C.c_float(huber_T),
                         self.frame_ptr.ctypes.data_as(C.c_void_p),
                         C.c_double(s["p_s"] if p_s is None else p_s),
                         C.c_double(s["p_t"] if p_t is None else p_t))
        if rc:
            raise ValueError("orc_setup rejected the geometry")
        self.angles = np.ascontiguousarray(s["angles"], np.float32)
        self.cos = np.zeros(N_p, np.float64)
        self.sin = np.zeros(N_p, np.float64)
        L.orc_sincos(N_p, self.angles.ctypes.data_as(C.c_void_p),
                     self.cos.ctypes.data_as(C.c_void_p), self.sin.ctypes.data_as(C.c_void_p))
        self.H = np.zeros((N_p, NB), np.float32)
        L.orc_detector_response(C.byref(g), self.angles.ctypes.data_as(C.c_void_p),
                                self.H.ctypes.data_as(C.c_void_p))
        self.y = sino_to_internal(s["projections"], N_p, N_t, s["proj_cols"]).astype(np.float64)
        self.w = sino_to_internal(s["weights"], N_p, N_t, s["proj_cols"]).astype(np.float64)
        self.mask = np.zeros((g.N_y, g.N_x), np.uint8)
        L.orc_mask(C.byref(g), self.mask.ctypes.data_as(C.c_void_p))

    def _p(self, a):
        return a.ctypes.data_as(C.c_void_p)

    @property
    def shape_x(self):
        g = self.g
        return (g.N_time, g.N_z, g.N_y, g.N_x)

    def amatrix_column(self, row, col, view):
        idx = np.zeros(256, np.int32)
        val = np.zeros(256, np.float64)
        n = self.L.orc_amatrix_column(C.byref(self.g), self._p(self.cos), self._p(self.sin),
                                      self._p(self.H), int(row), int(col), int(view), self._p(idx),
                                      self._p(val))
        return idx[:n].copy(), val[:n].copy()

    def amatrix_all(self):
        g = self.g
        first = np.zeros((g.N_y * g.N_x, g.N_p), np.int32)
        val = np.zeros((g.N_y * g.N_x, g.N_p, g.W), np.float64)
        self.L.orc_amatrix_all(C.byref(g), self._p(self.cos), self._p(self.sin), self._p(self.H),
                               self._p(first), self._p(val))
        return first, val

    def forward_project(self, x):
        g = self.g
        x = np.ascontiguousarray(x, np.float64).reshape(self.shape_x)
        ax = np.zeros((g.N_p, g.N_r, g.N_t), np.float64)
        self.L.orc_forward_project(C.byref(g), self._p(self.cos), self._p(self.sin),
                                   self._p(self.H), self._p(self.frame_ptr),
                                   self._p(self.frame_views), self._p(x), self._p(ax))
        return ax

    def init_error(self, x):
        return self.y - self.forward_project(x)

    def cost(self, e, x):
        out = np.zeros(2, np.float64)
        e = np.ascontiguousarray(e, np.float64)
        x = np.ascontiguousarray(x, np.float64).reshape(self.shape_x)
        c = self.L.orc_cost(C.byref(self.g), self._p(e), self._p(self.w), self._p(x), self._p(out))
        return c, out

    def substitution(self, V, th1, th2, nb, nbf, tn, tf):
        nb = np.ascontiguousarray(nb, np.float64).ravel()
        nbf = np.ascontiguousarray(nbf, np.uint8).ravel()
        tn = np.ascontiguousarray(tn, np.float64)
        tf = np.ascontiguousarray(tf, np.uint8)
        return self.L.orc_substitution(C.byref(self.g), C.c_double(V), C.c_double(th1),
                                       C.c_double(th2), self._p(nb), self._p(nbf),
                                       self._p(tn), self._p(tf))

    def qggmrf_weight(self, d, spatial=True):
        return self.L.orc_qggmrf_weight(C.byref(self.g), C.c_double(d), int(spatial))

    def qggmrf_value(self, d, spatial=True):
        return self.L.orc_qggmrf_value(C.byref(self.g), C.c_double(d), int(spatial))

    def thetas(self, e, t, z, row, col, w=None):
        th1, th2 = C.c_double(), C.c_double()
        e = np.ascontiguousarray(e, np.float64)
        w = self.w if w is None else np.ascontiguousarray(w, np.float64)
        self.L.orc_voxel_thetas(C.byref(self.g), self._p(self.cos), self._p(self.sin),
                                self._p(self.H), self._p(self.frame_ptr),
                                self._p(self.frame_views), self._p(e), self._p(w),
                                t, z, row, col, C.byref(th1), C.byref(th2))
        return th1.value, th2.value

    def icd_equit(self, x, e, iteration, num_z_blocks=2, K=10):
        stats = np.zeros(3, np.float64)
        self.L.orc_icd_equit(C.byref(self.g), self._p(self.cos), self._p(self.sin),
                             self._p(self.H), self._p(self.frame_ptr), self._p(self.frame_views),
                             self._p(self.mask), num_z_blocks, K, iteration, self._p(x),
                             self._p(e), self._p(self.w), self._p(stats))
        return stats

    def icd_equit_tiled(self, x, e, iteration, T, alpha, seed):
        self.L.orc_icd_equit_tiled(C.byref(self.g), self._p(self.cos), self._p(self.sin),
                                   self._p(self.H), self._p(self.frame_ptr),
                                   self._p(self.frame_views), self._p(self.mask), T,
                                   C.c_double(alpha), C.c_uint32(seed), iteration, self._p(x),
                                   self._p(e), self._p(self.w))

    def icd_equit_async(self, x, e, iteration, T, alpha, seed, call_index=0, lag_same=0,
                        lag_cross=0, groups=8, placement=0):
        """One equit of the GPU sweep's schedule (orc_icd_equit_async): x
        [N_time][N_z][N_y][N_x] and e [N_p][N_r][N_t] (float64) updated in
        place; returns stats (sum |dx|, sum |x|, skipped)."""
        if not (lag_cross < 0 or lag_cross >= lag_same) or lag_same < 0:
            raise ValueError("need lag_same >= 0 and lag_cross < 0 or >= lag_same")
        assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == np.prod(self.shape_x)
        assert e.dtype == np.float64 and e.flags.c_contiguous and e.shape == self.y.shape
        op = AsyncOpts(T, alpha, seed & 0xFFFFFFFF, iteration, call_index, lag_same, lag_cross,
                       groups, placement)
        stats = np.zeros(3, np.float64)
        self.L.orc_icd_equit_async(C.byref(self.g), self._p(self.cos), self._p(self.sin),
                                   self._p(self.H), self._p(self.frame_ptr),
                                   self._p(self.frame_views), self._p(self.mask), C.byref(op),
                                   self._p(x), self._p(e), self._p(self.w), self._p(stats))
        return stats

    # --- PCMBIR tomographic ICD (row a14)
    def pc_substitution(self, P, Vd, Vb, th1d, th1b, th2d, th2b, nbd, nbb, nbf, tnd, tnb, tf):
        """orc_pc_substitution on one voxel; returns (delta, beta)."""
        vd, vb = C.c_double(Vd), C.c_double(Vb)
        arr = lambda a, t=np.float64: np.ascontiguousarray(a, t).ravel()
        a = [arr(nbd), arr(nbb), arr(nbf, np.uint8), arr(tnd), arr(tnb), arr(tf, np.uint8)]
        self.L.orc_pc_substitution(C.byref(P), C.byref(vd), C.byref(vb), C.c_double(th1d),
                                   C.c_double(th1b), C.c_double(th2d), C.c_double(th2b),
                                   *[self._p(v) for v in a])
        return vd.value, vb.value

    def pc_init_error(self, P, target_d, target_b, xd, xb):
        """Error sinograms e = target - k A x of both objects (internal
        layout [N_p][N_r][N_t])."""
        k = P.wavenumber
        return (np.asarray(target_d, np.float64) - k * self.forward_project(xd),
                np.asarray(target_b, np.float64) - k * self.forward_project(xb))

    def pc_cost(self, P, ed, eb, xd, xb):
        out = np.zeros(2, np.float64)
        c = [np.ascontiguousarray(a, np.float64) for a in (ed, eb, xd, xb)]
        self.L.orc_pc_cost.restype = C.c_double
        tot = self.L.orc_pc_cost(C.byref(self.g), C.byref(P), *[self._p(v) for v in c],
                                 self._p(out))
        return tot, out

    def pc_equit(self, P, xd, xb, ed, eb, iteration, T, seed, call_index=0, mask=None):
        """One equit of the GPU sweep's order on both objects (float64,
        updated in place); mask None: every column (PCMBIR updates all
        voxels, Q18).  Returns stats (sum |dx|, 0, skipped)."""
        for a in (xd, xb, ed, eb):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        m = np.ones_like(self.mask) if mask is None else np.ascontiguousarray(mask, np.uint8)
        stats = np.zeros(3, np.float64)
        self.L.orc_pc_equit(C.byref(self.g), self._p(self.cos), self._p(self.sin),
                            self._p(self.H), self._p(self.frame_ptr), self._p(self.frame_views),
                            self._p(m), C.byref(P), T, C.c_uint32(seed & 0xFFFFFFFF), iteration,
                            call_index, self._p(xd), self._p(xb), self._p(ed), self._p(eb),
                            self._p(stats))
        return stats

    def srandom2(self, seed):
        self.L.orc_srandom2(C.c_uint32(seed))

    def random2(self):
        return self.L.orc_random2()

    def run_icd(self, n_equits, x=None, num_z_blocks=2, K=10, seed=761521):
        """Sequential-schedule ICD from x (zeros by default); returns x, e, costs."""
        x = np.zeros(self.shape_x, np.float64) if x is None else np.array(x, np.float64)
        e = self.init_error(x)
        self.srandom2(seed)
        costs = [self.cost(e, x)[0]]
        for it in range(1, n_equits + 1):
            self.icd_equit(x, e, it, num_z_blocks, K)
            costs.append(self.cost(e, x)[0])
        return x, e, np.array(costs)


def _orc():
    L = _lib(ORACLE_SO)
    L.orc_offset_windows.restype = C.c_int32
    return L


def offset_windows(N_r, N_t):
    """orc_offset_windows: [M][N_r][N_t] window weights."""
    L = _orc()
    M = L.orc_offset_windows(C.c_int32(N_r), C.c_int32(N_t), None)
    win = np.zeros((M, N_r, N_t), np.float64)
    L.orc_offset_windows(C.c_int32(N_r), C.c_int32(N_t), win.ctypes.data_as(C.c_void_p))
    return win


def offset_update(mode, w, e, d, sel, T, delta):
    """orc_offset_update on float64 copies ([p][r][t], [r][t]); returns
    (e, d, sel)."""
    L = _orc()
    N_p, N_r, N_t = e.shape
    w = np.ascontiguousarray(w, np.float64)
    e = np.array(e, np.float64)
    d = np.array(d, np.float64)
    sel = np.array(sel, np.uint8)
    L.orc_offset_update(C.c_int32(mode), C.c_int32(N_p), C.c_int32(N_r), C.c_int32(N_t),
                        C.c_double(T), C.c_double(delta), w.ctypes.data_as(C.c_void_p),
                        e.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p),
                        sel.ctypes.data_as(C.c_void_p))
    return e, d, sel


def variance_update(var, w, e, sel, T, delta):
    """orc_variance_update on float64 copies; returns (var, w, sel)."""
    L = _orc()
    w = np.array(w, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    sel = np.array(sel, np.uint8)
    v = C.c_double(var)
    L.orc_variance_update(C.c_int64(e.size), C.c_double(T), C.c_double(delta),
                          w.ctypes.data_as(C.c_void_p), e.ctypes.data_as(C.c_void_p),
                          sel.ctypes.data_as(C.c_void_p), C.byref(v))
    return v.value, w, sel


def phantom_project(mag, phase, angles, proj_rows, proj_cols, vox_wid, rot_center, wavenumber):
    """orc_phantom_project: the phase-contrast simulator's projections of
    the phantoms [n_z][n_xy][n_xy] (PCMBIR_3D/XT_genSinogram.c:146-186);
    returns (real, imag) [n_p][n_z][n_xy] in float64."""
    L = _orc()
    mag = np.ascontiguousarray(mag, np.float32)
    phase = np.ascontiguousarray(phase, np.float32)
    ang = np.ascontiguousarray(angles, np.float32)
    n_z, n_xy = mag.shape[0], mag.shape[-1]
    out_r = np.zeros((ang.size, n_z, n_xy))
    out_i = np.zeros_like(out_r)
    rc = L.orc_phantom_project(
        mag.ctypes.data_as(C.c_void_p), phase.ctypes.data_as(C.c_void_p), C.c_int32(n_xy),
        C.c_int32(n_z), C.c_int32(ang.size), ang.ctypes.data_as(C.c_void_p),
        C.c_int32(proj_rows), C.c_int32(proj_cols), C.c_double(vox_wid), C.c_double(rot_center),
        C.c_double(wavenumber), out_r.ctypes.data_as(C.c_void_p),
        out_i.ctypes.data_as(C.c_void_p))
    if rc:
        raise ValueError("n_xy must be a multiple of proj_cols")
    return out_r, out_i


NHICD_ITERATIONS = 10  # NHICD_Iterations (XT_Init.c:464)


def vsc(mag):
    """ComputeVSC (XT_NHICD.c:126-160): the magnitude map(s) [..., N_y, N_x]
    filtered with the 5 x 5 Hamming window, the outer product of the
    5-point window 0.54 - 0.46 cos(2 pi n / 4), truncated at the edges;
    accumulated in float64, stored as float32 as the reference does."""
    mag = np.asarray(mag, np.float64)
    h = 0.54 - 0.46 * np.cos(2.0 * np.pi * np.arange(5) / 4.0)
    W = np.outer(h, h)
    ny, nx = mag.shape[-2:]
    pad = np.zeros(mag.shape[:-2] + (ny + 4, nx + 4))
    pad[..., 2:-2, 2:-2] = mag
    out = np.zeros_like(mag)
    for p in range(5):
        for q in range(5):
            out += W[p, q] * pad[..., p:p + ny, q:q + nx]
    return out.astype(np.float32)


def vsc_select(filt, count, iterations=NHICD_ITERATIONS):
    """VSC_based_Voxel_Line_Select (XT_NHICD.c:164-216) for one frame's
    filtered map [N_y, N_x]: k = floor(count / iterations) lines, the
    threshold the (N - k)-th smallest value of the whole map
    (RandomizedSelect, XT_NHICD.c:60-80), every line strictly above it
    selected (fewer than k when values tie at the threshold)."""
    flat = np.asarray(filt).ravel()
    k = count // iterations
    thresh = np.sort(flat)[flat.size - k - 1] if k > 0 else np.inf
    return np.asarray(filt) > thresh


def aps_normalize(proj, white, dark, ratio, bh_coef=0.5):
    """read_data's flat-field step (reconstruct/real_data/APS_Data_Format/
    XT_HDFIO.c:138-178) in the reference's arithmetic: float32 sums in its
    order (frames; then block rows outer, channels inner), the quotient in
    float32, log and the polynomial in float64.  Returns (projections,
    weights) [views][rows][cols]."""
    f32 = np.float32
    proj, white, dark = (np.asarray(a, np.uint16) for a in (proj, white, dark))
    mw = np.zeros(white.shape[1:], f32)
    for fr in white:
        mw = mw + fr.astype(f32)
    mw = mw / f32(white.shape[0])
    md = np.zeros(dark.shape[1:], f32)
    for fr in dark:
        md = md + fr.astype(f32)
    md = md / f32(dark.shape[0])
    n, R, Cc = proj.shape
    rows, cols = R // ratio, Cc // ratio
    w = np.zeros((n, rows, cols), f32)
    wd = np.zeros((rows, cols), f32)
    for m in range(ratio):
        for q in range(ratio):
            dk = md[m::ratio, q::ratio][:rows, :cols]
            w = w + np.abs(proj[:, m::ratio, q::ratio][:, :rows, :cols].astype(f32) - dk)
            wd = wd + np.abs(mw[m::ratio, q::ratio][:rows, :cols] - dk)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.log((wd[None] / w).astype(np.float64)).astype(f32).astype(np.float64)
    return (bh_coef * t * t + t).astype(f32), w


def std_downsample(proj, weights, ratio, bh_coef=0.0):
    """read_ProjWeightData's arithmetic (reconstruct/real_data/
    Standard_Data_Format/XT_HDFIO.c:90-111): float32 block sums (block rows
    outer, channels inner), the mean divided in float32, the polynomial in
    float64.  Returns (projections, weights) [views][rows][cols]."""
    f32 = np.float32
    proj = np.asarray(proj, f32)
    weights = np.asarray(weights, f32)
    n, R, Cc = proj.shape
    rows, cols = R // ratio, Cc // ratio
    t = np.zeros((n, rows, cols), f32)
    w = np.zeros((n, rows, cols), f32)
    for m in range(ratio):
        for q in range(ratio):
            t = t + proj[:, m::ratio, q::ratio]
            w = w + weights[:, m::ratio, q::ratio]
    td = (t / f32(ratio * ratio)).astype(np.float64)
    return (bh_coef * td * td + td).astype(f32), w


def reference_interlaced_views(K, n_theta, n_views, proj_start=0, min_acq_time=0.0,
                               rot_speed=0.0):
    """The APS driver's own gen_interlaced_views_0_to_Inf (XT_Views.c:7-46),
    from the reference build; returns (views, times, count)."""
    L = _lib(VIEWS_REF_SO)
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen(b"/dev/null", b"w")
    views = np.zeros(n_views, np.float32)
    times = np.zeros(n_views, np.float32)
    L.gen_interlaced_views_0_to_Inf.restype = C.c_int32
    n = L.gen_interlaced_views_0_to_Inf(
        views.ctypes.data_as(C.c_void_p), times.ctypes.data_as(C.c_void_p), C.c_int32(K),
        C.c_int32(n_theta), C.c_int32(proj_start), C.c_int32(n_views), C.c_float(min_acq_time),
        C.c_float(rot_speed), C.c_void_p(fp))
    libc.fclose(fp)
    return views, times, n


def views_reference_available():
    return os.path.exists(VIEWS_REF_SO)


def reference_available():
    return os.path.exists(REF_SO)


_REF = None


def _ref_lib():
    global _REF
    if _REF is None:
        L = _lib(REF_SO)
        L.ref_create.restype = C.c_void_p
        L.ref_substitution.restype = C.c_double
        L.ref_cost.restype = C.c_double
        L.ref_random2.restype = C.c_double
        L.ref_amatrix_column.restype = C.c_int32
        L.ref_update_voxel.restype = C.c_long
        L.ref_init_error.restype = C.c_int32
        L.ref_reconstruct.restype = C.c_int32
        L.ref_variance_update.restype = C.c_double
        L.ref_variance_update.argtypes = [C.c_void_p, C.c_double] + [C.c_void_p] * 3
        _REF = L
    return _REF


class Reference:
    """The reference build (oracle/_ref) for one stage of one scan."""

    def __init__(self, scan, mult_xy=1, mult_z=1, quad_convex=1, remove_rings=0,
                 huber_delta=0.5, huber_T=4.0, convg=1.0, threads=0):
        self.L = L = _ref_lib()
        self.scan = s = scan
        self.tmp = tempfile.mkdtemp(prefix="mbir_ref_")
        self._keep = [np.ascontiguousarray(s[k], np.float32)
                      for k in ("projections", "weights", "angles", "proj_times", "recon_times")]
        P = [a.ctypes.data_as(C.c_void_p) for a in self._keep]
        f = C.c_float
        self.h = L.ref_create(self.tmp.encode(), mult_xy, mult_z, *P,
                              s["proj_rows"], s["proj_cols"], s["proj_num"], s["recon_num"],
                              f(s["vox_wid"]), f(s["rot_center"]), f(s["sig_s"]), f(s["sig_t"]),
                              f(s["c_s"]), f(s["c_t"]), f(convg), remove_rings, quad_convex,
                              f(huber_delta), f(huber_T), threads)
        if not self.h:
            raise RuntimeError("reference initStructures failed")
        self.h = C.c_void_p(self.h)
        ints = np.zeros(8, np.int32)
        dbls = np.zeros(24 + 27, np.float64)
        L.ref_geometry(self.h, ints.ctypes.data_as(C.c_void_p), dbls.ctypes.data_as(C.c_void_p))
        self.ints, self.dbls = ints, dbls
        (self.N_r, self.N_t, self.N_p, self.N_x, self.N_y, self.N_z, self.N_time,
         self.num_z_blocks) = [int(v) for v in ints]

    def _p(self, a):
        return a.ctypes.data_as(C.c_void_p)

    @property
    def shape_x(self):
        return (self.N_time, self.N_z, self.N_y, self.N_x)

    @property
    def shape_e(self):
        return (self.N_p, self.N_r, self.N_t)

    def view_map(self):
        ptr = np.zeros(self.N_time + 1, np.int32)
        views = np.zeros(max(self.N_p, 1), np.int32)
        self.L.ref_view_map(self.h, self._p(ptr), self._p(views))
        return ptr, views[:ptr[-1]]

    def detector_response(self):
        H = np.zeros((self.N_p, NB), np.float32)
        self.L.ref_detector_response(self.h, self._p(H))
        return H

    def amatrix_column(self, row, col, view):
        idx = np.zeros(256, np.int32)
        val = np.zeros(256, np.float64)
        n = self.L.ref_amatrix_column(self.h, C.c_int32(row), C.c_int32(col),
                                      C.c_int32(view), self._p(idx), self._p(val))
        return idx[:n].copy(), val[:n].copy()

    def set_object(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(self.shape_x)
        self.L.ref_set_object(self.h, self._p(x))

    def get_object(self):
        x = np.zeros(self.shape_x, np.float32)
        self.L.ref_get_object(self.h, self._p(x))
        return x

    def init_error(self):
        e = np.zeros(self.shape_e, np.float32)
        if self.L.ref_init_error(self.h, self._p(e)):
            raise RuntimeError("initErrorSinogam failed")
        return e

    def get_error(self):
        e = np.zeros(self.shape_e, np.float32)
        self.L.ref_get_error(self.h, self._p(e))
        return e

    def get_weight(self):
        w = np.zeros(self.shape_e, np.float32)
        self.L.ref_get_weight(self.h, self._p(w))
        return w

    def substitution(self, V, th1, th2, nb, nbf, tn, tf):
        nb = np.ascontiguousarray(nb, np.float64).ravel()
        nbf = np.ascontiguousarray(nbf, np.uint8).ravel()
        tn = np.ascontiguousarray(tn, np.float64)
        tf = np.ascontiguousarray(tf, np.uint8)
        return self.L.ref_substitution(self.h, C.c_double(V), C.c_double(th1),
                                       C.c_double(th2), self._p(nb), self._p(nbf),
                                       self._p(tn), self._p(tf))

    def update_voxel(self, t, z, row, col, iteration):
        return self.L.ref_update_voxel(self.h, C.c_int32(t), C.c_int32(z), C.c_int32(row),
                                       C.c_int32(col), C.c_int32(iteration))

    def cost(self):
        c = self.L.ref_cost(self.h)
        parts = np.zeros(2, np.float64)
        self.L.ref_cost_parts(self.h, self._p(parts))
        return c, parts

    def icd_run(self, n, first_iter=1):
        costs = np.zeros(n, np.float64)
        conv = np.zeros(n, np.int32)
        secs = np.zeros(n, np.float64)
        self.L.ref_icd_run(self.h, C.c_int32(n), C.c_int32(first_iter), self._p(costs),
                           self._p(conv), self._p(secs))
        return costs, conv, secs

    def offset_update(self, mode, w, e, d, sel):
        """update_Sinogram_Offset on float copies; returns (e, d, sel)."""
        w = np.ascontiguousarray(w, np.float32)
        e = np.array(e, np.float32)
        d = np.array(d, np.float32)
        sel = np.array(sel, np.uint8)
        self.L.ref_offset_update(self.h, C.c_int32(mode), self._p(w), self._p(e), self._p(d),
                                 self._p(sel))
        return e, d, sel

    def variance_update(self, var, w, e, sel):
        """update_variance_parameter; returns (var, w, sel)."""
        w = np.array(w, np.float32)
        e = np.ascontiguousarray(e, np.float32)
        sel = np.array(sel, np.uint8)
        v = self.L.ref_variance_update(self.h, float(var), self._p(w), self._p(e), self._p(sel))
        return v, w, sel

    def vsc(self, mag):
        """ComputeVSC of one map [N_y, N_x]."""
        mag = np.ascontiguousarray(mag, np.float32)
        out = np.zeros_like(mag)
        self.L.ref_vsc(self.h, self._p(mag), self._p(out))
        return out

    def vsc_select(self, mag, count):
        """VSC_based_Voxel_Line_Select on maps [N_time, N_y, N_x] with
        UpdateSelectNum = count; returns the selected lines as a bool map."""
        mag = np.ascontiguousarray(mag, np.float32)
        sel = np.zeros(mag.shape, np.uint8)
        self.L.ref_vsc_select(self.h, self._p(mag), C.c_int32(count), self._p(sel))
        return sel.astype(bool)

    def close(self):
        if self.h:
            self.L.ref_destroy(self.h)
            self.h = None


def reference_reconstruct(scan, convg=1.0, quad_convex=1, remove_rings=0, huber_delta=0.5,
                          huber_T=4.0, threads=0, scratch=None):
    """The reference's full reconstruct() (multi-resolution); returns the
    object [recon_num][rows+2][cols][cols] including the halo slices."""
    L = _ref_lib()
    s = scan
    keep = [np.ascontiguousarray(s[k], np.float32)
            for k in ("projections", "weights", "angles", "proj_times", "recon_times")]
    obj = np.zeros((s["recon_num"], s["proj_rows"] + 2, s["proj_cols"], s["proj_cols"]), np.float32)
    d = scratch or tempfile.mkdtemp(prefix="mbir_refrec_")
    f = C.c_float
    rc = L.ref_reconstruct(d.encode(), obj.ctypes.data_as(C.c_void_p),
                           *[a.ctypes.data_as(C.c_void_p) for a in keep],
                           s["proj_rows"], s["proj_cols"], s["proj_num"], s["recon_num"],
                           f(s["vox_wid"]), f(s["rot_center"]), f(s["sig_s"]), f(s["sig_t"]),
                           f(s["c_s"]), f(s["c_t"]), f(convg), remove_rings, quad_convex,
                           f(huber_delta), f(huber_T), threads)
    if rc:
        raise RuntimeError("reference reconstruct() returned -1")
    return obj, d
