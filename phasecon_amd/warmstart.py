"""Learned warm start of the MBIR solver (include/unet_hip.h, libmbir4d.so).

The U-Net maps images computed directly from the scan to an
estimate of the converged MBIR image; the solver starts its full-resolution
stage from that estimate (clipped at zero, zero outside the update mask) and
still runs to convergence, so the network changes how much work the solver
does, never its fixed point.  Network inputs (``network_input``): "fbp2"
(default) the filtered back-projection of each frame (``Stage.fbp``: Ram-Lak
filter, A^T, scaled so an object uniform inside the update mask keeps its
value) and, as a second channel, its mean over the frames (all the scan's
views: structure a frame's own few views leave streaky); "fbp3" adds the
mean over the frame and its two neighbours between those two; "fbp" the
first channel alone; "bp" the normalised back-projection (``Stage.bp_normalized``:
A^T(w (y - d)) / A^T(w A m)).  Training pairs come from the GPU MBIR itself:
(normalised back-projection, converged image) of scans other than the one
reconstructed.  There is no fallback: without the native library the
calls raise.

Volumes are scaled by ``SCALE`` before they enter the network (attenuation
values of 0.005-0.02 per voxel become 0.25-1).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _lib

SCALE = 50.0


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(L, rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {L.unet_last_error().decode(errors='replace')}")


class UNet:
    """3-level 3D U-Net with BF16 MFMA kernels on the GPU (unet_create)."""

    def __init__(self, channels=16, seed=0, in_channels=1):
        self.L = _lib.load()
        h = C.c_void_p()
        _check(self.L, self.L.unet_create(C.byref(h), int(channels), int(in_channels), int(seed)),
               "unet_create")
        self.h = h
        self.channels = channels
        self.in_channels = in_channels

    @property
    def n_params(self):
        return int(self.L.unet_param_count(self.h))

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(self.L, self.L.unet_get_params(self.h, _p(p)), "unet_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(self.L, self.L.unet_set_params(self.h, _p(p)), "unet_set_params")

    def _in(self, x):
        """[N][D][H][W] (one input channel) or [N][D][H][W][in_channels]."""
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim == 4:
            x = x[..., None]
        assert x.ndim == 5 and x.shape[-1] == self.in_channels, "inputs are [N][D][H][W][C]"
        return x

    @staticmethod
    def _vol(t):
        t = np.ascontiguousarray(t, np.float32)
        assert t.ndim == 4, "volumes are [N][D][H][W]"
        return t

    def forward(self, x):
        x = self._in(x)
        y = np.empty(x.shape[:4], np.float32)
        _check(self.L, self.L.unet_forward(self.h, *x.shape[:4], _p(x), _p(y)), "unet_forward")
        return y

    def gradients(self, x, t):
        x, t = self._in(x), self._vol(t)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check(self.L, self.L.unet_gradients(self.h, *x.shape[:4], _p(x), _p(t), _p(g),
                                             C.byref(loss)), "unet_gradients")
        return loss.value, g

    def train_step(self, x, t, lr):
        x, t = self._in(x), self._vol(t)
        loss = C.c_double()
        _check(self.L, self.L.unet_train_step(self.h, *x.shape[:4], _p(x), _p(t), float(lr),
                                              C.byref(loss)), "unet_train_step")
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            self.L.unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def converge(stage, equits, alpha_seed=761521, refresh=10):
    """Runs `equits` equits of the parallel ICD on a stage whose object is
    already set (A built here), refreshing the error sinogram every
    `refresh` equits; returns the object."""
    stage.build_amatrix()
    stage.init_error()
    for it in range(1, equits + 1):
        stage.icd(1, it, alpha_seed)
        if refresh and it % refresh == 0:
            stage.init_error()
    return stage.get_object()


IN_CHANNELS = {"fbp3": 3, "fbp2": 2, "fbp": 1, "bp": 1}


def network_input(stage, source="fbp2"):
    """The network's input [N_time][N_z][N_y][N_x] (one channel) or
    [...][2] ("fbp2"); A must be built."""
    if source in ("fbp2", "fbp3"):
        f = stage.fbp()
        chans = [f]
        if source == "fbp3":
            # mean over the frame and its two neighbours (edges clamped)
            T = f.shape[0]
            lo = f[[max(i - 1, 0) for i in range(T)]]
            hi = f[[min(i + 1, T - 1) for i in range(T)]]
            chans.append((lo + f + hi) / 3.0)
        chans.append(np.broadcast_to(f.mean(axis=0), f.shape))
        return np.stack(chans, axis=-1)
    if source == "fbp":
        return stage.fbp()
    if source == "bp":
        return stage.bp_normalized()
    raise ValueError(f"unknown network input {source!r}")


def training_pair(scan, equits=400, tile=8, alpha=1.9, source="fbp2"):
    """(network input, converged image) of one scan, both
    [N_time][N_z][N_y][N_x], from the GPU MBIR started at zero."""
    from .mbir import Stage
    st = Stage(scan, tile=tile, alpha=alpha)
    try:
        st.set_object(np.zeros(st.shape_x, np.float32))
        x = converge(st, equits)
        bp = network_input(st, source)
    finally:
        st.close()
    return bp, x


def converged_pair(scan, coarse_equits=(60, 100), full_equits=200, alpha=1.9, tile=8,
                   coarse_tile=4, source="fbp2"):
    """(network input, image) of one scan at full resolution
    after the multi-resolution schedule: the coarse stages from zero, each
    up-sampled into the next, then `full_equits` at full resolution."""
    from .mbir import Stage, multires_schedule
    stages = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    prev = None
    made = []
    try:
        for (mxy, mz), n in zip(stages[:-1], coarse_equits):
            cs = Stage(scan, mult_xy=mxy, mult_z=mz, tile=coarse_tile, alpha=alpha)
            made.append(cs)
            if prev is None:
                cs.set_object(np.zeros(cs.shape_x, np.float32))
            else:
                cs.upsample_from(prev, masked=True)
            converge(cs, n)
            prev = cs
        st = Stage(scan, tile=tile, alpha=alpha)
        made.append(st)
        if prev is None:
            st.set_object(np.zeros(st.shape_x, np.float32))
        else:
            st.upsample_from(prev, masked=True)
        x = converge(st, full_equits)
        bp = network_input(st, source)
    finally:
        for m in made:
            m.close()
    return bp, x


def random_patches(pairs, n, size, rng):
    """n patches [n][D][size][size] of (input, target) drawn from the pairs'
    frames (all of D, a size x size window in y, x)."""
    xs, ts = [], []
    for _ in range(n):
        bp, x = pairs[rng.integers(len(pairs))]
        f = rng.integers(bp.shape[0])
        y0 = rng.integers(bp.shape[2] - size + 1)
        x0 = rng.integers(bp.shape[3] - size + 1)
        xs.append(bp[f, :, y0:y0 + size, x0:x0 + size])
        ts.append(x[f, :, y0:y0 + size, x0:x0 + size])
    return (np.stack(xs) * SCALE).astype(np.float32), (np.stack(ts) * SCALE).astype(np.float32)


def train(net, pairs, steps, batch=8, size=64, lr=1e-3, seed=0, log=None):
    """Adam on random patches; returns the loss history."""
    rng = np.random.default_rng(seed)
    hist = []
    for s in range(steps):
        xb, tb = random_patches(pairs, batch, size, rng)
        hist.append(net.train_step(xb, tb, lr if s < steps * 3 // 4 else lr * 0.3))
        if log and (s % 100 == 0 or s == steps - 1):
            log(f"step {s}: loss {hist[-1]:.4g}")
    return np.array(hist)


def _down2(x, zf):
    """Mean over 2 x 2 (x 2 when zf == 2) blocks of [T][Z][Y][X]."""
    T, Z, Y, X = x.shape
    if zf == 2:
        return x.reshape(T, Z // 2, 2, Y // 2, 2, X // 2, 2).mean(axis=(2, 4, 6))
    return x.reshape(T, Z, Y // 2, 2, X // 2, 2).mean(axis=(3, 5))


def _up_axis(c, axis):
    """2x along `axis` by the multi-resolution rule: the ends copy, interior
    fine samples weigh the two nearest coarse samples 3/4 and 1/4."""
    n = c.shape[axis]
    i = np.arange(2 * n)
    k = i >> 1
    other = np.where(i & 1, k + 1, k - 1)
    w1 = np.where((i == 0) | (i == 2 * n - 1), 0.0, 0.25)
    other = np.clip(other, 0, n - 1)
    a = np.take(c, k, axis=axis)
    b = np.take(c, other, axis=axis)
    shape = [1] * c.ndim
    shape[axis] = 2 * n
    w1 = w1.reshape(shape)
    return (1.0 - w1) * a + w1 * b


def _up2(x, zf):
    y = _up_axis(_up_axis(x, 2), 3)
    return _up_axis(y, 1) if zf == 2 else y


def learned_start(stage, net, source="fbp2", coarse=None, coarse_equits=0, seed=761521):
    """Sets the stage's object to the network's estimate from the stage's
    network input (A must be built): clipped at zero and zero outside the
    update mask.  With a coarse stage (x and y halved, z halved or kept) and
    coarse_equits > 0, the estimate's low frequencies are corrected first:
    the coarse stage runs that many equits from the block mean of the
    estimate, and the up-sampled change is added to the estimate (a
    two-grid correction)."""
    y = net.forward(network_input(stage, source) * SCALE) / SCALE
    m = stage.mask[None, None].astype(np.float32)
    y = np.maximum(y, 0.0) * m
    if coarse is not None and coarse_equits > 0:
        zf = stage.g.N_z // coarse.g.N_z
        yd = _down2(y, zf).astype(np.float32)
        coarse.set_object(yd * coarse.mask[None, None].astype(np.float32))
        xc = converge(coarse, coarse_equits, seed)
        y = np.maximum(y + _up2(xc - yd, zf), 0.0) * m
    stage.set_object(y.astype(np.float32))
    return y
