"""Learned warm start of the MBIR solver (include/unet_hip.h, libmbir4d.so).

The U-Net maps the normalised back-projection of a scan
(``Stage.bp_normalized``: per voxel sum(w a (y - d)) / sum(w a^2)) to an
estimate of the converged MBIR image; the solver starts its full-resolution
stage from that estimate (clipped at zero, zero outside the update mask) and
still runs to convergence, so the network changes how much work the solver
does, never its fixed point.  Training pairs come from the GPU MBIR itself:
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

    def __init__(self, channels=16, seed=0):
        self.L = _lib.load()
        h = C.c_void_p()
        _check(self.L, self.L.unet_create(C.byref(h), int(channels), int(seed)), "unet_create")
        self.h = h
        self.channels = channels

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

    @staticmethod
    def _vol(x):
        x = np.ascontiguousarray(x, np.float32)
        assert x.ndim == 4, "volumes are [N][D][H][W]"
        return x

    def forward(self, x):
        x = self._vol(x)
        y = np.empty_like(x)
        _check(self.L, self.L.unet_forward(self.h, *x.shape, _p(x), _p(y)), "unet_forward")
        return y

    def gradients(self, x, t):
        x, t = self._vol(x), self._vol(t)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check(self.L, self.L.unet_gradients(self.h, *x.shape, _p(x), _p(t), _p(g),
                                             C.byref(loss)), "unet_gradients")
        return loss.value, g

    def train_step(self, x, t, lr):
        x, t = self._vol(x), self._vol(t)
        loss = C.c_double()
        _check(self.L, self.L.unet_train_step(self.h, *x.shape, _p(x), _p(t), float(lr),
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


def training_pair(scan, equits=400, tile=8, alpha=1.9):
    """(normalised back-projection, converged image) of one scan, both
    [N_time][N_z][N_y][N_x], from the GPU MBIR started at zero."""
    from .mbir import Stage
    st = Stage(scan, tile=tile, alpha=alpha)
    try:
        st.set_object(np.zeros(st.shape_x, np.float32))
        x = converge(st, equits)
        bp = st.bp_normalized()
    finally:
        st.close()
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


def learned_start(stage, net):
    """Sets the stage's object to the network's estimate from the stage's
    normalised back-projection (A must be built): clipped at zero and zero
    outside the update mask."""
    bp = stage.bp_normalized()
    y = net.forward(bp.reshape(-1, *bp.shape[1:]) * SCALE) / SCALE
    y = np.maximum(y, 0.0) * stage.mask[None, None].astype(np.float32)
    stage.set_object(y.astype(np.float32))
    return y
