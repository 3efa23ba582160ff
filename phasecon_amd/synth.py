"""Seeded synthetic scans for the MBIR workloads (numpy only, deterministic).

The reference ships no sinograms (its example blobs are missing and its
attenuation simulators do not compile), so every workload is generated here:

* view schedules: uniform over pi, or the time-interlaced rule of
  reconstruct/real_data/APS_Data_Format/XT_Views.c:7-46;
* phantoms: the modified Shepp-Logan ellipse table, and a two-phase
  Cahn-Hilliard microstructure that coarsens over time;
* measurements: a pixel-driven projector (linear splat onto detector bins,
  2x2 sub-pixels) that is deliberately *not* the reference system matrix,
  followed by Poisson counts.  projections = ln(I0 / counts), weights =
  counts, as Examples/shepp-logan-3D/gen_data_shepplogan.m:26-39 does.

Arrays are returned in the reference's file layout: projections/weights are
[view][row][col] float32 (col fastest), angles/times float32.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "uniform_views", "interlaced_views", "shepp_logan", "cahn_hilliard",
    "project", "make_scan", "WORKLOADS",
]


def uniform_views(n):
    """n angles k*pi/n with times k (Examples/shepp-logan-3D layout)."""
    k = np.arange(n, dtype=np.float32)
    return (k * np.float32(np.pi / n)).astype(np.float32), k


def interlaced_views(K, n_theta, n_views, proj_start=0):
    """Time-interlaced views, restating XT_Views.c:7-46 with min_acq_time=0.

    View i has angle (i*K + bitrev_log2K((i // L) % K)) * pi / n_theta with
    L = n_theta // K and time angle_index / K, computed in float32 as the
    reference does.
    """
    L = n_theta // K
    bits = int(np.floor(np.log(K) / np.log(2.0) + 0.5))
    dtheta = np.float32(np.pi) / np.float32(n_theta)
    n = proj_start + n_views
    i = np.arange(n)
    sub = (i // L) % K
    rev = np.zeros(n, dtype=np.int64)
    for j in range(bits):
        rev |= ((sub >> j) & 1) << (bits - 1 - j)
    idx = (i * K + rev).astype(np.float32)
    times = (idx / np.float32(K)).astype(np.float32)
    angles = (idx * dtheta).astype(np.float32)
    return angles[proj_start:], times[proj_start:]


# Modified Shepp-Logan (Toft): value, semi-axes a, b, centre x, y, angle(deg)
_SHEPP = np.array([
    [1.0, .69, .92, 0, 0, 0],
    [-.8, .6624, .8740, 0, -.0184, 0],
    [-.2, .1100, .3100, .22, 0, -18],
    [-.2, .1600, .4100, -.22, 0, 18],
    [.1, .2100, .2500, 0, .35, 0],
    [.1, .0460, .0460, 0, .1, 0],
    [.1, .0460, .0460, 0, -.1, 0],
    [.1, .0460, .0230, -.08, -.605, 0],
    [.1, .0230, .0230, 0, -.606, 0],
    [.1, .0230, .0460, .06, -.605, 0],
])


def shepp_logan(n, supersample=4):
    """n x n modified Shepp-Logan image in [0, 1], row index = y (downward)."""
    s = supersample
    c = (np.arange(n * s) + 0.5) / (n * s) * 2.0 - 1.0
    X, Y = np.meshgrid(c, -c)
    img = np.zeros_like(X)
    for v, a, b, x0, y0, ph in _SHEPP:
        t = np.deg2rad(ph)
        xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += v
    img = img.reshape(n, s, n, s).mean(axis=(1, 3))
    return np.clip(img, 0.0, None)


def cahn_hilliard(n, nz, nt, seed, lo=0.005, hi=0.02, steps0=60, steps_per_frame=6):
    """Two-phase field coarsening over nt frames: [nt][nz][n][n] float32.

    Semi-implicit spectral Cahn-Hilliard, c_t = lap(c^3 - c - g lap c), on a
    periodic n x n x nz grid from seeded spinodal noise; each frame is the
    field thresholded at 0 into {lo, hi}, inside a disc of radius 0.45 n so
    the object stays within the reconstruction circle.
    """
    rng = np.random.default_rng(seed)
    c = 0.05 * rng.standard_normal((nz, n, n))
    kz = 2 * np.pi * np.fft.fftfreq(nz)
    ky = 2 * np.pi * np.fft.fftfreq(n)
    kx = 2 * np.pi * np.fft.fftfreq(n)
    K2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    gamma, dt = 1.0, 0.5
    den = 1.0 + dt * gamma * K2 * K2

    def step(c):
        ch = np.fft.fftn(c)
        nl = np.fft.fftn(c ** 3 - c)
        return np.real(np.fft.ifftn((ch - dt * K2 * nl) / den))

    for _ in range(steps0):
        c = step(c)
    yy, xx = np.mgrid[0:n, 0:n]
    disc = ((xx + 0.5 - n / 2) ** 2 + (yy + 0.5 - n / 2) ** 2) < (0.45 * n) ** 2
    out = np.zeros((nt, nz, n, n), np.float32)
    for f in range(nt):
        out[f] = np.where(c > 0, hi, lo) * disc[None]
        for _ in range(steps_per_frame):
            c = step(c)
    return out


def project(x, angles, frame_of_view, vox_wid, rot_center):
    """Pixel-driven projection of x [nt][nz][n][n] -> [view][row][col].

    Each pixel is split into 2x2 sub-pixels whose line integral
    (value * area / detector pitch) is split linearly between the two nearest
    detector bins; the geometry (R0 = -rot_center*pitch, y0 = -n*pitch/2)
    matches XT_Init.c:318-360 at full resolution.
    """
    nt, nz, n, _ = x.shape
    pitch = float(vox_wid)
    R0 = -float(rot_center) * pitch
    y0 = -n * pitch / 2.0
    sub = (np.arange(2) + 0.5) / 2.0
    cx = R0 + (np.arange(n)[:, None] + sub[None, :]).ravel() * pitch
    cy = y0 + (np.arange(n)[:, None] + sub[None, :]).ravel() * pitch
    XX, YY = np.meshgrid(cx, cy)            # [2n][2n], row = y
    w_area = (pitch / 2) ** 2 / pitch
    out = np.zeros((len(angles), nz, n), np.float64)
    for v, th in enumerate(np.asarray(angles, np.float64)):
        f = frame_of_view[v]
        if f < 0:
            continue
        u = (XX * np.cos(th) - YY * np.sin(th) - R0) / pitch - 0.5
        j0 = np.floor(u).astype(np.int64)
        fr = u - j0
        img = np.repeat(np.repeat(x[f], 2, axis=1), 2, axis=2) * w_area  # [nz][2n][2n]
        for jj, wt in ((j0, 1.0 - fr), (j0 + 1, fr)):
            ok = (jj >= 0) & (jj < n)
            idx = jj[ok]
            for z in range(nz):
                out[v, z] += np.bincount(idx, weights=(img[z][ok] * wt[ok]), minlength=n)
    return out.astype(np.float32)


# Workload table.  sig/c values are fixed here once for CPU and GPU alike.
WORKLOADS = {
    # Plain 3D: 128x128x4, 180 views over 180 deg, I0 = 1e4.
    "shepp3d": dict(rows=4, cols=128, views=180, frames=1, I0=1e4, K=None,
                    sig_s=0.005, sig_t=1.0, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.2),
    # Plain 4D: 256x256x8 x 16 frames, 256 views interlaced K=4, I0 = 5e3.
    "ch4d": dict(rows=8, cols=256, views=256, frames=16, I0=5e3, K=4,
                 sig_s=0.01, sig_t=0.001, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.1),
    # Small 4D case sized for CPU parity tests (same rule as ch4d).
    "ch4d_small": dict(rows=4, cols=64, views=64, frames=4, I0=5e3, K=4,
                       sig_s=0.01, sig_t=0.001, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.2),
    # Tiny 3D case for unit parity tests.
    "tiny3d": dict(rows=4, cols=32, views=24, frames=1, I0=1e4, K=None,
                   sig_s=0.005, sig_t=1.0, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.2),
    # Tall 3D case: 16 detector rows and 24 views per frame, so a column's
    # sinogram window is not one load batch per lane (the sweep's
    # non-pipelined theta path).
    "tall3d": dict(rows=16, cols=32, views=24, frames=1, I0=1e4, K=None,
                   sig_s=0.005, sig_t=1.0, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.2),
    # The mixed-precision workload's detector height (32 rows) and views per
    # frame (30) on a small grid, for parity of its coarse stages (window
    # widths up to 17, up to 8 rows per voxel).
    "cfg4_small": dict(rows=32, cols=64, views=60, frames=2, I0=5e3, K=None,
                       sig_s=0.01, sig_t=0.001, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.2),
}


def make_scan(name, seed=0, **over):
    """Builds a workload: dict with the reconstruct() inputs and parameters."""
    w = dict(WORKLOADS[name])
    w.update(over)
    rows, cols, nv, nf = w["rows"], w["cols"], w["views"], w["frames"]
    vox_wid, rot_center = 1.0, cols / 2.0
    if w["K"] is None:
        angles, times = uniform_views(nv)
        recon_times = np.array([0.0, float(nv)], np.float32)
        if nf != 1:
            recon_times = np.linspace(0, nv, nf + 1).astype(np.float32)
    else:
        angles, times = interlaced_views(w["K"], nv, nv)
        recon_times = np.linspace(0.0, float(nv), nf + 1).astype(np.float32)
    frame_of_view = np.full(nv, -1, np.int64)
    for f in range(nf):
        sel = (times >= recon_times[f]) & (times < recon_times[f + 1])
        frame_of_view[sel] = f
    if name.startswith("ch4d"):
        phantom = cahn_hilliard(cols, rows, nf, seed)
    else:
        sl = shepp_logan(cols).astype(np.float32) * np.float32(0.02)
        phantom = np.broadcast_to(sl, (nf, rows, cols, cols)).astype(np.float32)
    line = project(phantom, angles, frame_of_view, vox_wid, rot_center)
    rng = np.random.default_rng(seed + 1000)
    counts = rng.poisson(w["I0"] * np.exp(-line.astype(np.float64))).astype(np.float64)
    counts = np.maximum(counts, 1.0)
    projections = np.log(w["I0"] / counts).astype(np.float32)
    weights = counts.astype(np.float32)
    return dict(
        name=name, projections=projections, weights=weights, angles=angles,
        proj_times=times, recon_times=recon_times, proj_rows=rows,
        proj_cols=cols, proj_num=nv, recon_num=nf, vox_wid=vox_wid,
        rot_center=rot_center, sig_s=w["sig_s"], sig_t=w["sig_t"],
        c_s=w["c_s"], c_t=w["c_t"], p_s=w["p_s"], p_t=w["p_t"],
        phantom=phantom, seed=seed,
    )


def geometry_scan(rows, cols, views, seed=0):
    """A parallel-beam 3D scan's geometry and reconstruct() parameters with
    no measurement (projections 0, weights 1): views uniform over 180
    degrees, one frame, a rows x cols detector and a cols^2 x rows grid.
    For the PCMBIR tomographic step, whose targets come from the phase
    retrieval rather than from the scan's projections."""
    angles, times = uniform_views(views)
    return dict(
        name="geometry", projections=np.zeros((views, rows, cols), np.float32),
        weights=np.ones((views, rows, cols), np.float32), angles=angles, proj_times=times,
        recon_times=np.array([0.0, float(views)], np.float32), proj_rows=rows, proj_cols=cols,
        proj_num=views, recon_num=1, vox_wid=1.0, rot_center=cols / 2.0, sig_s=0.005,
        sig_t=1.0, c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.2, phantom=None, seed=seed,
    )


def drifting_ellipsoids(n, nz, nt, seed, lo=0.005, hi=0.02, count=40):
    """[nt][nz][n][n] two-phase phantom: value lo inside a disk of radius
    0.85 (of the half-width), hi inside seeded ellipsoids that drift a little
    from frame to frame, 0 outside the disk."""
    rng = np.random.default_rng(seed)
    c0 = rng.uniform(-0.55, 0.55, (count, 2))
    rxy = rng.uniform(0.04, 0.14, (count, 2))
    zc = rng.uniform(0, nz, count)
    rz = rng.uniform(nz / 6, nz / 2, count)
    vel = rng.normal(0.0, 0.004, (count, 2))
    ang = rng.uniform(0, np.pi, count)
    u = (np.arange(n) + 0.5) / n * 2 - 1
    yy, xx = np.meshgrid(u, u, indexing="ij")
    disk = (xx ** 2 + yy ** 2) <= 0.85 ** 2
    z = np.arange(nz) + 0.5
    out = np.zeros((nt, nz, n, n), np.float32)
    for t in range(nt):
        inside = np.zeros((nz, n, n), bool)
        for e in range(count):
            cx, cy = c0[e] + vel[e] * t
            ca, sa = np.cos(ang[e]), np.sin(ang[e])
            dx, dy = xx - cx, yy - cy
            q = ((ca * dx + sa * dy) / rxy[e, 0]) ** 2 + ((-sa * dx + ca * dy) / rxy[e, 1]) ** 2
            dz2 = ((z - zc[e]) / rz[e]) ** 2
            inside |= (q[None] + dz2[:, None, None]) <= 1.0
        out[t] = np.where(disk[None], np.where(inside, hi, lo), 0.0)
    return out


def uneven_frames(n_views, n_frames, lo, hi, seed):
    """Seeded views per frame in [lo, hi] summing to n_views."""
    rng = np.random.default_rng(seed)
    k = rng.integers(lo, hi + 1, n_frames)
    while k.sum() != n_views:
        i = rng.integers(n_frames)
        if k.sum() > n_views and k[i] > lo:
            k[i] -= 1
        elif k.sum() < n_views and k[i] < hi:
            k[i] += 1
    return k


def make_cfg4_scan(seed=0, cols=512, rows=32, views=720, frames=24, K=8, I0=5e3):
    """The mixed-precision 4D workload: cols x cols x rows voxels x frames,
    `views` time-interlaced views (K-way), seeded uneven views per frame in
    [28, 32], a drifting-ellipsoid phantom projected by the GPU forward
    projector (the CPU projector is too slow at this size), Poisson counts."""
    from .mbir import Stage
    angles, times = interlaced_views(K, views, views)
    per = uneven_frames(views, frames, 28, 32, seed + 77)
    first = np.concatenate([[0], np.cumsum(per)[:-1]])
    recon_times = np.concatenate([times[first], [float(views)]]).astype(np.float32)
    recon_times[0] = 0.0
    phantom = drifting_ellipsoids(cols, rows, frames, seed)
    scan = dict(
        name="cfg4", projections=np.zeros((views, rows, cols), np.float32),
        weights=np.ones((views, rows, cols), np.float32), angles=angles, proj_times=times,
        recon_times=recon_times, proj_rows=rows, proj_cols=cols, proj_num=views,
        recon_num=frames, vox_wid=1.0, rot_center=cols / 2.0, sig_s=0.01, sig_t=0.001,
        c_s=0.01, c_t=0.01, p_s=1.2, p_t=1.1, seed=seed, views_per_frame=per)
    st = Stage(scan)
    try:
        line = st.forward_project(phantom)            # [N_p][N_r][N_t]
    finally:
        st.close()
    line = np.ascontiguousarray(line.transpose(0, 2, 1)).astype(np.float64)
    rng = np.random.default_rng(seed + 1000)
    counts = np.maximum(rng.poisson(I0 * np.exp(-line)).astype(np.float64), 1.0)
    scan["projections"] = np.log(I0 / counts).astype(np.float32)
    scan["weights"] = counts.astype(np.float32)
    scan["phantom"] = phantom
    return scan

