"""What the learned warm start gains: a U-Net trained on GPU-MBIR pairs of
four other scans starts the solver on a held-out scan, which then reaches
the same MAP cost target in fewer equits than from the plain (zero)
initialiser.  The target is (1 + 1e-5) C*, C* from a 600-equit run of the
same solver (over-relaxation 1.9, 1.8e-9 of the FP64 oracle's C* after 120
equits on this workload)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _equits_to(st, target, max_equits, seed=17):
    st.build_amatrix()
    st.init_error()
    for it in range(1, max_equits + 1):
        st.icd(1, it, seed)
        if it % 10 == 0:
            st.init_error()
        if st.cost(estimate=True)[0] <= target and st.cost(fresh=True)[0] <= target:
            return it
    return max_equits + 1


def test_bp_normalized_formula(small4d_scan):
    """The network's input equals A^T(w y) / A^T(w A m) built from the
    projector pair (each checked against the oracle), m the update mask."""
    from phasecon_amd.mbir import Stage
    st = Stage(small4d_scan)
    st.build_amatrix()
    bp = st.bp_normalized().astype(np.float64)
    m = np.broadcast_to(st.mask[None, None].astype(np.float32), st.shape_x).copy()
    am = st.forward_project(m)
    num = st.backproject((st.w * st.y).astype(np.float32)).astype(np.float64)
    den = st.backproject((st.w * am).astype(np.float32)).astype(np.float64)
    st.close()
    ref = np.where(den > 0, num / np.where(den > 0, den, 1), 0)
    inside = m > 0
    print(f"bp range {bp[inside].min():.4g} .. {bp[inside].max():.4g}", flush=True)
    assert np.max(np.abs(bp - ref)[inside]) <= 1e-5 * np.max(np.abs(ref[inside]))


def _ramlak(n):
    k = np.arange(-(n - 1), n)
    h = np.zeros(k.size)
    h[k == 0] = 0.25
    odd = (k % 2) != 0
    h[odd] = -1.0 / (np.pi * k[odd]) ** 2
    return h


def _ramp(v):
    """Ram-Lak filter along the channel axis r of a [p][r][t] sinogram."""
    h = _ramlak(v.shape[1])
    out = np.zeros(v.shape)
    for p in range(v.shape[0]):
        for t in range(v.shape[2]):
            full = np.convolve(v[p, :, t].astype(np.float64), h)
            out[p, :, t] = full[v.shape[1] - 1:2 * v.shape[1] - 1]
    return out


def test_fbp_formula_and_uniform_object(small4d_scan):
    """fbp = A^T(h * y) / K with h the Ram-Lak filter (numpy convolution)
    and K the mask mean of A^T(h * A m); an exact scan of an object uniform
    inside the mask reconstructs to a mask mean equal to its value."""
    from phasecon_amd.mbir import Stage
    st = Stage(small4d_scan)
    st.build_amatrix()
    fbp = st.fbp().astype(np.float64)
    m = np.broadcast_to(st.mask[None, None].astype(np.float32), st.shape_x).copy()
    am = st.forward_project(m)
    K = st.backproject(_ramp(am).astype(np.float32))[m > 0].astype(np.float64).mean()
    ref = st.backproject(_ramp(st.y).astype(np.float32)).astype(np.float64) / K
    inside = m > 0
    err = np.max(np.abs(fbp - ref)[inside]) / np.max(np.abs(ref[inside]))
    print(f"fbp vs numpy ramp: {err:.3g}; range {fbp[inside].min():.4g} .. "
          f"{fbp[inside].max():.4g}", flush=True)
    assert err <= 1e-4
    # exact projections of 0.013 m: the mask mean of the reconstruction is 0.013
    scan = dict(small4d_scan)
    c = 0.013
    N_p, N_t, N_r = scan["proj_num"], scan["proj_rows"], scan["proj_cols"]
    scan["projections"] = np.ascontiguousarray(
        (c * am).reshape(N_p, N_r, N_t).transpose(0, 2, 1)).astype(np.float32)
    su = Stage(scan)
    su.build_amatrix()
    fu = su.fbp().astype(np.float64)
    su.close()
    st.close()
    print(f"uniform {c}: mask mean {fu[inside].mean():.6g}", flush=True)
    assert abs(fu[inside].mean() - c) <= 1e-5 * c


def test_learned_start_saves_equits(small4d_scan):
    from phasecon_amd.mbir import Stage
    from phasecon_amd.synth import make_scan
    from phasecon_amd.warmstart import UNet, learned_start, train, training_pair
    pairs = [training_pair(make_scan("ch4d_small", seed=s), equits=300) for s in (10, 11, 12, 13)]
    net = UNet(channels=16, seed=0, in_channels=2)
    hist = train(net, pairs, steps=300, batch=8, size=64, lr=1e-3, seed=0)
    print(f"training loss {hist[:10].mean():.4g} -> {hist[-10:].mean():.4g}", flush=True)
    assert hist[-10:].mean() < 0.5 * hist[:10].mean()

    st = Stage(small4d_scan, tile=8, alpha=1.9)
    st.set_object(np.zeros(st.shape_x, np.float32))
    st.build_amatrix()
    st.init_error()
    for it in range(1, 601):
        st.icd(1, it, 5)
        if it % 10 == 0:
            st.init_error()
    cstar = st.cost(fresh=True)[0]
    target = cstar * (1 + 1e-5)

    st.set_object(np.zeros(st.shape_x, np.float32))
    plain = _equits_to(st, target, 300)
    st.build_amatrix()
    y = learned_start(st, net)
    learned = _equits_to(st, target, 300)
    x_end = st.get_object()
    net.close()
    st.close()
    print(f"equits to (1+1e-5) C*: plain {plain}, learned {learned}; "
          f"start RMSE {np.sqrt(np.mean((y - x_end) ** 2)) / np.sqrt(np.mean(x_end ** 2)):.3g}",
          flush=True)
    assert learned < plain
