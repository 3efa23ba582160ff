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


def test_learned_start_saves_equits(small4d_scan):
    from phasecon_amd.mbir import Stage
    from phasecon_amd.synth import make_scan
    from phasecon_amd.warmstart import UNet, learned_start, train, training_pair
    pairs = [training_pair(make_scan("ch4d_small", seed=s), equits=300) for s in (10, 11, 12, 13)]
    net = UNet(channels=16, seed=0)
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
