"""Pins the U-Net oracle (oracle/unet_oracle.py) on the CPU: the convolution
and its gradients against torch.nn.functional.conv3d and autograd in float64,
the whole network (without BF16 storage) against the same network written in
torch, pooling / up-sampling against closed forms and adjointness, and the
BF16 rounding against torch's bfloat16 conversion."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet_oracle as U


def _to_torch_w(w):
    # [cout][27][cin] -> [cout][cin][3][3][3] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1))
    co, _, ci = w.shape
    return torch.tensor(w.reshape(co, 3, 3, 3, ci).transpose(0, 4, 1, 2, 3).copy())


def _to_torch_x(x):
    return torch.tensor(np.ascontiguousarray(x.transpose(0, 4, 1, 2, 3)))


def _from_torch_x(t):
    return t.detach().numpy().transpose(0, 2, 3, 4, 1)


@pytest.mark.parametrize("cin,cout", [(1, 16), (5, 3), (48, 16)])
def test_conv3d_and_gradients_match_torch(cin, cout):
    rng = np.random.default_rng(cin * 7 + cout)
    x = rng.standard_normal((2, 4, 6, 5, cin))
    w = rng.standard_normal((cout, 27, cin))
    b = rng.standard_normal(cout)
    g = rng.standard_normal((2, 4, 6, 5, cout))
    tx = _to_torch_x(x).requires_grad_(True)
    tw = _to_torch_w(w).requires_grad_(True)
    ty = F.conv3d(tx, tw, torch.tensor(b), padding=1)
    ty.backward(_to_torch_x(g))
    y = U.conv3d(x, w, b)
    assert np.allclose(y, _from_torch_x(ty), rtol=1e-12, atol=1e-10)
    assert np.allclose(U.conv3d_bwd_data(g, w), _from_torch_x(tx.grad), rtol=1e-12, atol=1e-10)
    dw = U.conv3d_bwd_weight(g, x)
    tdw = tw.grad.numpy().transpose(0, 2, 3, 4, 1).reshape(cout, 27, cin)
    assert np.allclose(dw, tdw, rtol=1e-12, atol=1e-9)


def test_pool_and_upsample():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 4, 6, 3))
    p = U.avgpool2(x)
    assert np.isclose(p[0, 1, 0, 2, 1], x[0, 2:4, 0:2, 4:6, 1].mean())
    u = U.up2(p)
    assert u.shape == x.shape and u[0, 3, 1, 5, 2] == p[0, 1, 0, 2, 2]
    g = rng.standard_normal(p.shape)
    gx = rng.standard_normal(x.shape)
    # adjoint identities <pool x, g> = <x, pool^T g>, <up p, gx> = <p, up^T gx>
    assert np.isclose(np.sum(p * g), np.sum(x * U.avgpool2_bwd(g)))
    assert np.isclose(np.sum(u * gx), np.sum(p * U.up2_bwd(gx)))


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(4)
    a = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-20, 20, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38]]).astype(np.float32)
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(U.bf16(a), ref)


def _torch_unet(params, x, t, c):
    relu = torch.relu
    P = [(torch.tensor(w.reshape(w.shape[0], 3, 3, 3, w.shape[2]).transpose(0, 4, 1, 2, 3).copy(),
                       requires_grad=True), torch.tensor(b, requires_grad=True)) for w, b in params]
    conv = lambda a, i: F.conv3d(a, P[i][0], P[i][1], padding=1)
    X = torch.tensor(x)[:, None]
    A1 = relu(conv(X, 0))
    S1 = relu(conv(A1, 1))
    A2 = relu(conv(F.avg_pool3d(S1, 2), 2))
    S2 = relu(conv(A2, 3))
    A3 = relu(conv(F.avg_pool3d(S2, 2), 4))
    B3 = relu(conv(A3, 5))
    up = lambda a: F.interpolate(a, scale_factor=2, mode="nearest")
    A4 = relu(conv(torch.cat([up(B3), S2], 1), 6))
    B4 = relu(conv(A4, 7))
    A5 = relu(conv(torch.cat([up(B4), S1], 1), 8))
    B5 = relu(conv(A5, 9))
    y = conv(B5, 10)[:, 0] + torch.tensor(x)
    loss = torch.mean((y - torch.tensor(t)) ** 2)
    loss.backward()
    grads = []
    for w, b in P:
        g = w.grad.numpy()  # [cout][cin][3][3][3]
        grads.append((g.transpose(0, 2, 3, 4, 1).reshape(g.shape[0], 27, g.shape[1]), b.grad.numpy()))
    return float(loss.detach()), grads, y.detach().numpy()


def test_unet_matches_torch_autograd():
    c = 8
    rng = np.random.default_rng(5)
    params = []
    for cin, cout, _ in U.layer_shapes(c):
        params.append((rng.standard_normal((cout, 27, cin)) * np.sqrt(2.0 / (27 * cin)),
                       rng.standard_normal(cout) * 0.1))
    x = rng.random((2, 4, 8, 8))
    t = x + 0.1 * rng.standard_normal(x.shape)
    loss, G, y = U.unet_forward_backward(params, x, t, c, bf16_storage=False)
    tl, tG, ty = _torch_unet(params, x, t, c)
    assert np.isclose(loss, tl, rtol=1e-12)
    assert np.allclose(y, ty, rtol=1e-11, atol=1e-12)
    for (dw, db), (tdw, tdb) in zip(G, tG):
        assert np.allclose(dw, tdw, rtol=1e-9, atol=1e-13)
        assert np.allclose(db, tdb, rtol=1e-9, atol=1e-13)


def test_bf16_storage_changes_little():
    """With BF16 storage the loss moves by BF16-sized relative amounts only."""
    c = 8
    rng = np.random.default_rng(6)
    params = [(rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / (27 * ci)), np.zeros(co))
              for ci, co, _ in U.layer_shapes(c)]
    x = rng.random((1, 4, 8, 8))
    t = x * 0.9
    l64, G64, _ = U.unet_forward_backward(params, x, t, c, bf16_storage=False)
    l16, G16, _ = U.unet_forward_backward(params, x, t, c, bf16_storage=True)
    assert abs(l16 - l64) <= 2e-2 * l64
    f64 = np.concatenate([np.concatenate([a.ravel(), b]) for a, b in G64])
    f16 = np.concatenate([np.concatenate([a.ravel(), b]) for a, b in G16])
    assert np.linalg.norm(f16 - f64) <= 5e-2 * np.linalg.norm(f64)
