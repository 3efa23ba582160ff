"""CPU oracle of the learned warm start's U-Net (include/unet_hip.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

The network has no counterpart in the reference (it is the learned warm
start of this library); the oracle restates its definition in numpy
float64:

  * conv3d / conv3d_bwd_data / conv3d_bwd_weight: 3x3x3 cross-correlation
    with zero padding, weights [cout][27][cin], tap = (dz+1)*9 + (dy+1)*3 +
    (dx+1); activations [N][D][H][W][C];
  * avgpool2 / up2 and their adjoints;
  * unet_forward_backward: the 11-layer network, loss mean((x_0 + net(x) -
    t)^2) (x_0 the first input channel) and the gradient of every parameter.  With bf16=True every value
    the GPU stores in BF16 (the input, the packed weights, every activation
    and activation gradient) is rounded to BF16 (round to nearest even) at
    the same point, so the GPU's FP32 accumulation is the only difference.

Pins (tests/test_unet_oracle.py): conv3d and its two gradients against
torch.nn.functional.conv3d and autograd (CPU, float64); the whole network
with bf16=False against the same network written with torch autograd;
avgpool2 / up2 against their closed forms and the adjoint identity; the BF16
rounding against torch's bfloat16 conversion.
"""
from __future__ import annotations

import numpy as np

LAYERS = ("enc1a", "enc1b", "enc2a", "enc2b", "bott_a", "bott_b", "dec2a", "dec2b",
          "dec1a", "dec1b", "out")


def layer_shapes(c, in_channels=1):
    """(cin, cout, relu) of the 11 layers for base channels c."""
    c1, c2, c3 = c, 2 * c, 4 * c
    return [(in_channels, c1, 1), (c1, c1, 1), (c1, c2, 1), (c2, c2, 1), (c2, c3, 1), (c3, c3, 1),
            (c3 + c2, c2, 1), (c2, c2, 1), (c2 + c1, c1, 1), (c1, c1, 1), (c1, 1, 0)]


def unflatten(flat, c, in_channels=1):
    out, i = [], 0
    for cin, cout, _ in layer_shapes(c, in_channels):
        nw = cout * 27 * cin
        w = np.asarray(flat[i:i + nw], np.float64).reshape(cout, 27, cin)
        i += nw
        b = np.asarray(flat[i:i + cout], np.float64)
        i += cout
        out.append((w, b))
    assert i == len(flat)
    return out


def flatten(params):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in params])


def bf16(a):
    """Round to BF16 (nearest even) and back, through float32."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _taps():
    for tap in range(27):
        yield tap, tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1


def conv3d(x, w, b=None):
    N, D, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((N, D, H, W, w.shape[0]))
    for tap, dz, dy, dx in _taps():
        sl = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]
        out += sl @ w[:, tap, :].T
    if b is not None:
        out += b
    return out


def conv3d_bwd_data(g, w):
    """Gradient with respect to the input of conv3d for output gradient g."""
    N, D, H, W, _ = g.shape
    gp = np.pad(g, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((N, D, H, W, w.shape[2]))
    for tap, dz, dy, dx in _taps():
        sl = gp[:, 1 - dz:1 - dz + D, 1 - dy:1 - dy + H, 1 - dx:1 - dx + W, :]
        out += sl @ w[:, tap, :]
    return out


def conv3d_bwd_weight(g, x):
    """dW[co][tap][ci] = sum_v g[v][co] x[v + d_tap][ci]."""
    N, D, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    g2 = g.reshape(-1, g.shape[-1])
    dw = np.zeros((g.shape[-1], 27, x.shape[-1]))
    for tap, dz, dy, dx in _taps():
        sl = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]
        dw[:, tap, :] = g2.T @ sl.reshape(-1, x.shape[-1])
    return dw


def avgpool2(x):
    N, D, H, W, C = x.shape
    return x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def avgpool2_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up2_bwd(g):
    N, D, H, W, C = g.shape
    return g.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def unet_forward_backward(params, x, t, c, bf16_storage=True, grads=True):
    """x: [N][D][H][W] or [N][D][H][W][in_channels], t: [N][D][H][W].
    Returns (loss, grads as a list of (dw, db), y) with y = x_0 + net(x),
    x_0 the first input channel."""
    r = bf16 if bf16_storage else (lambda a: np.asarray(a, np.float64))
    relu = lambda a: np.maximum(a, 0.0)
    W_ = [(r(w), b) for w, b in params]
    x = np.asarray(x, np.float64)
    if x.ndim == 4:
        x = x[..., None]
    X0 = r(x)
    A1 = r(relu(conv3d(X0, *W_[0])))
    S1 = r(relu(conv3d(A1, *W_[1])))
    P1 = r(avgpool2(S1))
    A2 = r(relu(conv3d(P1, *W_[2])))
    S2 = r(relu(conv3d(A2, *W_[3])))
    P2 = r(avgpool2(S2))
    A3 = r(relu(conv3d(P2, *W_[4])))
    B3 = r(relu(conv3d(A3, *W_[5])))
    C2 = np.concatenate([up2(B3), S2], axis=-1)
    A4 = r(relu(conv3d(C2, *W_[6])))
    B4 = r(relu(conv3d(A4, *W_[7])))
    C1 = np.concatenate([up2(B4), S1], axis=-1)
    A5 = r(relu(conv3d(C1, *W_[8])))
    B5 = r(relu(conv3d(A5, *W_[9])))
    net = conv3d(B5, *W_[10])[..., 0]
    y = net + x[..., 0]
    d = y - np.asarray(t, np.float64)
    loss = float(np.mean(d * d))
    if not grads:
        return loss, None, y
    V = d.size
    G = [None] * 11

    def layer(i, g, act_out, inp, need_in=True):
        if act_out is not None:
            g = g * (act_out > 0)
        G[i] = (conv3d_bwd_weight(g, inp), g.reshape(-1, g.shape[-1]).sum(axis=0))
        return r(conv3d_bwd_data(g, W_[i][0])) if need_in else None

    GO = r(2.0 * d / V)[..., None]
    GB5 = layer(10, GO, None, B5)
    GA5 = layer(9, GB5, B5, A5)
    GC1 = layer(8, GA5, A5, C1)
    c2 = 2 * c
    GB4 = r(up2_bwd(GC1[..., :c2]))
    GA4 = layer(7, GB4, B4, A4)
    GC2 = layer(6, GA4, A4, C2)
    c3 = 4 * c
    GB3 = r(up2_bwd(GC2[..., :c3]))
    GA3 = layer(5, GB3, B3, A3)
    GP2 = layer(4, GA3, A3, P2)
    GS2 = r(GC2[..., c3:] + avgpool2_bwd(GP2))
    GA2 = layer(3, GS2, S2, A2)
    GP1 = layer(2, GA2, A2, P1)
    GS1 = r(GC1[..., c2:] + avgpool2_bwd(GP1))
    GA1 = layer(1, GS1, S1, A1)
    layer(0, GA1, A1, X0, need_in=False)
    return loss, G, y
