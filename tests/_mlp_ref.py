"""numpy restatement of the learned-bound MLP (test infrastructure).

Same parameter layout and the same BF16 rounding points as the device:
GEMM operands (features, W1, W2, relu(Z1), dZ2, dZ1) are rounded to BF16
(nearest even); everything else is float32 (here accumulated in float64 and
compared with a tolerance)."""
import numpy as np

F, H = 64, 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_PARAMS = OFF_B3 + 1


def bf16_bits(x):
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f32(h):
    return (np.asarray(h, np.uint32) << 16).view(np.float32)


def rbf(x):
    return bf16_to_f32(bf16_bits(x)).astype(np.float64)


F3 = 96  # the 3D network's feature width (82 used)


def layout(n_params):
    """Offsets (w1, b1, w2, b2, w3, b3) and feature width of a parameter block."""
    f = (n_params - 1 - 3 * H - H * H) // H
    assert f * H + 3 * H + H * H + 1 == n_params
    w1, b1 = 0, f * H
    w2 = b1 + H
    b2 = w2 + H * H
    w3 = b2 + H
    return (w1, b1, w2, b2, w3, w3 + H), f


def unpack(p):
    p = np.asarray(p, np.float64)
    (w1, b1, w2, b2, w3, b3), f = layout(p.size)
    return (p[w1:b1].reshape(f, H), p[b1:w2], p[w2:b2].reshape(H, H), p[b2:w3], p[w3:b3], p[b3])


def forward(p, Xbits, bf=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = rbf if bf else (lambda a: np.asarray(a, np.float64))
    X = bf16_to_f32(Xbits).astype(np.float64)
    Z1 = X @ r(W1) + b1
    H1 = r(np.maximum(Z1, 0))
    Z2 = H1 @ r(W2) + b2
    H2 = np.maximum(Z2, 0)
    y = H2 @ w3 + b3
    return X, Z1, H1, Z2, H2, y


def grad(p, Xbits, T, bf=True):
    """(gradient in the parameter layout, mean squared error)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = rbf if bf else (lambda a: np.asarray(a, np.float64))
    X, Z1, H1, Z2, H2, y = forward(p, Xbits, bf)
    m = len(T)
    d = y - np.asarray(T, np.float64)
    loss = float(np.mean(d * d))
    dy = 2 * d / m
    (ow1, ob1, ow2, ob2, ow3, ob3), _ = layout(np.size(p))
    g = np.zeros(np.size(p))
    g[ow3:ob3] = H2.T @ dy
    g[ob3] = dy.sum()
    dZ2 = dy[:, None] * w3[None, :] * (Z2 > 0)
    g[ob2:ow3] = dZ2.sum(0)
    dZ2r = r(dZ2)
    g[ow2:ob2] = (H1.T @ dZ2r).ravel()
    dH1 = dZ2r @ r(W2).T
    dZ1 = dH1 * (Z1 > 0)
    g[ob1:ow2] = dZ1.sum(0)
    g[ow1:ob1] = (X.T @ r(dZ1)).ravel()
    return g, loss
