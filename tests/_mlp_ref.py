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


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


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
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = H2.T @ dy
    g[OFF_B3] = dy.sum()
    dZ2 = dy[:, None] * w3[None, :] * (Z2 > 0)
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    dZ2r = r(dZ2)
    g[OFF_W2:OFF_B2] = (H1.T @ dZ2r).ravel()
    dH1 = dZ2r @ r(W2).T
    dZ1 = dH1 * (Z1 > 0)
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    g[OFF_W1:OFF_B1] = (X.T @ r(dZ1)).ravel()
    return g, loss
