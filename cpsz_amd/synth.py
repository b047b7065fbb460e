"""Seeded synthetic vector fields for the benchmark workloads and tests.

All fields are generated on the host in float64 with numpy's PCG64 generator
and rounded once to float32, so every consumer (GPU path, CPU oracle, the
reference build) receives identical bytes for a given (shape, seed).
"""
from __future__ import annotations

import math

import numpy as np


def field_2d(r1: int, r2: int, seed: int = 0, n_vortex: int = 64, n_modes: int = 8):
    """2D u,v (row-major r1 x r2): Gaussian vortices + Fourier modes + noise.

    64 vortices: centre uniform on the grid, radius U[20,200] cells,
    circulation U[-1,1], stream function psi = G exp(-rho^2 / 2R^2),
    u = dpsi/dy, v = -dpsi/dx. 8 Fourier modes per component with amplitude
    0.1, integer wavenumbers in {0..4}^2 (not both zero), random phase.
    N(0, 1e-3) noise per value.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    y = np.arange(r1, dtype=np.float64)[:, None]
    x = np.arange(r2, dtype=np.float64)[None, :]
    u = np.zeros((r1, r2))
    v = np.zeros((r1, r2))
    for _ in range(n_vortex):
        cx = rng.uniform(0, r2)
        cy = rng.uniform(0, r1)
        R = rng.uniform(20, 200)
        G = rng.uniform(-1, 1)
        dx = x - cx
        dy = y - cy
        e = G * np.exp(-(dx * dx + dy * dy) / (2 * R * R)) / (R * R)
        u += -dy * e
        v += dx * e
    for comp in (u, v):
        for _ in range(n_modes):
            while True:
                kx, ky = rng.integers(0, 5, size=2)
                if kx or ky:
                    break
            ph = rng.uniform(0, 2 * math.pi)
            comp += 0.1 * np.cos(2 * math.pi * (kx * x / r2 + ky * y / r1) + ph)
    u += rng.normal(0, 1e-3, size=(r1, r2))
    v += rng.normal(0, 1e-3, size=(r1, r2))
    return u.astype(np.float32), v.astype(np.float32)


def field_2d_device(r1: int, r2: int, seed: int = 0, device="cuda", n_vortex: int = 64, n_modes: int = 8):
    """field_2d's recipe evaluated on the GPU with torch in float64 (parameters
    drawn on the host from the same PCG64 stream, noise from a seeded torch
    generator), for batches too large to generate on the host. Returns two
    float32 device tensors. Deterministic per (shape, seed, device type), not
    byte-identical to field_2d (different noise generator)."""
    import torch

    rng = np.random.Generator(np.random.PCG64(seed))
    y = torch.arange(r1, dtype=torch.float64, device=device)[:, None]
    x = torch.arange(r2, dtype=torch.float64, device=device)[None, :]
    u = torch.zeros((r1, r2), dtype=torch.float64, device=device)
    v = torch.zeros_like(u)
    for _ in range(n_vortex):
        cx, cy = rng.uniform(0, r2), rng.uniform(0, r1)
        R, G = rng.uniform(20, 200), rng.uniform(-1, 1)
        dx, dy = x - cx, y - cy
        e = G * torch.exp(-(dx * dx + dy * dy) / (2 * R * R)) / (R * R)
        u -= dy * e
        v += dx * e
    for comp in (u, v):
        for _ in range(n_modes):
            while True:
                kx, ky = rng.integers(0, 5, size=2)
                if kx or ky:
                    break
            ph = rng.uniform(0, 2 * math.pi)
            comp += 0.1 * torch.cos(2 * math.pi * (float(kx) * x / r2 + float(ky) * y / r1) + ph)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    u += 1e-3 * torch.randn((r1, r2), dtype=torch.float64, device=device, generator=g)
    v += 1e-3 * torch.randn((r1, r2), dtype=torch.float64, device=device, generator=g)
    return u.float().contiguous(), v.float().contiguous()


def batch_2d_shapes(count: int = 96, seed: int = 0, lo: int = 256, hi: int = 4096):
    """Heights and widths drawn log-uniform in [lo, hi] (the batched 2D workload)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hw = np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), size=(count, 2)))).astype(int)
    return [(int(h), int(w)) for h, w in hw]


def field_3d(r1: int, r2: int, r3: int, seed: int = 0, n_tubes: int = 32, step: int = 0):
    """3D u,v,w (r1 slowest .. r3 fastest): ABC flow + vortex tubes + noise.

    ABC: u = A sin Z + C cos Y, v = B sin X + A cos Z, w = C sin Y + B cos X
    with X,Y,Z = 4 pi (k, j, i) / n (two periods per axis), A,B,C ~ U[0.5,1.5].
    32 Gaussian vortex tubes: random unit axis a, centre c, radius U[4,24],
    circulation U[-1,1]; velocity G exp(-d^2 / 2R^2) (a x (p - c)) / R with d
    the distance to the axis. ``step`` advects the tube centres by 2 cells per
    step along one seeded direction (time series). N(0, 1e-3) noise.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    A, B, C = rng.uniform(0.5, 1.5, size=3)
    i = np.arange(r1, dtype=np.float64)[:, None, None]
    j = np.arange(r2, dtype=np.float64)[None, :, None]
    k = np.arange(r3, dtype=np.float64)[None, None, :]
    X = 4 * math.pi * k / r3
    Y = 4 * math.pi * j / r2
    Z = 4 * math.pi * i / r1
    shape = (r1, r2, r3)
    u = np.broadcast_to(A * np.sin(Z) + C * np.cos(Y), shape).copy()
    v = np.broadcast_to(B * np.sin(X) + A * np.cos(Z), shape).copy()
    w = np.broadcast_to(C * np.sin(Y) + B * np.cos(X), shape).copy()
    drift = rng.normal(size=3)
    drift /= np.linalg.norm(drift)
    for _ in range(n_tubes):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        c = rng.uniform(0, 1, size=3) * np.array([r3, r2, r1]) + 2.0 * step * drift
        R = rng.uniform(4, 24)
        G = rng.uniform(-1, 1)
        px = k - c[0]
        py = j - c[1]
        pz = i - c[2]
        dot = px * a[0] + py * a[1] + pz * a[2]
        d2 = px * px + py * py + pz * pz - dot * dot
        g = G * np.exp(-d2 / (2 * R * R)) / R
        u += g * (a[1] * pz - a[2] * py)
        v += g * (a[2] * px - a[0] * pz)
        w += g * (a[0] * py - a[1] * px)
    u += rng.normal(0, 1e-3, size=shape)
    v += rng.normal(0, 1e-3, size=shape)
    w += rng.normal(0, 1e-3, size=shape)
    return u.astype(np.float32), v.astype(np.float32), w.astype(np.float32)


def field_3d_device(r1: int, r2: int, r3: int, seed: int = 0, n_tubes: int = 32, step: int = 0, device="cuda"):
    """field_3d's recipe evaluated on the GPU with torch in float64 (parameters
    drawn on the host from the same PCG64 stream, noise from a seeded torch
    generator), for timestep batches too large to generate on the host.
    Returns three float32 device tensors. Deterministic per (shape, seed,
    step, device type), not byte-identical to field_3d (different noise)."""
    import torch

    rng = np.random.Generator(np.random.PCG64(seed))
    A, B, C = rng.uniform(0.5, 1.5, size=3)
    f64 = dict(dtype=torch.float64, device=device)
    i = torch.arange(r1, **f64)[:, None, None]
    j = torch.arange(r2, **f64)[None, :, None]
    k = torch.arange(r3, **f64)[None, None, :]
    X = 4 * math.pi * k / r3
    Y = 4 * math.pi * j / r2
    Z = 4 * math.pi * i / r1
    shape = (r1, r2, r3)
    u = (A * torch.sin(Z) + C * torch.cos(Y)).expand(shape).clone()
    v = (B * torch.sin(X) + A * torch.cos(Z)).expand(shape).clone()
    w = (C * torch.sin(Y) + B * torch.cos(X)).expand(shape).clone()
    drift = rng.normal(size=3)
    drift /= np.linalg.norm(drift)
    for _ in range(n_tubes):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        c = rng.uniform(0, 1, size=3) * np.array([r3, r2, r1]) + 2.0 * step * drift
        R = rng.uniform(4, 24)
        G = rng.uniform(-1, 1)
        px, py, pz = k - float(c[0]), j - float(c[1]), i - float(c[2])
        dot = px * float(a[0]) + py * float(a[1]) + pz * float(a[2])
        d2 = px * px + py * py + pz * pz - dot * dot
        g = G * torch.exp(-d2 / (2 * R * R)) / R
        u += g * (float(a[1]) * pz - float(a[2]) * py)
        v += g * (float(a[2]) * px - float(a[0]) * pz)
        w += g * (float(a[0]) * py - float(a[1]) * px)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed * 1000003 + step)
    for comp in (u, v, w):
        comp += 1e-3 * torch.randn(shape, generator=gen, **f64)
    return u.float().contiguous(), v.float().contiguous(), w.float().contiguous()


def batch_shapes_2d(count: int = 96, seed: int = 0, lo: int = 256, hi: int = 4096):
    """(h, w) pairs drawn log-uniform in [lo, hi] (the batched 2D workload)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        h = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        w = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        out.append((h, w))
    return out
