"""3D learned-bound proposer on the GPU: analytic bounds, features and labels
against the oracle, the 3D MLP's forward and gradient against the numpy
restatement, training, and the batched learned compressor: streams the plain
decoder (and the oracle's) reads, every field with the originals' critical
points, and a higher compression ratio than the plain method."""
import ctypes as C

import numpy as np
import pytest
import torch

import _mlp_ref as M
import cpsz_amd as cz
from _libs import load_oracle
from cpsz_amd.synth import field_3d

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_analytic_eb_features_labels_3d_match_oracle(orc):
    shape = (20, 22, 24)
    print(f"[learn3d] features/labels {shape}", flush=True)
    U, V, W = field_3d(*shape, 5)
    eb = 0.05
    n = U.size
    dU, dV, dW = _dev(U), _dev(V), _dev(W)
    dE = torch.empty(n, dtype=torch.float64, device="cuda")
    cz.learn_analytic_eb_3d_dev(dU.data_ptr(), dV.data_ptr(), dW.data_ptr(), *shape, eb, dE.data_ptr())
    torch.cuda.synchronize()
    E = dE.cpu().numpy().reshape(shape)
    assert np.array_equal(E.view(np.uint64), orc.eb_3d(U, V, W, eb).view(np.uint64))
    dX = torch.empty(n * 96, dtype=torch.int16, device="cuda")
    cz.learn_features_3d_dev(dU.data_ptr(), dV.data_ptr(), dW.data_ptr(), dE.data_ptr(), *shape, eb, 0, n,
                             dX.data_ptr())
    torch.cuda.synchronize()
    X = dX.cpu().numpy().view(np.uint16).reshape(n, 96)
    Xo = orc.learn_features_3d(U, V, W, E, eb)
    assert np.array_equal(X[:, :81], Xo[:, :81]) and np.array_equal(X[:, 82:], Xo[:, 82:])
    # log2 of the bound: device and host libm may differ in the last bit of the double
    assert np.abs(X[:, 81].astype(np.int32) - Xo[:, 81].astype(np.int32)).max() <= 1
    for cap in (eb, 0.5):
        E2 = orc.eb_3d(U, V, W, cap)
        dE2 = _dev(E2.ravel())
        idx = np.arange(0, n, 7, dtype=np.uint32)
        didx = _dev(idx.view(np.int32))
        dL = torch.empty(idx.size, dtype=torch.uint8, device="cuda")
        cz.learn_labels_3d_dev(dU.data_ptr(), dV.data_ptr(), dW.data_ptr(), dE2.data_ptr(), *shape, cap,
                               didx.data_ptr(), idx.size, dL.data_ptr())
        torch.cuda.synchronize()
        L = dL.cpu().numpy()
        assert np.array_equal(L, orc.learn_labels_3d(U, V, W, E2, cap, idx)), cap
        print(f"[learn3d] cap {cap}: labels histogram {np.bincount(L)[:12]}", flush=True)


def test_mlp3_forward_and_gradient_match_numpy():
    rng = np.random.default_rng(4)
    p = cz.learn_init_params_3d(3)
    (w1, b1, w2, b2, w3, b3), f = M.layout(p.size)
    assert f == 96
    p[b1:w2] = rng.standard_normal(128) * 0.1
    p[b3] = 0.5
    m = 1024
    Xf = rng.standard_normal((m, 96)).astype(np.float32)
    Xf[:, 82:] = 0
    Xb = M.bf16_bits(Xf)
    T = rng.uniform(0, 6, m).astype(np.float32)
    dX, dT = _dev(Xb.view(np.int16)), _dev(T)
    dY = torch.empty(m, dtype=torch.float32, device="cuda")
    cz.mlp_forward_3d_dev(p, dX.data_ptr(), m, dY.data_ptr())
    torch.cuda.synchronize()
    y_ref = M.forward(p.astype(np.float64), Xb)[-1]
    assert np.allclose(dY.cpu().numpy(), y_ref, rtol=2e-3, atol=2e-3)
    g, loss = cz.mlp_grad_3d_dev(p, dX.data_ptr(), dT.data_ptr(), m)
    g_ref, loss_ref = M.grad(p.astype(np.float64), Xb, T)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    for a, b, name in [(w1, b1, "W1"), (b1, w2, "b1"), (w2, b2, "W2"), (b2, w3, "b2"), (w3, b3, "w3"),
                       (b3, p.size, "b3")]:
        err = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-12)
        print(f"[grad3] {name} rel err {err:.2e}", flush=True)
        assert err < 2e-2, name


@pytest.fixture(scope="module")
def trained3():
    fields = [field_3d(48, 48, 48, 200 + s) for s in range(4)]
    dev = [tuple(_dev(x) for x in f) for f in fields]
    p0 = cz.learn_init_params_3d(7)
    p, losses = cz.learn_train_3d_dev([(u.data_ptr(), v.data_ptr(), w.data_ptr(), 48, 48, 48) for u, v, w in dev],
                                      0.05, p0, samples_per_field=8192, epochs=6, batch=2048, lr=2e-3, seed=5)
    print(f"[train3] loss per epoch {losses}", flush=True)
    return p, losses


def test_training3_reduces_loss(trained3):
    _, losses = trained3
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.9 * losses[0]


def test_learned_batch3d_keeps_critical_points(orc, trained3):
    p, _ = trained3
    shapes = [(40, 44, 48), (32, 40, 36), (40, 44, 48)]
    fields = [field_3d(*s, 300, step=k) for k, s in enumerate(shapes)]  # advected timesteps
    dev = [tuple(_dev(x) for x in f) for f in fields]
    base, off, size, it, boosted, fb = cz.compress_3d_learned_batch_dev(
        [(u.data_ptr(), v.data_ptr(), w.data_ptr(), *s) for (u, v, w), s in zip(dev, shapes)], 0.05, p)
    print(f"[learned3] sizes {size.tolist()} rounds {it} boosted {boosted} fell back {fb}", flush=True)
    allb = np.empty(int(off[-1] + size[-1]), np.uint8)
    torch.cuda.synchronize()
    C.CDLL("libamdhip64.so").hipMemcpy(C.c_void_p(allb.ctypes.data), C.c_void_p(base), C.c_size_t(allb.size), 2)
    lossy_l = lossy_p = 0
    for k, ((U, V, W), s) in enumerate(zip(fields, shapes)):
        stream = allb[off[k]:off[k] + size[k]]
        d = cz.sz_decompress_cp_preserve_3d_online_log(stream, *s)
        o = orc.decompress_3d(stream, *s)
        for a, b in zip(d, o):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert orc.cp_3d(U, V, W) == orc.cp_3d(*d), s
        for a, b in zip((U, V, W), d):  # every value within the cap: |x'| / |x| in [1/1.05, 1.05], sign kept
            a64, b64 = a.astype(np.float64), b.astype(np.float64)
            assert np.all(np.abs(b64) <= np.abs(a64) * 1.05 * (1 + 1e-6))
            assert np.all(np.abs(b64) >= np.abs(a64) / 1.05 * (1 - 1e-6))
            assert np.all(np.sign(a64) * np.sign(b64) >= 0)
        lossy_l += cz.sz_lossless_compress(stream, 3).size
        lossy_p += cz.sz_lossless_compress(cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05), 3).size
    print(f"[learned3] zstd bytes learned {lossy_l} plain {lossy_p}", flush=True)
    assert lossy_l < lossy_p
