"""The benchmark configurations at their full sizes, on the GPU.

* Batched uneven 2D: the 96 fields (h, w log-uniform in [256, 4096], seeded;
  220M vertices) in one batched compress and one batched decompress: streams
  byte-equal to the oracle's on a deterministic subset, the critical-point set
  of every decoded field equal to the original's (GPU judge), and every
  decoded value within its vertex's analytic bound.
* Learned 2D: the network trained on 64 seeded 1024x1024 fields, then the
  2400x3600 field compressed with learned bounds: 0 critical-point errors, a
  higher ratio than the plain method, values within the learned bounds.
* Learned + batched 3D: 16 timesteps of 192^3, plain batched streams byte-equal
  to the oracle's on two timesteps, then the learned batch: 0 critical-point
  errors on every timestep and a higher ratio than the plain batch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import cpsz_amd as cz
from _libs import load_oracle, mask_2d_padding
from cpsz_amd.synth import batch_2d_shapes, field_2d, field_2d_device, field_3d_device

pytestmark = pytest.mark.gpu
HIP = C.CDLL("libamdhip64.so")


def _host(ptr, n):
    out = np.empty(n, np.uint8)
    torch.cuda.synchronize()
    HIP.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(n), 2)
    return out


def _cp_errors(c0, c1):
    return (sum(1 for k in c1 if k not in c0) + sum(1 for k in c0 if k not in c1) +
            sum(1 for k in c0 if k in c1 and c0[k] != c1[k]))


def _zstd_bytes(allb, off, size):
    return sum(cz.sz_lossless_compress(allb[o:o + n], 3).size for o, n in zip(off, size))


def test_config2_batched_96_fields():
    shapes = batch_2d_shapes(96, seed=0)
    nv = sum(r1 * r2 for r1, r2 in shapes)
    print(f"[config 2] 96 fields, {nv} vertices", flush=True)
    assert nv == 220143645
    dev = [field_2d_device(r1, r2, seed=1000 + k) for k, (r1, r2) in enumerate(shapes)]
    outs = [(torch.empty_like(u), torch.empty_like(v)) for u, v in dev]
    torch.cuda.synchronize()  # the library runs on its own stream
    base, off, size = cz.compress_2d_batch_dev([(u.data_ptr(), v.data_ptr(), r1, r2)
                                                for (u, v), (r1, r2) in zip(dev, shapes)], 0.1)
    allb = _host(base, int(off[-1] + size[-1]))
    cz.decompress_2d_batch_dev([(base + int(o), int(n)) for o, n in zip(off, size)], shapes,
                               [(x.data_ptr(), y.data_ptr()) for x, y in outs])
    torch.cuda.synchronize()
    orc = load_oracle()
    # oracle streams on every 12th field (8 fields of all sizes)
    for k in range(0, 96, 12):
        r1, r2 = shapes[k]
        U, V = dev[k][0].cpu().numpy(), dev[k][1].cpu().numpy()
        ref = orc.compress_2d(U, V, 0.1)
        s = allb[off[k]:off[k] + size[k]]
        assert np.array_equal(mask_2d_padding(s, r1, r2), mask_2d_padding(ref, r1, r2)), (k, r1, r2)
    print("[config 2] oracle streams equal on 8 fields", flush=True)
    total_cp = 0
    for k, ((u, v), (x, y), (r1, r2)) in enumerate(zip(dev, outs, shapes)):
        c0 = cz.critical_points_2d_dev(u.data_ptr(), v.data_ptr(), r1, r2)
        c1 = cz.critical_points_2d_dev(x.data_ptr(), y.data_ptr(), r1, r2)
        assert _cp_errors(c0, c1) == 0, k
        total_cp += len(c0)
        # every value within its vertex's analytic bound: |x' - x| <= E |x|. The
        # stream's precision 4^id eps may exceed E |x| by a few ulps where libm's
        # log2 rounds eb/eps up to a power of 4 (the reference's own id), hence
        # the relative slack of 1e-9
        E = torch.empty(r1 * r2, dtype=torch.float64, device="cuda")
        cz.learn_analytic_eb_2d_dev(u.data_ptr(), v.data_ptr(), r1, r2, 0.1, E.data_ptr())
        torch.cuda.synchronize()  # the library ran on its context's own stream
        Ev = E.view(r1, r2)
        for a, b in ((u, x), (v, y)):
            a64, b64 = a.double(), b.double()
            err, lim = (b64 - a64).abs(), Ev * a64.abs()
            if not bool(torch.all(err <= lim * (1 + 1e-9))):
                bad = err > lim * (1 + 1e-9)
                i = int(torch.nonzero(bad.flatten())[0])
                print(f"[config 2] field {k}: {int(bad.sum())} values beyond the bound; first at {i}: "
                      f"x={a64.flatten()[i].item()!r} x'={b64.flatten()[i].item()!r} E={Ev.flatten()[i].item()!r}",
                      flush=True)
            assert bool(torch.all(err <= lim * (1 + 1e-9))), k
    print(f"[config 2] 0 critical-point errors over {total_cp} critical points", flush=True)
    assert total_cp > 1000


@pytest.fixture(scope="module")
def learned2d_params():
    train = [field_2d_device(1024, 1024, seed=5000 + k) for k in range(64)]
    torch.cuda.synchronize()  # the library runs on its own stream
    p, losses = cz.learn_train_2d_dev([(u.data_ptr(), v.data_ptr(), 1024, 1024) for u, v in train], 0.1,
                                      cz.learn_init_params(1), samples_per_field=16384, epochs=8, batch=4096,
                                      lr=2e-3, seed=11)
    print(f"[config 3] loss per epoch {[round(float(x), 4) for x in losses]}", flush=True)
    return p


def test_config3_learned_2400x3600(learned2d_params):
    p = learned2d_params
    U, V = field_2d(2400, 3600, seed=0)
    du, dv = torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda()
    ptr, n, it, boosted, fb = cz.compress_2d_learned_dev(du.data_ptr(), dv.data_ptr(), 2400, 3600, 0.1, p)
    learned = _host(ptr, n)
    print(f"[config 3] learned stream {n} B, {it} rounds, {boosted} boosted, {fb} fell back", flush=True)
    plain = cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
    r_l = 2 * U.nbytes / cz.sz_lossless_compress(learned, 3).size
    r_p = 2 * U.nbytes / cz.sz_lossless_compress(plain, 3).size
    print(f"[config 3] ratio learned {r_l:.3f} plain {r_p:.3f}", flush=True)
    assert r_l > 1.5 * r_p
    ou, ov = torch.empty_like(du), torch.empty_like(dv)
    dst = torch.empty(n, dtype=torch.uint8, device="cuda")
    dst.copy_(torch.from_numpy(learned))
    cz.decompress_2d_dev(dst.data_ptr(), n, 2400, 3600, ou.data_ptr(), ov.data_ptr())
    c0 = cz.critical_points_2d_dev(du.data_ptr(), dv.data_ptr(), 2400, 3600)
    c1 = cz.critical_points_2d_dev(ou.data_ptr(), ov.data_ptr(), 2400, 3600)
    assert len(c0) > 0 and _cp_errors(c0, c1) == 0
    # every value within the largest bound the proposer may choose (max_pwr_eb)
    for a, b in ((du, ou), (dv, ov)):
        a64, b64 = a.double(), b.double()
        assert bool(torch.all((b64 - a64).abs() <= 0.1 * a64.abs()))


def test_config4_learned_batched_3d_16x192():
    nt, shape, eb = 16, (192, 192, 192), 0.05
    dev = [field_3d_device(*shape, seed=0, step=t) for t in range(nt)]
    fl = [(u.data_ptr(), v.data_ptr(), w.data_ptr(), *shape) for u, v, w in dev]
    outs = [tuple(torch.empty_like(x) for x in f) for f in dev]
    torch.cuda.synchronize()  # the library runs on its own stream
    base, off, size = cz.compress_3d_batch_dev(fl, eb)
    allb = _host(base, int(off[-1] + size[-1]))
    orc = load_oracle()
    for t in (0, nt - 1):
        U, V, W = (x.cpu().numpy() for x in dev[t])
        ref = orc.compress_3d(U, V, W, eb)
        assert np.array_equal(allb[off[t]:off[t] + size[t]], ref), t
    print("[config 4] plain batched streams equal the oracle's on timesteps 0 and 15", flush=True)
    z_plain = _zstd_bytes(allb, off, size)
    train = [field_3d_device(128, 128, 128, seed=7000 + k) for k in range(8)]
    torch.cuda.synchronize()
    p, losses = cz.learn_train_3d_dev([(u.data_ptr(), v.data_ptr(), w.data_ptr(), 128, 128, 128)
                                       for u, v, w in train], eb, cz.learn_init_params_3d(1),
                                      samples_per_field=16384, epochs=8, batch=4096, lr=2e-3, seed=11)
    del train
    lb, lo, ls, it, boosted, fb = cz.compress_3d_learned_batch_dev(fl, eb, p)
    lall = _host(lb, int(lo[-1] + ls[-1]))
    z_learned = _zstd_bytes(lall, lo, ls)
    print(f"[config 4] ratio learned {12 * nt * 192 ** 3 / z_learned:.3f} plain {12 * nt * 192 ** 3 / z_plain:.3f}, "
          f"{it} rounds, {boosted} boosted, {fb} fell back", flush=True)
    assert z_learned < z_plain
    cz.decompress_3d_batch_dev([(lb + int(o), int(n)) for o, n in zip(lo, ls)], [shape] * nt,
                               [tuple(x.data_ptr() for x in o) for o in outs])
    torch.cuda.synchronize()
    total = 0
    for f, o in zip(dev, outs):
        c0 = cz.critical_points_3d_dev(*(x.data_ptr() for x in f), *shape)
        c1 = cz.critical_points_3d_dev(*(x.data_ptr() for x in o), *shape)
        assert _cp_errors(c0, c1) == 0
        total += len(c0)
        for a, b in zip(f, o):  # the log codec keeps signs and relative errors below max_pwr_eb
            a64, b64 = a.double(), b.double()
            assert bool(torch.all(a64 * b64 >= 0))
            assert bool(torch.all((b64 - a64).abs() <= eb * a64.abs() * (1 + 1e-5)))
    print(f"[config 4] 0 critical-point errors over {total} critical points", flush=True)
    assert total > 0
