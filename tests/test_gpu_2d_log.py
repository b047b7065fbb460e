"""2D offline_log codec on the GPU against the compiled reference's fixtures
and the oracle: streams byte-equal (the reference's three uninitialised
meanInfo padding bytes masked), decoded fields bit-equal, at the golden
shapes, random shapes, edge shapes, fields with exact zeros and at full size."""
import glob
import os

import numpy as np
import pytest
import torch

import cpsz_amd as cz
from _libs import load_oracle, mask_2d_log_padding
from cpsz_amd.synth import field_2d

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _bits(a):
    return a.view(np.uint32)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "c2l_*.npz"))), ids=os.path.basename)
def test_log_golden(path):
    z = np.load(path)
    U, V = z["U"], z["V"]
    r1, r2 = U.shape
    print(f"[2d log] golden {os.path.basename(path)}", flush=True)
    s = cz.sz_compress_cp_preserve_2d_offline_log(U, V, float(z["eb"]))
    assert np.array_equal(mask_2d_log_padding(s, r1, r2), z["stream"])
    dU, dV = cz.sz_decompress_cp_preserve_2d_offline_log(z["stream"], r1, r2)
    assert np.array_equal(_bits(dU), _bits(z["dU"]))
    assert np.array_equal(_bits(dV), _bits(z["dV"]))


@pytest.mark.parametrize("shape,seed,eb,zeros", [((1, 70), 1, 0.1, False), ((70, 1), 2, 0.1, False),
                                                 ((2, 2), 3, 0.1, False), ((65, 129), 4, 0.05, True),
                                                 ((300, 97), 5, 0.3, True), ((128, 256), 6, 0.01, False),
                                                 ((257, 64), 7, 0.1, True)])
def test_log_vs_oracle(orc, shape, seed, eb, zeros):
    r1, r2 = shape
    print(f"[2d log] {shape} seed {seed} eb {eb} zeros {zeros}", flush=True)
    U, V = field_2d(r1, r2, seed)
    if zeros:
        rng = np.random.default_rng(seed)
        U[rng.random(U.shape) < 0.03] = 0
        V[rng.random(V.shape) < 0.03] = 0
    s = cz.sz_compress_cp_preserve_2d_offline_log(U, V, eb)
    ref = orc.compress_2d_log(U, V, eb)
    assert np.array_equal(mask_2d_log_padding(s, r1, r2), mask_2d_log_padding(ref, r1, r2))
    dU, dV = cz.sz_decompress_cp_preserve_2d_offline_log(ref, r1, r2)
    oU, oV = orc.decompress_2d_log(ref, r1, r2)
    assert np.array_equal(_bits(dU), _bits(oU))
    assert np.array_equal(_bits(dV), _bits(oV))
    # signs kept, zeros exact, relative error within the log-domain bound (plus
    # the float roundings of log2f and exp2f, a few 1e-7 relative)
    for a, b in ((U, dU), (V, dV)):
        assert np.all((a == 0) == (b == 0)) and np.all(np.sign(a) == np.sign(b))
        nz = a != 0
        assert np.all(np.abs(b[nz].astype(np.float64) / a[nz] - 1) <= eb + 2e-6)


def test_log_full_size_device(orc):
    """2400x3600 through the device entry points; stream equal to the oracle's and
    0 critical-point errors after the round trip."""
    U, V = field_2d(2400, 3600, seed=0)
    du, dv = torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda()
    import ctypes as C
    p, n = C.c_void_p(), C.c_size_t()
    assert cz.lib().cpsz_compress_2d_log_f32_dev(0, None, du.data_ptr(), dv.data_ptr(), 2400, 3600, 0.1,
                                                 C.byref(p), C.byref(n)) == 0
    s = torch.empty(n.value, dtype=torch.uint8, device="cuda")
    C.CDLL("libamdhip64.so").hipMemcpy(C.c_void_p(s.data_ptr()), p, n, 3)
    ou, ov = torch.empty_like(du), torch.empty_like(dv)
    assert cz.lib().cpsz_decompress_2d_log_f32_dev(0, None, s.data_ptr(), n.value, 2400, 3600, ou.data_ptr(),
                                                   ov.data_ptr()) == 0
    torch.cuda.synchronize()
    host = s.cpu().numpy()
    ref = orc.compress_2d_log(U, V, 0.1)
    assert np.array_equal(mask_2d_log_padding(host, 2400, 3600), mask_2d_log_padding(ref, 2400, 3600))
    c0 = cz.critical_points_2d_dev(du.data_ptr(), dv.data_ptr(), 2400, 3600)
    c1 = cz.critical_points_2d_dev(ou.data_ptr(), ov.data_ptr(), 2400, 3600)
    print(f"[2d log] 2400x3600: {n.value} B, {len(c0)} critical points", flush=True)
    assert len(c0) > 0 and c0 == c1
