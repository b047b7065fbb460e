"""The double instantiation of the 2D offline pair on the GPU against the
compiled reference's fixture and the oracle's double path: streams byte-equal
(the reference's seven uninitialised meanInfo<double> padding bytes masked),
decoded fields bit-equal."""
import os

import numpy as np
import pytest

import cpsz_amd as cz
from _libs import load_oracle, mask_2d_padding
from cpsz_amd.synth import field_2d

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def test_f64_golden():
    z = np.load(os.path.join(GOLD, "c2d_f64_33x41_s9.npz"))
    U, V = z["U"], z["V"]
    assert U.dtype == np.float64
    s = cz.sz_compress_cp_preserve_2d_offline(U, V, float(z["eb"]))
    assert np.array_equal(mask_2d_padding(s, 33, 41, 8), z["stream"])
    dU, dV = cz.sz_decompress_cp_preserve_2d_offline(z["stream"], 33, 41, np.float64)
    assert np.array_equal(dU.view(np.uint64), z["dU"].view(np.uint64))
    assert np.array_equal(dV.view(np.uint64), z["dV"].view(np.uint64))


@pytest.mark.parametrize("shape,seed,eb", [((1, 70), 1, 0.1), ((70, 1), 2, 0.1), ((2, 2), 3, 0.1),
                                           ((65, 129), 4, 0.05), ((130, 64), 5, 0.3), ((300, 97), 6, 0.01),
                                           ((700, 900), 7, 0.1)])
def test_f64_vs_oracle(orc, shape, seed, eb):
    r1, r2 = shape
    print(f"[2d f64] {shape} seed {seed} eb {eb}", flush=True)
    U, V = field_2d(r1, r2, seed)
    rng = np.random.default_rng(seed)
    # doubles that are not floats: the products and predictions differ from the float path
    U = U.astype(np.float64) * (1 + rng.standard_normal(U.shape) * 1e-9)
    V = V.astype(np.float64) * (1 + rng.standard_normal(V.shape) * 1e-9)
    s = cz.sz_compress_cp_preserve_2d_offline(U, V, eb)
    ref = orc.compress_2d(U, V, eb)
    assert np.array_equal(mask_2d_padding(s, r1, r2, 8), mask_2d_padding(ref, r1, r2, 8))
    dU, dV = cz.sz_decompress_cp_preserve_2d_offline(ref, r1, r2, np.float64)
    oU, oV = orc.decompress_2d(ref, r1, r2, np.float64)
    assert np.array_equal(dU.view(np.uint64), oU.view(np.uint64))
    assert np.array_equal(dV.view(np.uint64), oV.view(np.uint64))
