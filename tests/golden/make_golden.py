"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run from the repo root after ``make -C oracle`` has built oracle/_ref/libcpsz_ref.so
from /root/reference. Each fixture is an .npz of inputs and the reference's
outputs (no code). The uninitialised meanInfo padding bytes of 2D streams are
zeroed. 3D streams come from the oracle (the reference 3D compressor does not
build without FTK); the decoded fields come from the reference 3D decoder.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from _libs import Reference, load_oracle, mask_2d_log_padding, mask_2d_padding  # noqa: E402
from cpsz_amd.synth import field_2d, field_3d  # noqa: E402


def offline_log():
    """2D offline_log streams and decoded fields from the compiled reference
    (src/sz_compress_cp_preserve_2d.cpp:285-374, src/sz_decompress_cp_preserve_2d.cpp:54-99),
    including exact zeros (log value -100, sign 0)."""
    ref = Reference()
    for r1, r2, seed, eb, zeros in [(40, 50, 1, 0.1, False), (77, 63, 2, 0.05, True), (19, 130, 3, 0.2, True)]:
        U, V = field_2d(r1, r2, seed)
        if zeros:
            U[::7, ::5] = 0
            V[3::11, :] = 0
        s = mask_2d_log_padding(ref.compress_2d_log(U, V, eb), r1, r2)
        dU, dV = ref.decompress_2d_log(s, r1, r2)
        np.savez_compressed(os.path.join(HERE, f"c2l_{r1}x{r2}_s{seed}.npz"), U=U, V=V, eb=np.float64(eb),
                            stream=s, dU=dU, dV=dV)


def main():
    offline_log()
    ref = Reference()
    orc = load_oracle()
    out = {}
    # 2D offline streams and decoded fields.
    cases = [(40, 50, 1, 0.1), (64, 97, 2, 0.1), (123, 77, 3, 0.1), (31, 200, 4, 0.1),
             (57, 45, 5, 0.01), (25, 26, 6, 0.5)]
    for r1, r2, seed, eb in cases:
        U, V = field_2d(r1, r2, seed)
        s = mask_2d_padding(ref.compress_2d(U, V, eb), r1, r2)
        dU, dV = ref.decompress_2d(s, r1, r2)
        np.savez_compressed(os.path.join(HERE, f"c2d_{r1}x{r2}_s{seed}.npz"), U=U, V=V,
                            eb=np.float64(eb), stream=s, dU=dU, dV=dV)
    # f64 instantiation
    U, V = field_2d(33, 41, 9)
    U, V = U.astype(np.float64), V.astype(np.float64)
    s = mask_2d_padding(ref.compress_2d(U, V, 0.1), 33, 41, 8)
    dU, dV = ref.decompress_2d(s, 33, 41, np.float64)
    np.savez_compressed(os.path.join(HERE, "c2d_f64_33x41_s9.npz"), U=U, V=V, eb=np.float64(0.1),
                        stream=s, dU=dU, dV=dV)
    # Edge cases: a constant field (zero vectors -> all unpredictable, single
    # symbol alphabets) and a field whose bounds land in (eps, 4 eps).
    U = np.zeros((24, 30), np.float32)
    V = np.zeros((24, 30), np.float32)
    s = mask_2d_padding(ref.compress_2d(U, V, 0.1), 24, 30)
    dU, dV = ref.decompress_2d(s, 24, 30)
    np.savez_compressed(os.path.join(HERE, "c2d_zero_24x30.npz"), U=U, V=V, eb=np.float64(0.1),
                        stream=s, dU=dU, dV=dV)
    U, V = field_2d(30, 30, 11)
    U = (U * np.float32(1e-14)).astype(np.float32)
    V = (V * np.float32(3e-15)).astype(np.float32)
    s = mask_2d_padding(ref.compress_2d(U, V, 0.1), 30, 30)
    dU, dV = ref.decompress_2d(s, 30, 30)
    np.savez_compressed(os.path.join(HERE, "c2d_tiny_30x30.npz"), U=U, V=V, eb=np.float64(0.1),
                        stream=s, dU=dU, dV=dV)
    # Huffman sections straight from the reference coder.
    rng = np.random.default_rng(12345)
    syms_list, secs = [], []
    for state, n, k in [(2048, 5000, 40), (2048, 1, 1), (2048, 7, 1), (131072, 20000, 3000),
                        (131072, 3000, 2), (512, 999, 25), (131072, 12000, 12000)]:
        if k >= n:
            syms = rng.permutation(n).astype(np.int32) % (2 * state)
        else:
            syms = (rng.geometric(0.05, size=n) % k).astype(np.int32) + (state // 2 - k // 2)
        syms_list.append(syms)
        secs.append(ref.huff_encode(state, syms))
    np.savez_compressed(os.path.join(HERE, "huffman.npz"),
                        **{f"syms{i}": s for i, s in enumerate(syms_list)},
                        **{f"sec{i}": s for i, s in enumerate(secs)},
                        states=np.array([2048, 2048, 2048, 131072, 131072, 512, 131072]))
    # 2D Lorenzo section with a given precision field.
    U, V = field_2d(45, 38, 21)
    eb = orc.eb_2d(U, V, 0.1)
    ids, prec = orc.eb_quantize_2d(np.abs(U).astype(np.float64) * eb)
    sec = ref.lorenzo_2d(U, prec)
    sec = sec.copy()
    sec[10:13] = 0
    np.savez_compressed(os.path.join(HERE, "lorenzo2d_45x38.npz"), data=U, prec=prec, sec=sec)
    # 3D: oracle streams decoded by the reference decoder.
    for r1, r2, r3, seed in [(12, 13, 14, 1), (14, 16, 18, 2), (9, 17, 33, 3)]:
        U, V, W = field_3d(r1, r2, r3, seed)
        s = orc.compress_3d(U, V, W, 0.05)
        dU, dV, dW = ref.decompress_3d(s, r1, r2, r3)
        np.savez_compressed(os.path.join(HERE, f"c3d_{r1}x{r2}x{r3}_s{seed}.npz"), U=U, V=V, W=W,
                            eb=np.float64(0.05), stream=s, dU=dU, dV=dV, dW=dW)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "log":
        offline_log()
    else:
        main()
