"""Batched 2D codec on the GPU: every field's stream equals the single-field
codec's (and the oracle's), decoded fields equal the single-field decoder's,
for uneven shapes in one persistent wavefront launch, with and without the
autotuner."""
import ctypes as C

import numpy as np
import pytest
import torch

import cpsz_amd as cz
from _libs import load_oracle, mask_2d_padding
from cpsz_amd.synth import field_2d

pytestmark = pytest.mark.gpu

SHAPES = [(1, 5), (70, 1), (64, 64), (65, 129), (130, 70), (200, 333), (257, 100), (33, 512), (400, 401),
          (129, 2)]


def _host(ptr, n):
    out = np.empty(n, np.uint8)
    torch.cuda.synchronize()
    C.CDLL("libamdhip64.so").hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(n), 2)
    return out


@pytest.fixture(scope="module")
def batch():
    fields = [field_2d(r1, r2, 30 + k) for k, (r1, r2) in enumerate(SHAPES)]
    dev = [(torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda()) for U, V in fields]
    return fields, dev


@pytest.mark.parametrize("tune", [False, True])
def test_batch_streams_match_single_field(batch, tune):
    print(f"[batch] {len(SHAPES)} fields tune={tune}", flush=True)
    fields, dev = batch
    base, off, size = cz.compress_2d_batch_dev([(u.data_ptr(), v.data_ptr(), *SHAPES[k]) for k, (u, v)
                                                in enumerate(dev)], 0.1, tune=tune)
    orc = load_oracle()
    allb = _host(base, int(off[-1] + size[-1]))
    for k, ((U, V), shape) in enumerate(zip(fields, SHAPES)):
        s = allb[off[k]:off[k] + size[k]]
        single = cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
        assert np.array_equal(s, single), shape
        if k % 3 == 0:
            ref = orc.compress_2d(U, V, 0.1)
            assert np.array_equal(mask_2d_padding(s, *shape), mask_2d_padding(ref, *shape)), shape
    w, g, _ = cz.batch_2d_config(SHAPES)
    assert w >= 1 and g >= 1


@pytest.mark.parametrize("device_trees", [True, False])
def test_batch_tree_placement_gives_identical_streams(batch, device_trees):
    """Huffman trees built on the GPU (one wave per tree) or on the host pool:
    the same bytes, equal to the single-field codec's."""
    print(f"[batch] device_trees={device_trees}", flush=True)
    fields, dev = batch
    cz.batch_2d_set_config(SHAPES, 4, 1024, device_trees)
    assert cz.batch_2d_config(SHAPES) == (4, 1024, device_trees)
    base, off, size = cz.compress_2d_batch_dev([(u.data_ptr(), v.data_ptr(), *SHAPES[k]) for k, (u, v)
                                                in enumerate(dev)], 0.1)
    allb = _host(base, int(off[-1] + size[-1]))
    for k, ((U, V), shape) in enumerate(zip(fields, SHAPES)):
        single = cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
        assert np.array_equal(allb[off[k]:off[k] + size[k]], single), shape


def _stage_names():
    return [k for k, _ in cz.last_stages(0)]


def test_batch_graph_replays_match_eager(batch):
    """Trees on the GPU: the first call runs eagerly and captures the batch;
    later calls with the same inputs replay the graph. Every replay (with
    other eager work on the context in between, and after the inputs change
    in place) gives the single-field codec's streams for the current data."""
    fields, dev = batch
    own = [(u.clone(), v.clone()) for u, v in dev]
    torch.cuda.synchronize()
    fl = [(u.data_ptr(), v.data_ptr(), *SHAPES[k]) for k, (u, v) in enumerate(own)]
    cz.batch_2d_set_config(SHAPES, 4, 1024, True)
    cz.set_timing(True, 0)
    host = [(U.copy(), V.copy()) for U, V in fields]
    try:
        for call in range(5):
            print(f"[batch graph] call {call}", flush=True)
            if call == 3:  # new data in the same buffers
                U, V = field_2d(*SHAPES[6], 77)
                own[6][0].copy_(torch.from_numpy(U))
                own[6][1].copy_(torch.from_numpy(V))
                host[6] = (U, V)
                torch.cuda.synchronize()
            base, off, size = cz.compress_2d_batch_dev(fl, 0.1)
            assert ("graph" in _stage_names()) == (call > 0), _stage_names()
            allb = _host(base, int(off[-1] + size[-1]))
            for k, ((U, V), shape) in enumerate(zip(host, SHAPES)):
                single = cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
                assert np.array_equal(allb[off[k]:off[k] + size[k]], single), (call, shape)
            # other eager work on the same context between the replays
            s = cz.sz_compress_cp_preserve_2d_offline(*field_2d(90, 110, call), 0.1)
            cz.sz_decompress_cp_preserve_2d_offline(s, 90, 110)
    finally:
        cz.set_timing(False, 0)
        cz.batch_2d_set_config(SHAPES, 4, 1024, False)


@pytest.mark.parametrize("device_trees", [False, True])
def test_batch_nonfinite_field_reported(batch, device_trees):
    """A field with a NaN makes the batched compress fail with the non-finite
    error (the wavefront skips that field); the next batch is unaffected."""
    fields, dev = batch
    bad = [(u.clone(), v.clone()) for u, v in dev]
    bad[4][0][7, 9] = float("nan")
    torch.cuda.synchronize()
    cz.batch_2d_set_config(SHAPES, 4, 1024, device_trees)
    try:
        with pytest.raises(cz.CpszError) as e:
            cz.compress_2d_batch_dev([(u.data_ptr(), v.data_ptr(), *SHAPES[k]) for k, (u, v) in enumerate(bad)], 0.1)
        assert e.value.code == -6
        base, off, size = cz.compress_2d_batch_dev([(u.data_ptr(), v.data_ptr(), *SHAPES[k]) for k, (u, v)
                                                    in enumerate(dev)], 0.1)
        allb = _host(base, int(off[-1] + size[-1]))
        U, V = fields[4]
        assert np.array_equal(allb[off[4]:off[4] + size[4]], cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1))
    finally:
        cz.batch_2d_set_config(SHAPES, 4, 1024, False)


def test_batch_round_trip_matches_single_decoder(batch):
    fields, dev = batch
    base, off, size = cz.compress_2d_batch_dev([(u.data_ptr(), v.data_ptr(), *SHAPES[k]) for k, (u, v)
                                                in enumerate(dev)], 0.1)
    outs = [(torch.empty(s, dtype=torch.float32, device="cuda"), torch.empty(s, dtype=torch.float32, device="cuda"))
            for s in SHAPES]
    cz.decompress_2d_batch_dev([(base + int(o), int(n)) for o, n in zip(off, size)], SHAPES,
                               [(a.data_ptr(), b.data_ptr()) for a, b in outs])
    torch.cuda.synchronize()
    allb = _host(base, int(off[-1] + size[-1]))
    for k, shape in enumerate(SHAPES):
        dU, dV = cz.sz_decompress_cp_preserve_2d_offline(allb[off[k]:off[k] + size[k]], *shape)
        assert np.array_equal(outs[k][0].cpu().numpy().view(np.uint32), dU.view(np.uint32)), shape
        assert np.array_equal(outs[k][1].cpu().numpy().view(np.uint32), dV.view(np.uint32)), shape


def test_batch_compress_waits_for_inputs_on_caller_stream(batch):
    """Inputs written by a kernel queued on the caller's stream (behind slow
    work) are what the batch compresses: the bounds passes, which run on the
    library's worker streams, wait for the caller's stream first."""
    fields, _ = batch
    print("[batch] inputs written on a side stream", flush=True)
    side = torch.cuda.Stream()
    src = [(torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda()) for U, V in fields]
    dst = [(torch.zeros_like(u), torch.zeros_like(v)) for u, v in src]
    big = torch.randn(4096, 4096, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        for _ in range(8):  # delays the copies below by milliseconds
            big = torch.tanh(big @ big)
        for (u, v), (a, b) in zip(src, dst):
            a.copy_(u)
            b.copy_(v)
    base, off, size = cz.compress_2d_batch_dev([(a.data_ptr(), b.data_ptr(), *SHAPES[k]) for k, (a, b)
                                                in enumerate(dst)], 0.1, stream=side.cuda_stream)
    side.synchronize()
    allb = _host(base, int(off[-1] + size[-1]))
    for k, ((U, V), shape) in enumerate(zip(fields, SHAPES)):
        assert np.array_equal(allb[off[k]:off[k] + size[k]], cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)), shape
