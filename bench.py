"""Benchmark: compress+decompress throughput of the cpSZ codec on one MI355X.

Workload (default, the largest single-GPU configuration): 96 u,v float32 2D
fields with heights and widths log-uniform in [256, 4096] (seeded; 220M
vertices, 1.76 GB), relative eb 0.1, each step one batched compress and one
batched decompress of all fields. The plain method (the drop-in) gives
`value`; the learned-eb method is timed on the same batch and reported beside
it (ratio and GB/s, plain vs learned). Other workloads: ``--workload 3d`` (3D
plain cpSZ on u,v,w 256^3, rel eb 0.05), ``2d`` (2400x3600, rel eb 0.1),
``learned2d`` and ``learned3d`` (16 timesteps of 192^3, batched).

A step = one device-resident compress (field in HBM -> pre-ZSTD stream in HBM)
plus one device-resident decompress (stream in HBM -> field in HBM). value =
uncompressed field bytes * steps / timed seconds (GB/s). Prints one JSON line.
The compression ratio (ZSTD level 3 on the host, as the reference drivers do),
the critical-point error count of the decoded field, the roofline of the
dominant kernel and the CPU reference timing are measured outside the timed
region.
"""
from __future__ import annotations

import argparse
import ctypes as C
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", choices=["batch2d", "3d", "2d", "learned2d", "learned3d"], default="batch2d")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-learned", action="store_true", help="batch2d: skip the learned method")
    p.add_argument("--no-graph", action="store_true", help="batch2d: skip the tree placement / graph probe")
    p.add_argument("--no-cp", action="store_true", help="skip the critical-point check")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a plain 256 MiB device copy (the rocprofv3 FETCH_SIZE / WRITE_SIZE "
                        "calibration: 256 MiB read and 256 MiB written per dispatch)")
    return p.parse_args()


def cpu_baseline(workload, fields, eb):
    """Times the CPU codec on a bounded sample, single-threaded."""
    with _QuietStdout():
        return _cpu_baseline(workload, fields, eb)


def _cpu_baseline(workload, fields, eb):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _libs

    if workload == "2d":
        U, V = fields
        r1, r2 = U.shape
        rows = min(r1, 600)  # a 600-row band of the field keeps this near ~10 s
        u, v = np.ascontiguousarray(U[:rows]), np.ascontiguousarray(V[:rows])
        if _libs.have_reference():
            ref = _libs.Reference()
            t0 = time.perf_counter()
            s = ref.compress_2d(u, v, eb)
            t1 = time.perf_counter()
            ref.decompress_2d(s, rows, r2)
            t2 = time.perf_counter()
            kind = "reference"
        else:
            orc = _libs.load_oracle()
            t0 = time.perf_counter()
            s = orc.compress_2d(u, v, eb)
            t1 = time.perf_counter()
            orc.decompress_2d(s, rows, r2)
            t2 = time.perf_counter()
            kind = "port"
        nbytes = u.nbytes + v.nbytes
        sample = f"{rows}x{r2} band of the 2D field, compress+decompress"
    else:
        U, V, W = fields
        r1, r2, r3 = U.shape
        slab = min(r1, 24)  # a 24x256x256 slab keeps this near ~10-20 s
        u, v, w = (np.ascontiguousarray(X[:slab]) for X in (U, V, W))
        orc = _libs.load_oracle()
        t0 = time.perf_counter()
        s = orc.compress_3d(u, v, w, eb)
        t1 = time.perf_counter()
        if _libs.have_reference():
            _libs.Reference().decompress_3d(s, slab, r2, r3)
            kind = "port"  # compressor is the restatement (the reference 3D compressor needs FTK)
        else:
            orc.decompress_3d(s, slab, r2, r3)
            kind = "port"
        t2 = time.perf_counter()
        nbytes = u.nbytes * 3
        sample = f"{slab}x{r2}x{r3} slab of the 3D field, compress+decompress"
    return {"value": nbytes / (t2 - t0) / 1e9, "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": sample, "compress_s": t1 - t0, "decompress_s": t2 - t1}


def _cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


class _QuietStdout:
    """Sends file descriptor 1 to /dev/null for the duration (the reference CPU
    library prints progress to stdout; the bench's own stdout is one JSON line)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        fd = os.open(os.devnull, os.O_WRONLY)
        os.dup2(fd, 1)
        os.close(fd)
        return self

    def __exit__(self, *exc):
        C.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def _ref_worker(args):
    """One field through the CPU codec in a worker process: (seconds, vertices)."""
    path, r1, r2, eb = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _libs
    uv = np.load(path)
    ref = _libs.Reference() if _libs.have_reference() else _libs.load_oracle()
    with _QuietStdout():
        t0 = time.perf_counter()
        ref.decompress_2d(ref.compress_2d(np.ascontiguousarray(uv[0]), np.ascontiguousarray(uv[1]), eb), r1, r2)
        dt = time.perf_counter() - t0
    return dt, r1 * r2


def cpu_baseline_batch2d(fields_host, shapes, eb):
    """The compiled reference (the oracle where it is absent) on a bounded
    sample of the batch: single-threaded, median of 5 over the first fields up
    to ~12M vertices; then one process per core over fields up to ~48M
    vertices, median of 3 (wall time, field bytes / second)."""
    import multiprocessing as mp
    import statistics
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _libs
    have_ref = _libs.have_reference()
    ref = _libs.Reference() if have_ref else _libs.load_oracle()

    def first(limit):
        pick, acc = [], 0
        for k, (r1, r2) in enumerate(shapes):
            if acc + r1 * r2 > limit and pick:
                break
            pick.append(k)
            acc += r1 * r2
        return pick, acc

    pick1, nv1 = first(12_000_000)
    runs = []
    with _QuietStdout():
        for _ in range(5):
            t0 = time.perf_counter()
            for k in pick1:
                U, V = fields_host[k]
                ref.decompress_2d(ref.compress_2d(U, V, eb), *shapes[k])
            runs.append(time.perf_counter() - t0)
    t1 = statistics.median(runs)
    model, ncpu, avail = _cpu_info()
    procs = max(1, min(16, avail))  # the GPU box's CPU share is 16 cores per GPU
    pickn, nvn = first(48_000_000)
    tmp = tempfile.mkdtemp(prefix="cpsz_bench_")
    jobs = []
    try:
        for k in pickn:
            path = os.path.join(tmp, f"f{k}.npy")
            np.save(path, np.stack(fields_host[k]))
            jobs.append((path, shapes[k][0], shapes[k][1], eb))
        jobs.sort(key=lambda j: -j[1] * j[2])  # largest first balances the pool
        walls = []
        with mp.get_context("spawn").Pool(procs) as pool:
            pool.map(_ref_worker, jobs[:procs])  # start-up and page-in, untimed
            for _ in range(3):
                t0 = time.perf_counter()
                pool.map(_ref_worker, jobs, chunksize=1)
                walls.append(time.perf_counter() - t0)
        tn = statistics.median(walls)
    finally:
        for j in jobs:
            try:
                os.remove(j[0])
            except OSError:
                pass
        os.rmdir(tmp)
    kind = "reference" if have_ref else "port"
    return {"value": 8 * nv1 / t1 / 1e9, "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"{len(pick1)} of the 96 fields ({nv1} vertices), compress+decompress one at a time, "
                      f"single thread, median of 5",
            "cpu_model": model, "nproc": ncpu, "cores_available": avail,
            "multi_process": {"value": 8 * nvn / tn / 1e9, "unit": "GB/s", "processes": procs,
                              "sample": f"{len(pickn)} of the 96 fields ({nvn} vertices), one field per task, "
                                        f"median of 3 wall times"}}


def graph_probe(cz, torch, shapes, fl, eb, sh):
    """Compress-only timings (ms, median of 5, outside the timed region) of
    the two tree placements: host trees run eagerly, GPU trees replay a
    captured HIP graph. On the benchmark batch and on 256 small fields
    (h, w log-uniform in [32, 256]), where launch and host round trips weigh
    more than the trees."""
    from cpsz_amd.synth import batch_2d_shapes, field_2d_device
    small = batch_2d_shapes(256, seed=3, lo=32, hi=256)
    sdev = [field_2d_device(r1, r2, seed=2000 + k) for k, (r1, r2) in enumerate(small)]
    sfl = [(u.data_ptr(), v.data_ptr(), r1, r2) for (u, v), (r1, r2) in zip(sdev, small)]
    torch.cuda.synchronize()
    out = {}
    for name, shp, flist in (("batch_96_fields", shapes, fl), ("small_256_fields", small, sfl)):
        w, g, dt = cz.batch_2d_config(shp)
        res = {}
        for trees in ("host", "device"):
            cz.batch_2d_set_config(shp, w, g, trees == "device")
            ts = []
            for r in range(7):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                cz.compress_2d_batch_dev(flist, eb, stream=sh)
                torch.cuda.synchronize()
                ts.append((time.perf_counter() - t0) * 1e3)
            stages = [k for k, _ in cz.last_stages(0)]
            res[trees] = {"ms": float(np.median(ts[2:])), "replayed_graph": stages == ["graph"]}
        cz.batch_2d_set_config(shp, w, g, dt)
        out[name] = res
    return out


def run_batch2d(a, torch):
    """96 fields, h and w log-uniform in [256, 4096] (seeded), rel eb 0.1. A
    step is one batched compress + one batched decompress of all fields. The
    plain method (the drop-in) gives `value`; the learned method is timed on
    the same batch for the plain-vs-learned ratio and throughput."""
    import cpsz_amd as cz
    from cpsz_amd.synth import batch_2d_shapes, field_2d_device

    shapes = batch_2d_shapes(96, seed=0)
    eb = 0.1
    dev = [field_2d_device(r1, r2, seed=1000 + k) for k, (r1, r2) in enumerate(shapes)]
    outs = [(torch.empty_like(u), torch.empty_like(v)) for u, v in dev]
    nv = sum(r1 * r2 for r1, r2 in shapes)
    field_bytes = 8 * nv
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    hip = C.CDLL("libamdhip64.so")
    cz.set_timing(True, 0)
    fl = [(u.data_ptr(), v.data_ptr(), r1, r2) for (u, v), (r1, r2) in zip(dev, shapes)]
    tuned = {"done": False}

    def decode(base, off, size):
        cz.decompress_2d_batch_dev([(base + int(o), int(n)) for o, n in zip(off, size)], shapes,
                                   [(x.data_ptr(), y.data_ptr()) for x, y in outs], stream=sh)

    def step():
        base, off, size = cz.compress_2d_batch_dev(fl, eb, tune=not tuned["done"], stream=sh)
        tuned["done"] = True
        st_c = cz.last_stages(0)
        decode(base, off, size)
        st_d = cz.last_stages(0)
        return base, off, size, st_c, st_d

    per_step = []  # ms of each timed step (events between steps)

    def timed(fn, steps, warmup):
        for _ in range(max(warmup, 1)):
            r = fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        acc = []
        # as timeit does: no garbage-collector pass inside the timed region (a
        # full pass stalls the host thread that drives the device for tens of ms)
        gc.collect()
        gc.disable()
        ev[0].record(stream)
        for k in range(steps):
            r = fn()
            acc.append(r)
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        gc.enable()
        per_step.extend(ev[k].elapsed_time(ev[k + 1]) for k in range(steps))
        return acc, ev[0].elapsed_time(ev[steps]) / 1e3 / steps

    def zbytes(base, off, size):
        # one ZSTD level-3 frame per field (the reference's sz_lossless_compress
        # bytes), the streams copied in chunks while host threads compress
        frames = cz.sz_lossless_compress_batch_dev(base, off, size, 3)
        return sum(f.size for f in frames), int(np.sum(size))

    def cp_errors():
        fp = fn = ft = tot = 0
        for (u, v), (x, y), (r1, r2) in zip(dev, outs, shapes):
            c0 = cz.critical_points_2d_dev(u.data_ptr(), v.data_ptr(), r1, r2)
            c1 = cz.critical_points_2d_dev(x.data_ptr(), y.data_ptr(), r1, r2)
            tot += len(c0)
            fp += sum(1 for k in c1 if k not in c0)
            fn += sum(1 for k in c0 if k not in c1)
            ft += sum(1 for k in c0 if k in c1 and c0[k] != c1[k])
        return {"cp_total": tot, "fp": fp, "fn": fn, "ft": ft}

    # ---- plain (timed region: a.steps batched compress + decompress)
    t_wall = time.perf_counter()
    acc, t_step = timed(step, a.steps, a.warmup)
    t_wall = time.perf_counter() - t_wall
    plain_steps = [round(x, 2) for x in per_step]
    slow = int(np.argmax(per_step))  # the slowest step's stages, to see where a stall sits
    slowest = {"step": slow, "compress": {k: round(v, 2) for k, v in acc[slow][3]},
               "decompress": {k: round(v, 2) for k, v in acc[slow][4]}}
    stages_c, stages_d = {}, {}
    for _, _, _, st_c, st_d in acc:
        for k, v in st_c:
            stages_c[k] = stages_c.get(k, 0.0) + v / a.steps
        for k, v in st_d:
            stages_d[k] = stages_d.get(k, 0.0) + v / a.steps
    base, off, size = acc[-1][:3]
    zb_plain, sb_plain = zbytes(base, off, size)
    cp_plain = None if a.no_cp else cp_errors()
    graph = None if a.no_graph else graph_probe(cz, torch, shapes, fl, eb, sh)
    # ---- learned: trained on 64 seeded 1024x1024 fields (bisection labels), then
    # learned batched compress (proposals, verification rounds, fallback) +
    # batched decompress of the same 96 fields
    learned = None
    if not a.no_learned:
        t0 = time.perf_counter()
        train = [field_2d_device(1024, 1024, seed=5000 + k) for k in range(64)]
        params, losses = cz.learn_train_2d_dev([(u.data_ptr(), v.data_ptr(), 1024, 1024) for u, v in train], eb,
                                               cz.learn_init_params(1), samples_per_field=16384, epochs=8,
                                               batch=4096, lr=2e-3, seed=11)
        train_s = time.perf_counter() - t0
        del train

        def lstep():
            base, off, size, it, boosted, fb = cz.compress_2d_learned_batch_dev(fl, eb, params, stream=sh)
            decode(base, off, size)
            return base, off, size, it, boosted, fb

        lacc, t_learned = timed(lstep, max(1, min(a.steps, 2)), 1)
        lb, lo, ls, it, boosted, fb = lacc[-1]
        zb_l, sb_l = zbytes(lb, lo, ls)
        learned = {"gbps": field_bytes / t_learned / 1e9, "ms_per_step": t_learned * 1e3,
                   "steps": len(lacc), "ratio": field_bytes / zb_l, "stream_bytes": sb_l, "zstd_bytes": zb_l,
                   "cp_errors": None if a.no_cp else cp_errors(), "verification_rounds": it,
                   "boosted_vertices": boosted, "fallback_vertices": fb, "train_seconds": train_s,
                   "train_loss_per_epoch": [float(x) for x in losses]}
    # ---- roofline of the dominant kernel. Its duration comes from the
    # library's HIP-event marks on its launching stream (the stage named here
    # is that one launch). Algorithmic bytes per vertex: the wavefronts read
    # and write 16 B (compress: 8 B field + 4 B eb codes in, 4 B codes out;
    # decompress: 4 B ids + 4 B codes in, 8 B field out).
    cand = [("k_lorenzo2d_batch<true> (batched 2D compress wavefront)", stages_c.get("lorenzo", float("nan")), 16),
            ("k_lorenzo2d_batch<false> (batched 2D decompress wavefront)", stages_d.get("lorenzo", float("nan")),
             16)]
    name, k_ms, per_v = max(cand, key=lambda x: x[1] if x[1] == x[1] else -1)
    achieved = per_v * nv / (k_ms / 1e3) if k_ms == k_ms and k_ms > 0 else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_batch2d.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        kern = name.split(" (")[0]
        if kern in tj.get("bytes_per_dispatch", {}):
            traffic = tj["bytes_per_dispatch"][kern]
    cpu = None
    if not a.no_cpu:
        host = [(u.cpu().numpy(), v.cpu().numpy()) for u, v in dev[:40]]
        cpu = cpu_baseline_batch2d(host, shapes[:40], eb)
    w, g, dt = cz.batch_2d_config(shapes)
    out = {
        "metric": "compress+decompress GB/s and ratio at 0 CP errors: plain vs learned-eb, 1 GPU",
        "value": field_bytes / t_step / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (generated on the GPU, seeded)",
        "config": {"workload": "batched 2D plain cpSZ: 96 u,v float32 fields, h,w log-uniform in [256, 4096], "
                               "rel eb 0.1 (220,143,645 vertices, 1.76 GB)", "method": "plain", "fields": 96,
                   "vertices": nv, "field_bytes": field_bytes, "stream_bytes": sb_plain, "zstd_bytes": zb_plain,
                   "ratio": field_bytes / zb_plain, "cp_errors": cp_plain, "wall_s": t_wall,
                   "ms_each_step": plain_steps, "slowest_step_stages_ms": slowest,
                   "autotuned": {"workers": w, "grid_cap": g, "huffman_trees": "device" if dt else "host"},
                   "compress_ms_by_tree_placement": graph,
                   "stages_compress_ms": stages_c, "stages_decompress_ms": stages_d,
                   "learned": learned,
                   "learned_vs_plain": None if learned is None else {
                       "ratio_learned": learned["ratio"], "ratio_plain": field_bytes / zb_plain,
                       "gbps_learned": learned["gbps"], "gbps_plain": field_bytes / t_step / 1e9}},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9 if achieved else None, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK if achieved else None, "traffic": traffic,
                     "kernel": name, "kernel_ms": k_ms, "alg_bytes_per_launch": per_v * nv},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_learned2d(a, torch):
    """Learned eb proposer on the 2400x3600 field: trained on the GPU on 64 seeded
    1024x1024 fields (bisection labels), then learned compress (with its exact
    verification rounds) + decompress per step; ratio learned vs plain."""
    import cpsz_amd as cz
    from cpsz_amd.synth import field_2d, field_2d_device

    eb = 0.1
    t0 = time.perf_counter()
    train = [field_2d_device(1024, 1024, seed=5000 + k) for k in range(64)]
    params, losses = cz.learn_train_2d_dev([(u.data_ptr(), v.data_ptr(), 1024, 1024) for u, v in train], eb,
                                           cz.learn_init_params(1), samples_per_field=16384, epochs=8,
                                           batch=4096, lr=2e-3, seed=11)
    train_s = time.perf_counter() - t0
    del train
    U, V = field_2d(2400, 3600, seed=0)
    du, dv = torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda()
    ou, ov = torch.empty_like(du), torch.empty_like(dv)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    sbuf = torch.empty(64 << 20, dtype=torch.uint8, device="cuda")
    hip = C.CDLL("libamdhip64.so")

    def step():
        p, n, it, boosted, fb = cz.compress_2d_learned_dev(du.data_ptr(), dv.data_ptr(), 2400, 3600, eb, params,
                                                           stream=sh)
        hip.hipMemcpyAsync(C.c_void_p(sbuf.data_ptr()), C.c_void_p(p), C.c_size_t(n), 3, C.c_void_p(sh))
        cz.decompress_2d_dev(sbuf.data_ptr(), n, 2400, 3600, ou.data_ptr(), ov.data_ptr(), 0, sh)
        return n, it, boosted, fb

    for _ in range(max(a.warmup, 1)):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(a.steps):
        n, it, boosted, fb = step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t_step = ev0.elapsed_time(ev1) / 1e3 / a.steps
    field_bytes = U.nbytes + V.nbytes
    learned = np.empty(n, np.uint8)
    hip.hipMemcpy(C.c_void_p(learned.ctypes.data), C.c_void_p(sbuf.data_ptr()), C.c_size_t(n), 2)
    plain = cz.sz_compress_cp_preserve_2d_offline(U, V, eb)
    zl, zp = cz.sz_lossless_compress(learned, 3).size, cz.sz_lossless_compress(plain, 3).size
    c0 = cz.critical_points_2d_dev(du.data_ptr(), dv.data_ptr(), 2400, 3600)
    c1 = cz.critical_points_2d_dev(ou.data_ptr(), ov.data_ptr(), 2400, 3600)
    cp_err = {"cp_total": len(c0), "fp": sum(1 for k in c1 if k not in c0), "fn": sum(1 for k in c0 if k not in c1),
              "ft": sum(1 for k in c0 if k in c1 and c0[k] != c1[k])}
    out = {
        "metric": "compress+decompress GB/s and ratio at 0 CP errors: plain vs learned-eb, 1 GPU",
        "value": field_bytes / t_step / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 codec, bf16 MLP (fp32 accumulate)", "data": "synthetic",
        "config": {"workload": "learned eb 2D: u,v float32 2400x3600, rel eb 0.1; MLP 51-128-128-1 trained on "
                               "64 seeded 1024x1024 fields", "method": "learned",
                   "ratio_learned": field_bytes / zl, "ratio_plain": field_bytes / zp, "cp_errors": cp_err,
                   "verification_rounds": it, "boosted_vertices": boosted, "fallback_vertices": fb,
                   "train_seconds": train_s, "train_loss_per_epoch": [float(x) for x in losses]},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def run_learned3d(a, torch):
    """Learned proposer plus batching, 3D: 16 timesteps of 192^3 u,v,w (the 3D
    recipe with the vortex-tube centres advected 2 cells per step), rel eb
    0.05. The 82-feature MLP is trained on the GPU on 8 other seeded 128^3
    fields. Per step: one learned batched compress (proposals, batched
    compress, batched decode and exact critical-point verification rounds,
    fallback) + one batched decompress. The plain batched codec is timed on the
    same batch for the ratio and throughput comparison."""
    import cpsz_amd as cz
    from cpsz_amd.synth import field_3d_device

    eb, nt, n3 = 0.05, 16, 192
    t0 = time.perf_counter()
    train = [field_3d_device(128, 128, 128, seed=7000 + k) for k in range(8)]
    params, losses = cz.learn_train_3d_dev([(u.data_ptr(), v.data_ptr(), w.data_ptr(), 128, 128, 128)
                                            for u, v, w in train], eb, cz.learn_init_params_3d(1),
                                           samples_per_field=16384, epochs=8, batch=4096, lr=2e-3, seed=11)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del train
    shape = (n3, n3, n3)
    dev = [field_3d_device(*shape, seed=0, step=t) for t in range(nt)]
    outs = [tuple(torch.empty_like(x) for x in f) for f in dev]
    fl = [(u.data_ptr(), v.data_ptr(), w.data_ptr(), *shape) for u, v, w in dev]
    nv = nt * n3 ** 3
    field_bytes = 12 * nv
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    steps = max(1, min(a.steps, 3))
    hip = C.CDLL("libamdhip64.so")
    cz.set_timing(True, 0)

    def timed(fn):
        for _ in range(max(1, min(a.warmup, 1))):
            r = fn()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            r = fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return r, ev0.elapsed_time(ev1) / 1e3 / steps

    def decode(base, off, size):
        cz.decompress_3d_batch_dev([(base + int(o), int(n)) for o, n in zip(off, size)], [shape] * nt,
                                   [tuple(x.data_ptr() for x in o) for o in outs], stream=sh)

    def plain_step():
        base, off, size = cz.compress_3d_batch_dev(fl, eb, tune=True, stream=sh)
        stc = cz.last_stages(0)
        decode(base, off, size)
        return base, off, size, stc, cz.last_stages(0)

    def learned_step():
        base, off, size, it, boosted, fb = cz.compress_3d_learned_batch_dev(fl, eb, params, stream=sh)
        decode(base, off, size)
        return base, off, size, it, boosted, fb

    def zbytes(base, off, size):
        total = int(off[-1] + size[-1])
        allb = np.empty(total, np.uint8)
        hip.hipMemcpy(C.c_void_p(allb.ctypes.data), C.c_void_p(base), C.c_size_t(total), 2)
        return sum(cz.sz_lossless_compress(allb[o:o + n], 3).size for o, n in zip(off, size))

    (pb, po, ps, stc, std), t_plain = timed(plain_step)
    z_plain = zbytes(pb, po, ps)
    (lb, lo, ls, it, boosted, fb), t_learned = timed(learned_step)
    z_learned = zbytes(lb, lo, ls)
    cp_err = None
    if not a.no_cp:  # the learned streams, decoded by the last step
        fp = fn = ft = tot = 0
        for f, o in zip(dev, outs):
            c0 = cz.critical_points_3d_dev(*(x.data_ptr() for x in f), *shape)
            c1 = cz.critical_points_3d_dev(*(x.data_ptr() for x in o), *shape)
            tot += len(c0)
            fp += sum(1 for k in c1 if k not in c0)
            fn += sum(1 for k in c0 if k not in c1)
            ft += sum(1 for k in c0 if k in c1 and c0[k] != c1[k])
        cp_err = {"cp_total": tot, "fp": fp, "fn": fn, "ft": ft}
    w, g = cz.batch_3d_config([shape] * nt)
    stages_c = {}
    for k, v in stc:
        stages_c[k] = stages_c.get(k, 0.0) + v
    stages_d = {}
    for k, v in std:
        stages_d[k] = stages_d.get(k, 0.0) + v
    k_ms = stages_c.get("lorenzo", float("nan"))
    achieved = 20 * nv / (k_ms / 1e3) if k_ms == k_ms and k_ms > 0 else None
    out = {
        "metric": "compress+decompress GB/s and ratio at 0 CP errors: plain vs learned-eb, 1 GPU",
        "value": field_bytes / t_learned / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": t_learned * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 codec, bf16 MLP (fp32 accumulate)", "data": "synthetic (generated on the GPU, seeded)",
        "config": {"workload": "learned eb + batching 3D: 16 timesteps of u,v,w float32 192^3 (vortex tubes "
                               "advected 2 cells/step), rel eb 0.05; MLP 82-128-128-1 trained on 8 seeded 128^3 "
                               "fields", "method": "learned", "fields": nt, "vertices": nv,
                   "field_bytes": field_bytes, "ratio_learned": field_bytes / z_learned,
                   "ratio_plain": field_bytes / z_plain, "plain_gbps": field_bytes / t_plain / 1e9,
                   "plain_ms_per_step": t_plain * 1e3, "cp_errors": cp_err, "verification_rounds": it,
                   "boosted_vertices": boosted, "fallback_vertices": fb, "train_seconds": train_s,
                   "train_loss_per_epoch": [float(x) for x in losses], "autotuned": {"workers": w, "grid_cap": g},
                   "plain_stages_compress_ms": stages_c, "plain_stages_decompress_ms": stages_d},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9 if achieved else None, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK if achieved else None, "traffic": None,
                     "kernel": "k_wave3d<true> (batched 3D compress wavefront, 16 fields in one launch)",
                     "kernel_ms": k_ms, "alg_bytes_per_launch": 20 * nv},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the codec benchmark runs on one GPU: --gpus must be 1"}))
        sys.exit(2)
    import torch

    if a.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4,
                          "dispatches": 5}))
        return
    if a.workload == "batch2d":
        torch.cuda.set_device(0)
        run_batch2d(a, torch)
        return
    if a.workload == "learned2d":
        torch.cuda.set_device(0)
        run_learned2d(a, torch)
        return
    if a.workload == "learned3d":
        torch.cuda.set_device(0)
        run_learned3d(a, torch)
        return
    import cpsz_amd as cz
    from cpsz_amd.synth import field_2d, field_3d

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    if a.workload == "2d":
        shape, eb = (2400, 3600), 0.1
        fields = field_2d(*shape, seed=0)
        workload = "2D plain cpSZ, u,v float32 2400x3600, rel eb 0.1"
    else:
        shape, eb = (256, 256, 256), 0.05
        fields = field_3d(*shape, seed=0)
        workload = "3D plain cpSZ, u,v,w float32 256^3, rel eb 0.05"
    dfs = [torch.from_numpy(f).to(dev) for f in fields]
    outs = [torch.empty_like(d) for d in dfs]
    field_bytes = sum(f.nbytes for f in fields)
    cz.set_timing(True, 0)

    def compress():
        if a.workload == "2d":
            return cz.compress_2d_dev(dfs[0].data_ptr(), dfs[1].data_ptr(), *shape, eb, 0, sh)
        return cz.compress_3d_dev(dfs[0].data_ptr(), dfs[1].data_ptr(), dfs[2].data_ptr(), *shape, eb,
                                  0, sh)

    # the stream is copied out of the library-owned buffer so compress and
    # decompress of a step are independent calls
    sbuf = torch.empty(64 << 20, dtype=torch.uint8, device=dev)
    hip = C.CDLL("libamdhip64.so")

    def step():
        p, n = compress()
        st_c = cz.last_stages(0)
        if n > sbuf.numel():
            raise RuntimeError("stream larger than the staging buffer")
        hip.hipMemcpyAsync(C.c_void_p(sbuf.data_ptr()), C.c_void_p(p), C.c_size_t(n), 3, C.c_void_p(sh))
        if a.workload == "2d":
            cz.decompress_2d_dev(sbuf.data_ptr(), n, *shape, outs[0].data_ptr(), outs[1].data_ptr(), 0, sh)
        else:
            cz.decompress_3d_dev(sbuf.data_ptr(), n, *shape, outs[0].data_ptr(), outs[1].data_ptr(),
                                 outs[2].data_ptr(), 0, sh)
        st_d = cz.last_stages(0)
        return n, st_c, st_d

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    stages_c, stages_d = {}, {}
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(a.steps):
        n, st_c, st_d = step()
        for k, v in st_c:
            stages_c[k] = stages_c.get(k, 0.0) + v
        for k, v in st_d:
            stages_d[k] = stages_d.get(k, 0.0) + v
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)
    t_step = ms / 1e3 / a.steps
    value = field_bytes / t_step / 1e9
    stages_c = {k: v / a.steps for k, v in stages_c.items()}
    stages_d = {k: v / a.steps for k, v in stages_d.items()}
    # ---- outside the timed region: ratio, CP errors, roofline
    p, n = compress()
    stream_host = np.empty(n, np.uint8)
    torch.cuda.synchronize()
    C.CDLL("libamdhip64.so").hipMemcpy(C.c_void_p(stream_host.ctypes.data), C.c_void_p(p), C.c_size_t(n), 2)
    z = cz.sz_lossless_compress(stream_host, 3)
    ratio = field_bytes / z.size
    cp_err = None
    if not a.no_cp:
        try:
            if a.workload == "2d":
                c0 = cz.critical_points_2d_dev(dfs[0].data_ptr(), dfs[1].data_ptr(), *shape)
                c1 = cz.critical_points_2d_dev(outs[0].data_ptr(), outs[1].data_ptr(), *shape)
            else:
                c0 = cz.critical_points_3d_dev(*[d.data_ptr() for d in dfs], *shape)
                c1 = cz.critical_points_3d_dev(*[o.data_ptr() for o in outs], *shape)
            fp = sum(1 for k in c1 if k not in c0)
            fn = sum(1 for k in c0 if k not in c1)
            ft = sum(1 for k in c0 if k in c1 and c0[k] != c1[k])
            cp_err = {"cp_total": len(c0), "fp": fp, "fn": fn, "ft": ft}
        except Exception as e:  # report a failed CP check instead of hiding the bench line
            cp_err = {"error": str(e)}
    # roofline of the dominant kernel (largest share of the step's GPU time).
    # Algorithmic bytes per vertex (each input read once, each output written once):
    #   3D bounds     k_eb3d<true> + k_eb3d_fix: 12 B field in, 1 B eb code out                = 13
    #   3D compress  k_wave3d<true>:   12 B log values + 1 B eb code in, 6 B codes + 1 B id out = 20
    #   3D decompress k_wave3d<false>: 1 B id + 6 B codes + 12 B initial values in, 12 B out     = 31
    #   2D bounds     k_eb2d:            8 B field in, 4 B eb codes out                          = 12
    #   2D compress  k_lorenzo2d<true>:  8 B field + 4 B eb codes in, 4 B codes out             = 16
    #   2D decompress k_lorenzo2d<false>: 4 B ids + 4 B codes in, 8 B field out                  = 16
    # Kernel times are the library's per-stage HIP-event marks on its stream
    # (each stage named here is that kernel and nothing else of weight).
    nv = int(np.prod(shape))
    if a.workload == "2d":
        cand = [("k_eb2d (2D bounds)", stages_c.get("eb", float("nan")), 12),
                ("k_lorenzo2d<true> (2D compress wavefront)", stages_c.get("lorenzo", float("nan")), 16),
                ("k_lorenzo2d<false> (2D decompress wavefront)", stages_d.get("lorenzo", float("nan")), 16)]
    else:
        cand = [("k_eb3d<true> + k_eb3d_fix (3D bounds)", stages_c.get("eb", float("nan")), 13),
                ("k_wave3d<true> (3D compress wavefront)", stages_c.get("lorenzo", float("nan")), 20),
                ("k_wave3d<false> (3D decompress wavefront)", stages_d.get("lorenzo", float("nan")), 31)]
    name, k_ms, per_v = max(cand, key=lambda x: x[1] if x[1] == x[1] else -1)
    alg = nv * per_v
    achieved = alg / (k_ms / 1e3) if k_ms == k_ms and k_ms > 0 else None
    kernels = [k.strip() for k in name.split(" (")[0].split("+")]
    # HBM bytes per dispatch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE on this
    # workload, scaled by the copy-kernel calibration (profiles/traffic_3d.json);
    # VALU instructions per dispatch from rocprofv3 SQ_INSTS_VALU (profiles/sq_3d.json)
    traffic, valu = None, None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{a.workload}.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        if tj.get("workload") == workload and all(k in tj["bytes_per_dispatch"] for k in kernels):
            traffic = sum(tj["bytes_per_dispatch"][k] for k in kernels)
    sfile = os.path.join(ROOT, "profiles", f"sq_{a.workload}.json")
    if os.path.exists(sfile) and achieved:
        with open(sfile) as fh:
            sj = json.load(fh)
        per = sj["valu_instructions_per_dispatch"]
        if sj.get("workload") == workload and all(k in per for k in kernels):
            instr = sum(per[k] for k in kernels)
            rate = instr / (k_ms / 1e3)
            valu = {"instructions_per_launch": instr, "achieved": rate / 1e9, "peak": sj["valu_peak_per_s"] / 1e9,
                    "unit": "G wave64 VALU instr/s", "frac": rate / sj["valu_peak_per_s"]}
    roof = {"bound": "hbm", "achieved": achieved / 1e9 if achieved else None, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK if achieved else None, "traffic": traffic,
            "kernel": name, "kernel_ms": k_ms, "alg_bytes_per_launch": alg, "valu_issue": valu,
            "other_kernels": {c[0]: {"ms": c[1], "GB/s": nv * c[2] / (c[1] / 1e3) / 1e9}
                              for c in cand if c[0] != name and c[1] == c[1] and c[1] > 0}}
    cpu = None if a.no_cpu else cpu_baseline(a.workload, fields, eb)
    out = {
        "metric": "compress+decompress GB/s and ratio at 0 CP errors: plain vs learned-eb, 1 GPU",
        "value": value, "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "method": "plain", "field_bytes": field_bytes,
                   "stream_bytes": n, "zstd_bytes": int(z.size), "ratio": ratio,
                   "cp_errors": cp_err, "stages_compress_ms": stages_c,
                   "stages_decompress_ms": stages_d, "wall_s": wall},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
