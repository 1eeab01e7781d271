#!/usr/bin/env python3
"""Fused select + agg_sum over compressed columns on one MI355X.

Workload (format cascade sweep): four uint64 columns of 2^27 values, each
compressed on the GPU and resident in HBM before timing starts:
  (a) dynamic_vbp_f<512,2048,8>, per-512-value block bit width U{1..32}
  (b) delta_f<1024,8,dynamic_vbp_f<512,64,8>> over a sorted column with
      geometric(p=0.1) gaps
  (c) for_f<1024,8,dynamic_vbp_f<512,64,8>>, base U[2^40,2^41) + offsets < 2^16
  (d) rle_f, run lengths geometric(mean 64), run values U[0,2^20)
A step runs SELECT SUM(x) WHERE x < c on each column (c gives about 50%
selectivity).  Two executions of the same step are timed:
  plain    one fused call per column (ms_select_sum: 2 kernels per column)
  batched  the four columns as one batch (ms_batch_run: one persistent
           kernel over a device work queue + a partial-sum fold)
`value` is the batched step: compressed bytes read per step (sum of
sizeUsed + 8 B per result) divided by the step time.

Second workload (uneven columns, --c4-cols, default 4096): columns of
log-uniform length in [10^3, 10^6], format uniform over static_vbp<b,8>,
dynamic_vbp, delta+dynamic_vbp, for+dynamic_vbp and rle, one select+sum per
column; timed per-column (eager, and replayed as one hipGraph) against one
autotuned batch.  Every result of every path is checked bit-exactly
against numpy before it is timed.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "fused select+sum GB/s over compressed cols, %HBM peak; batched vs plain"
CMP_LT = 2
# The sweep's batch shape: fixed, so that the profiles and the traffic
# measurement under profiles/ belong to exactly the timed launch.  The
# autotuner's own pick and its gain over the library default are reported
# beside it (autotune).
SWEEP_CFG = {"wg_per_cu": 4, "units_per_item": 32, "tiles_per_item": 8, "runs_per_item": 4096,
             "values_per_item": 2048, "waves_per_simd": 0, "units_per_item_for": 0,
             "units_per_item_narrow": 0, "unroll": 0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def masked_sum(x, val):
    with np.errstate(over="ignore"):
        return int(np.sum(np.where(x < np.uint64(val), x, np.uint64(0)), dtype=np.uint64))


def make_layouts(n, seed, with_uncompr=False):
    from engine_amd import datagen
    from engine_amd.formats import (default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f, dynamic_vbp_f,
                                    rle_f, uncompr_f)
    x = datagen.dvbp_random_bw(seed, n)
    yield "dynamic_vbp<512,2048,8>", dynamic_vbp_f(512, 2048, 8), x, 1 << 16
    x = datagen.sorted_geometric(seed, n)
    yield "delta<1024,8,dynamic_vbp<512,64,8>>", default_delta_dynamic_vbp_f(8), x, int(x[n // 2])
    x = datagen.for_base_offset(seed, n)
    yield "for<1024,8,dynamic_vbp<512,64,8>>", default_for_dynamic_vbp_f(8), x, int(x.min()) + (1 << 15)
    x = datagen.rle_runs(seed, n)
    yield "rle", rle_f(), x, 1 << 19
    if with_uncompr:
        # reference point: the same select + sum over uncompressed values
        yield "uncompr", uncompr_f(), datagen.uniform_bits(seed, n, 32), 1 << 31


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cols, threads_all, reps=5):
    """The same 4-column step on the host cores: the CPU baseline's
    streaming select + sum (oracle/ms_cpu.c) over the GPU-produced images,
    in its AVX-512 style (scalar where the CPU lacks AVX-512) and its scalar
    style, on 1 thread (the reference's execution model) and on all the
    process's cores; median of `reps` steps after one warm-up."""
    from oracle import oracle as orc
    hosts = []
    for c in cols:
        img, meta = c["col"].image()
        hosts.append((orc.from_image(c["fmt"], img, *meta), c["val"], c["sum"]))
    styles = [("avx512", 1)] if orc.cpu_has_avx512() else []
    styles.append(("scalar", 0))
    out = {}
    for style, simd in styles:
        for threads in (1, threads_all):
            def step():
                return [orc.cpu_select_sum(CMP_LT, v, h, threads, simd) for h, v, _ in hosts]
            got = step()
            if got != [w for _, _, w in hosts]:
                raise RuntimeError(f"CPU baseline ({style}, {threads} threads) result mismatch")
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                step()
                ts.append(time.perf_counter() - t0)
            out[(style, threads)] = statistics.median(ts)
    return out, styles[0][0]


def timed(fn, steps, warmup, hip):
    """Average ms of fn() over `steps` calls, bracketed by HIP events on the
    default stream (every call below issues its work there)."""
    for _ in range(warmup):
        fn()
    hip.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    return e0.elapsed_ms(e1) / steps


def codec_rates(cols, n, args, ea, hip, datagen):
    """Column-level morph of each sweep layout: decompression (format ->
    uncompr_f) and compression (uncompr_f -> format), through ms_morph with
    the columns resident in HBM.  Algorithmic bytes of either direction:
    sizeUsed + 8 n.  The decompressed values are checked against the input
    hash and the recompressed image against the original's hash first."""
    from engine_amd.formats import uncompr_f
    out = {}
    for c in cols:
        raw = ea.morph(uncompr_f(), c["col"])
        vals = raw.values()
        if datagen.column_hash(vals) != c["hash"]:
            raise RuntimeError(f"{c['name']}: decompressed values differ from the input")
        del vals
        again = ea.morph(c["fmt"], raw)
        if datagen.column_hash(again.image()[0].view(np.uint8)) != c["image_hash"]:
            raise RuntimeError(f"{c['name']}: recompressed image differs")
        del again
        ms_d = timed(lambda: ea.morph(uncompr_f(), c["col"]), args.steps, args.warmup, hip)
        ms_c = timed(lambda: ea.morph(c["fmt"], raw), args.steps, args.warmup, hip)
        alg = c["size_used"] + 8 * n
        out[c["name"]] = {
            "algorithmic_bytes": alg,
            "decompress_us": round(ms_d * 1e3, 2), "decompress_GBps": round(alg / ms_d / 1e6, 1),
            "decompress_frac": round(alg / ms_d / 1e6 / HBM_PEAK_GBS, 4),
            "compress_us": round(ms_c * 1e3, 2), "compress_GBps": round(alg / ms_c / 1e6, 1),
            "compress_frac": round(alg / ms_c / 1e6 / HBM_PEAK_GBS, 4)}
        log(f"{c['name']}: decompress {ms_d * 1e3:.1f} us ({alg / ms_d / 1e6:.1f} GB/s), "
            f"compress {ms_c * 1e3:.1f} us ({alg / ms_c / 1e6:.1f} GB/s)")
        del raw
    return out


def c1_workload(args, ea, hip, datagen):
    """2^24 values uniform in [0, 2^12) as static_vbp<12,1>: the reference's
    query shape select(x < 2^11) -> positions -> project -> agg_sum,
    materialised (three operator calls) and fused (ms_select_sum)."""
    from engine_amd.formats import static_vbp_f, uncompr_f
    n = 1 << 24
    x = datagen.uniform_bits(datagen.BASE_SEED + 0, n, 12)
    col = ea.morph(static_vbp_f(12, 1), ea.upload_values(x))
    want = masked_sum(x, 1 << 11)
    su = col.meta().size_used

    def query():
        pos = ea.select(CMP_LT, 1 << 11, uncompr_f(), col, 1)
        dat = ea.project(uncompr_f(), col, pos, 1)
        return ea.agg_sum(dat)
    got = int(query().values()[0])
    res = ea.DeviceU64()
    fused = int(ea.select_sum(CMP_LT, 1 << 11, col, res).get()[0])
    if got != want or fused != want:
        raise RuntimeError(f"c1 sums differ: {got} {fused} {want}")
    ms_q = timed(query, max(2, args.steps // 2), 1, hip)
    ms_f = timed(lambda: ea.select_sum(CMP_LT, 1 << 11, col, res), args.steps, args.warmup, hip)
    log(f"c1: materialised query {ms_q * 1e3:.1f} us, fused {ms_f * 1e3:.1f} us ({su / ms_f / 1e6:.1f} GB/s)")
    return {"workload": "2^24 uint64 values uniform in [0, 2^12), static_vbp<12,1>, select(x < 2^11) -> "
                        "project -> agg_sum", "size_used": su, "input_hash": datagen.column_hash(x),
            "materialised_us": round(ms_q * 1e3, 2), "fused_us": round(ms_f * 1e3, 2),
            "fused_GBps": round(su / ms_f / 1e6, 1)}


def c3_workload(args, ea, hip, datagen):
    """Two-column plan over 2^k rows (key U[0,2^20), measure U[0,2^16), flag
    Zipf(1.1) over 8 values, each dynamic_vbp<512,2048,8>):
    select(flag==3) -> delta+dvbp positions, intersect with select(key<2^18),
    project(measure), agg_sum -- recorded once, replayed as one hipGraph."""
    from engine_amd.formats import default_delta_dynamic_vbp_f, dynamic_vbp_f, uncompr_f
    n = 1 << args.c3_log2n
    seed = datagen.BASE_SEED + 2
    f = dynamic_vbp_f(512, 2048, 8)
    fl = datagen.zipf_flags(seed, n)
    ky = datagen.uniform_bits(seed + 1, n, 20)
    ms = datagen.uniform_bits(seed + 2, n, 16)
    sel = (fl == 3) & (ky < (1 << 18))
    with np.errstate(over="ignore"):
        want = int(np.sum(ms[sel], dtype=np.uint64))
    nsel = int(sel.sum())
    st = hip.Stream()
    cols = [ea.morph(f, ea.upload_values(v), st.handle) for v in (fl, ky, ms)]
    del fl, ky, ms, sel
    in_bytes = sum(c.meta(st.handle).size_used for c in cols)
    pos_f = default_delta_dynamic_vbp_f(8)

    def query(stream):
        p1 = ea.select(0, 3, pos_f, cols[0], 8, stream)
        p2 = ea.select(CMP_LT, 1 << 18, pos_f, cols[1], 8, stream)
        i = ea.intersect_sorted(pos_f, p1, p2, stream)
        m = ea.project(uncompr_f(), cols[2], i, 8, stream)
        return ea.agg_sum(m, stream), (p1, p2, i, m)
    r, keep = query(st.handle)
    got = int(r.values(st.handle)[0])
    if got != want:
        raise RuntimeError(f"c3 sum {got} != {want}")
    inter = [k.meta(st.handle).size_used for k in keep]
    e0, e1 = hip.Event(), hip.Event()
    st.synchronize()
    e0.record(st.handle)
    for _ in range(3):
        query(st.handle)
    e1.record(st.handle)
    ms_eager = e0.elapsed_ms(e1) / 3
    plan = ea.Plan()
    with plan.record(st.handle):
        query(st.handle)
    with plan.record(st.handle):
        res, keep2 = query(st.handle)
    plan.launch(st.handle)
    st.synchronize()
    if int(res.values(st.handle)[0]) != want:
        raise RuntimeError("c3 plan replay differs")
    plan.launch(st.handle)
    st.synchronize()
    e0.record(st.handle)
    reps = max(2, args.steps // 2)
    for _ in range(reps):
        plan.launch(st.handle)
    e1.record(st.handle)
    ms_graph = e0.elapsed_ms(e1) / reps
    if int(res.values(st.handle)[0]) != want:
        raise RuntimeError("c3 plan replay differs after timing")
    # algorithmic bytes: inputs once, compressed position lists written and read
    algo = in_bytes + 2 * sum(inter[:3]) + inter[3] * 2
    log(f"c3: {n} rows, {nsel} selected, eager {ms_eager:.3f} ms, plan (hipGraph) {ms_graph:.3f} ms, "
        f"{algo / ms_graph / 1e6:.1f} GB/s algorithmic")
    plan.free()
    return {"workload": f"2^{args.c3_log2n} rows x 3 dynamic_vbp columns, select(flag==3) & select(key<2^18) "
                        f"-> intersect -> project(measure) -> agg_sum as one hipGraph",
            "input_bytes": in_bytes, "intermediate_bytes": inter, "selected": nsel,
            "eager_ms": round(ms_eager, 3), "graph_ms": round(ms_graph, 3),
            "algorithmic_GBps": round(algo / ms_graph / 1e6, 1)}


def c4_workload(args, ea, hip, datagen):
    """Uneven mixed-format columns: per-column calls vs one tuned batch."""
    t0 = time.perf_counter()
    cols, expect, bytes_used, nvals, kinds = [], [], 0, 0, {}
    for kind, fmt, x, val in datagen.mixed_columns(datagen.BASE_SEED + 3, args.c4_cols):
        col = ea.morph(fmt, ea.upload_values(x))
        cols.append((col, val))
        expect.append(masked_sum(x, val))
        m = col.meta()
        bytes_used += m.size_used
        nvals += len(x)
        kinds[kind] = kinds.get(kind, 0) + 1
    nc = len(cols)
    log(f"c4: {nc} columns, {nvals} values, {bytes_used} compressed bytes, prepared in "
        f"{time.perf_counter() - t0:.1f}s")
    res = ea.DeviceU64(nc)
    slots = [res.at(k) for k in range(nc)]

    def plain():
        for k, (col, val) in enumerate(cols):
            ea.select_sum(CMP_LT, val, col, slots[k])

    def verify(tag):
        got = res.get().tolist()
        bad = [k for k in range(nc) if got[k] != expect[k]]
        if bad:
            log(f"c4 MISMATCH ({tag}) at entries {bad[:5]}")
            sys.exit(1)

    plain()
    verify("plain")
    ms_plain = timed(plain, max(2, args.steps // 10), 1, hip)
    # the same per-column calls recorded once and replayed as one hipGraph
    # (left out of profiling runs: --no-graph-legs)
    ms_graph = None
    if not args.no_graph_legs:
        st = hip.Stream()
        plan = ea.Plan()

        def plain_on(stream):
            for k, (col, val) in enumerate(cols):
                ea.select_sum(CMP_LT, val, col, slots[k], stream)

        with plan.record(st.handle):
            plain_on(st.handle)
        with plan.record(st.handle):
            plain_on(st.handle)
        res.zero()                 # the replay must write every result
        plan.launch(st.handle)
        st.synchronize()
        verify("plain graph")
        for _ in range(2):
            plan.launch(st.handle)
        st.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        reps = max(2, args.steps // 4)
        e0.record(st.handle)
        for _ in range(reps):
            plan.launch(st.handle)
        e1.record(st.handle)
        ms_graph = e0.elapsed_ms(e1) / reps
        plan.free()
    # one batch, autotuned
    b = ea.Batch()
    for col, val in cols:
        b.enqueue(col, CMP_LT, val)
    res.zero()
    b.run(res)
    verify("batched default")
    t1 = time.perf_counter()
    cfg = b.tune(res, reps=3)
    tune_s = time.perf_counter() - t1
    b.run(res)
    verify("batched tuned")
    ms_batch = timed(lambda: b.run(res), args.steps, args.warmup, hip)
    _, items, _ = b.stats()
    gb = lambda ms: round((bytes_used + 8 * nc) / ms / 1e6, 1) if ms else None  # noqa: E731
    log(f"c4: plain {ms_plain:.3f} ms ({gb(ms_plain)} GB/s), plain graph {ms_graph} ms ({gb(ms_graph)} GB/s), "
        f"batched {ms_batch:.3f} ms ({gb(ms_batch)} GB/s), tuned {cfg.as_dict()} in {tune_s:.1f}s")
    return {
        "workload": f"{nc} uint64 columns, lengths log-uniform in [10^3, 10^6], formats {kinds}, "
                    f"select(x<c)+agg_sum per column",
        "values": nvals, "compressed_bytes": bytes_used,
        "plain_ms": round(ms_plain, 4), "plain_GBps": gb(ms_plain),
        "plain_graph_ms": round(ms_graph, 4) if ms_graph else None, "plain_graph_GBps": gb(ms_graph),
        "batched_ms": round(ms_batch, 4), "batched_GBps": gb(ms_batch),
        "batched_pct_hbm_peak": round(100 * gb(ms_batch) / HBM_PEAK_GBS, 2),
        "speedup_vs_plain": round(ms_plain / ms_batch, 2), "speedup_vs_plain_graph": round(ms_graph / ms_batch, 2) if ms_graph else None,
        "work_items": items, "tuned_config": cfg.as_dict(), "tune_seconds": round(tune_s, 1),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--c4-cols", type=int, default=4096, help="columns of the uneven workload (0: skip it)")
    ap.add_argument("--c3-log2n", type=int, default=28, help="rows of the two-column plan workload (0: skip it)")
    ap.add_argument("--no-c1", action="store_true", help="skip the 2^24-value static_vbp query")
    ap.add_argument("--c5-cols", type=int, default=1024,
                    help="columns of 64 x 2^16 values of the learned-selection workload (0: skip it)")
    ap.add_argument("--c5-train", type=int, default=32768,
                    help="training chunks of the selector (each compressed and timed in every candidate format)")
    ap.add_argument("--batch-config", default=None,
                    help="JSON ms_batch_config for the sweep's batch instead of autotuning (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner comparison of the sweep")
    ap.add_argument("--no-codecs", action="store_true", help="skip the decompression / compression rates")
    ap.add_argument("--no-graph-legs", action="store_true",
                    help="skip the per-column calls replayed as one hipGraph (profiling runs)")
    ap.add_argument("--only", default=None, help="only the layouts whose name starts with this (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards per column/page "
              f"across GPUs by design, but this build measures one GPU's share.", file=sys.stderr)
        sys.exit(2)

    import engine_amd as ea
    from engine_amd import datagen, hip

    ea.init(0)
    n = 1 << args.log2n
    seed = datagen.BASE_SEED + 1
    cols = []
    for name, fmt, x, val in make_layouts(n, seed, with_uncompr=args.only == "uncompr"):
        if args.only and not name.startswith(args.only):
            continue
        t0 = time.perf_counter()
        raw = ea.upload_values(x)
        raw.meta()
        up_s = time.perf_counter() - t0
        col = ea.morph(fmt, raw) if fmt.kind else raw
        del raw
        m = col.meta()
        expect = masked_sum(x, val)
        got = int(ea.select_sum(CMP_LT, val, col).get()[0])
        if got != expect:
            log(f"MISMATCH {name}: gpu {got} numpy {expect}")
            sys.exit(1)
        cols.append(dict(name=name, fmt=fmt, col=col, val=val, size_used=m.size_used, count=m.count,
                         hash=datagen.column_hash(x), sum=expect, upload_GBps=8 * n / up_s / 1e9,
                         image_hash=datagen.column_hash(col.image()[0].view(np.uint8)) if fmt.kind else None))
        log(f"{name}: n={n} sizeUsed={m.size_used} ({8 * n / m.size_used:.2f}x) prepared in "
            f"{time.perf_counter() - t0:.1f}s, sum verified")
        del x
    nc = len(cols)
    res = ea.DeviceU64(nc)
    slots = [res.at(k) for k in range(nc)]
    # per layout: one fused call
    per = {}
    for k, c in enumerate(cols):
        ms = timed(lambda: ea.select_sum(CMP_LT, c["val"], c["col"], slots[k]), args.steps, args.warmup, hip)
        per[c["name"]] = ms
        log(f"{c['name']}: {ms * 1e3:.1f} us/call, {c['size_used'] / ms / 1e6:.1f} GB/s compressed, "
            f"{8 * n / ms / 1e6:.1f} GB/s uncompressed-equivalent")

    def plain_step():
        for k, c in enumerate(cols):
            ea.select_sum(CMP_LT, c["val"], c["col"], slots[k])

    ms_plain = timed(plain_step, args.steps, args.warmup, hip)
    # batched: the same four select+sums as one persistent launch
    batch = ea.Batch()
    for c in cols:
        batch.enqueue(c["col"], CMP_LT, c["val"])
    batch.run(res)
    if res.get().tolist() != [c["sum"] for c in cols]:
        log("MISMATCH batched step")
        sys.exit(1)
    from engine_amd.formats import BatchConfig
    # the library default, then the autotuner's pick, then the fixed shape
    # (a profiling run, --no-tune, launches the fixed shape only)
    autotune = None
    if not args.no_tune:
        ms_default = timed(lambda: batch.run(res), args.steps, args.warmup, hip)
        t1 = time.perf_counter()
        tuned = batch.tune(res, reps=5)
        tune_s = time.perf_counter() - t1
        batch.run(res)
        if res.get().tolist() != [c["sum"] for c in cols]:
            log("MISMATCH batched step (tuned)")
            sys.exit(1)
        ms_tuned = timed(lambda: batch.run(res), args.steps, args.warmup, hip)
        autotune = {"default_us": round(ms_default * 1e3, 2), "tuned_us": round(ms_tuned * 1e3, 2),
                    "tuned_over_default": round(ms_default / ms_tuned, 3), "tuned_config": tuned.as_dict(),
                    "tune_seconds": round(tune_s, 2)}
        log(f"autotune: default {ms_default * 1e3:.1f} us, tuned {ms_tuned * 1e3:.1f} us {tuned.as_dict()}")
    batch.config = BatchConfig(**(json.loads(args.batch_config) if args.batch_config else SWEEP_CFG))
    cfg = batch.config
    batch.run(res)
    if res.get().tolist() != [c["sum"] for c in cols]:
        log("MISMATCH batched step (tuned)")
        sys.exit(1)
    wall0 = time.perf_counter()
    ms_step = timed(lambda: batch.run(res), args.steps, args.warmup, hip)
    wall = time.perf_counter() - wall0
    # the persistent kernel alone (HIP events around its launches, on the
    # stream it runs on), over the same number of runs
    batch.timing(args.steps)
    for _ in range(args.steps):
        batch.run(res)
    kms = batch.kernel_times(args.steps)
    batch.timing(0)
    ms_kernel = statistics.mean(kms) if kms else ms_step
    total_bytes = sum(c["size_used"] + 8 for c in cols)
    value = total_bytes / ms_step / 1e6
    value_plain = total_bytes / ms_plain / 1e6
    log(f"step: plain {ms_plain * 1e3:.1f} us ({value_plain:.1f} GB/s), batched {ms_step * 1e3:.1f} us "
        f"({value:.1f} GB/s), tuned {cfg.as_dict()}")
    achieved = total_bytes / ms_kernel / 1e6
    # HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    # profiles/) at this exact shape; null when the shape differs
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_batch_select_sum.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("batch_config") == cfg.as_dict() and t.get("algorithmic_bytes") == total_bytes:
            traffic = t.get("bytes_per_launch")
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "format cascade sweep, fused select(x<c)+agg_sum over %d uint64 columns of 2^%d "
                               "values resident in HBM, one batched persistent launch per step" % (nc, args.log2n),
                   "columns": {c["name"]: {"size_used": c["size_used"], "ratio": round(8 * n / c["size_used"], 3),
                                           "plain_us_per_call": round(per[c["name"]] * 1e3, 2),
                                           "plain_GBps_compressed": round(c["size_used"] / per[c["name"]] / 1e6, 1),
                                           "input_hash": c["hash"]} for c in cols},
                   "batch_config": cfg.as_dict()},
        "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 2),
        "plain": {"ms_per_step": round(ms_plain, 4), "value": round(value_plain, 1),
                  "pct_hbm_peak": round(100 * value_plain / HBM_PEAK_GBS, 2)},
        "batched_vs_plain": round(ms_plain / ms_step, 3),
        "uncompr_equiv_GBps": round(nc * 8 * n / ms_step / 1e6, 1),
        "host_upload_GBps": round(statistics.median(c["upload_GBps"] for c in cols), 2) if cols else None,
        "wall_s_timed": round(wall, 3),
        "roofline": {"bound": "hbm", "kernel": "k_batch_select_sum",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel_us": round(ms_kernel * 1e3, 2), "algorithmic_bytes": total_bytes},
        "autotune": autotune,
    }
    if not args.no_codecs and cols:
        try:
            out["codecs"] = codec_rates(cols, n, args, ea, hip, datagen)
        except Exception as e:     # the sweep's line above stands on its own
            log(f"codec measurement failed: {e!r}")
            out["codecs"] = {"error": repr(e)}
    extra = [("static_vbp_query", c1_workload, not args.no_c1), ("plan_query", c3_workload, args.c3_log2n > 0),
             ("uneven_columns", c4_workload, args.c4_cols > 0)]
    for key, fn, on in extra:
        if on and not args.only:
            try:
                out[key] = fn(args, ea, hip, datagen)
            except Exception as e:     # the sweep's line above stands on its own
                log(f"{key} workload failed: {e!r}")
                out[key] = {"error": repr(e)}
    if args.c5_cols > 0 and not args.only:
        from engine_amd import learn
        try:
            out["learned_selection"] = learn.run_selection_workload(ncols=args.c5_cols, ntrain=args.c5_train,
                                                                    steps=max(2, args.steps // 2), log=log,
                                                                    graph_leg=not args.no_graph_legs)
        except Exception as e:     # the sweep's line above stands on its own
            log(f"learned-selection workload failed: {e!r}")
            out["learned_selection"] = {"error": repr(e)}
    if not args.no_cpu and cols:
        threads_all = min(len(os.sched_getaffinity(0)), 16)
        r, best = cpu_baseline(cols, threads_all)
        gbs = lambda sec: round(total_bytes / sec / 1e9, 3)  # noqa: E731
        out["cpu_baseline"] = {
            "value": gbs(r[(best, 1)]), "unit": "GB/s", "cores": 1, "kind": "port",
            "style": best, "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"the same step: select(x<c)+sum over the {nc} sweep columns of 2^{args.log2n} values "
                      f"({total_bytes} compressed bytes), streaming restatement of the reference's "
                      f"decompress_and_process_batch ({best}), median of 5 steps",
            "all_cores": {"cores": threads_all, "value": gbs(r[(best, threads_all)])},
            "styles": {f"{st}_{th}t": gbs(sec) for (st, th), sec in r.items()},
        }
        log(f"cpu baseline ({out['cpu_baseline']['cpu']}): {out['cpu_baseline']['styles']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
