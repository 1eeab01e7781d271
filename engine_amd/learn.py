"""Learned per-chunk format selection: label generation, training,
selection (host side of the ms_gen_chunks / ms_chunk_* / ms_mlp_* /
ms_select_formats entry points).

The reference chooses formats from calibration runs (src/calibration/
bw_prof.cpp:71-292: time and size per format and bit width) and an
exhaustive size search (exhaustive_size_format_selector.h:35-141).  Here the
engine
  1. labels chunks of 2^16 values generated on the device: exact size of
     every candidate (closed forms from per-block bit widths) and decode
     time, measured by compressing the chunks in every candidate and
     timing them in the batched select+sum kernel's timed variant
     (labels="measured"), or from a per-(format, bit width) timing table
     (labels="table": ns per 512-value unit, ns per rle run);
  2. trains the FP32 MLP on the device (24 features -> log2 size and log2
     time of the 5 candidates);
  3. selects per chunk, rule 0: the smallest predicted candidate, or one
     predicted within `size_tol` of it and faster by more than
     `time_margin`; exact sizes check the pick and replace it by the exact
     smallest when it is more than `exact_tol` larger.  Rule 1 (the
     selection workload's): the exact smallest, replaced by the candidate
     predicted fastest among those whose exact size is within `exact_tol`
     of it, when that one is predicted faster by more than `time_margin`
     (the exact sizes come from the same statistics pass as the features).
Whatever is picked, the data stays bit-exact: only the format changes.
"""
import ctypes

import numpy as np

import engine_amd as ea

from . import hip
from .datagen import splitmix64

CHUNK = 1 << 16
NCAND, NFEAT, NOUT = 5, 24, 10
GEO_LEN = 1024
(GEN_UNIFORM, GEN_DVBP, GEN_SORTED, GEN_FOR, GEN_RLE) = range(5)
CAND_NAMES = ("static_vbp", "dynamic_vbp", "delta", "for", "rle")
NSTAT = 16
ST_MAX_BW = 1
MLP_NPARAM = 24 * 64 + 64 + 64 * 64 + 64 + 64 * 10 + 10
# target normalisation: y = (log2 value - off) / scale
SIZE_OFF, SIZE_SCALE = 14.0, 4.0
TIME_OFF, TIME_SCALE = 5.0, 3.0

U32, U64 = ctypes.c_uint32, ctypes.c_uint64


class ChunkSpec(ctypes.Structure):
    _fields_ = [("kind", U32), ("a", U32), ("b", U32), ("table", U32), ("seed", U64), ("base", U64)]


class SelectParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("size_off", "size_scale", "time_off", "time_scale", "size_tol",
                                               "time_margin", "exact_tol")] + [("rule", ctypes.c_uint32)]


def geo_table(a):
    """Thresholds of a geometric variate with mean ~a: thr[0] = 2^64 - 1,
    thr[k] = floor(thr[k-1] * (a - 1) / a) (exact integer recurrence)."""
    t = [(1 << 64) - 1]
    for _ in range(1, GEO_LEN):
        t.append(t[-1] * (a - 1) // a)
    return np.array(t, dtype=np.uint64)


class GeoTables:
    """Geometric tables by mean, indexed in the order they are asked for."""

    def __init__(self):
        self.index, self.tabs = {}, []

    def __call__(self, a):
        if a not in self.index:
            self.index[a] = len(self.tabs)
            self.tabs.append(geo_table(a))
        return self.index[a]

    def array(self):
        return np.concatenate(self.tabs) if self.tabs else np.zeros(GEO_LEN, np.uint64)


SPEC_DTYPE = np.dtype([("kind", "<u4"), ("a", "<u4"), ("b", "<u4"), ("table", "<u4"), ("seed", "<u8"),
                       ("base", "<u8")])
assert SPEC_DTYPE.itemsize == ctypes.sizeof(ChunkSpec)


def spec_array(specs):
    """A ms_chunk_spec array from a list of tuples (or a spec array)."""
    if isinstance(specs, np.ndarray) and specs.dtype == SPEC_DTYPE:
        return specs
    return np.array([tuple(int(v) for v in sp) for sp in specs], dtype=SPEC_DTYPE)


def gen_chunks(specs, tables, d_out, offset_chunks=0, stream=None):
    """Writes len(specs) chunks into d_out (a hip.DeviceArray of uint64)."""
    arr = np.ascontiguousarray(spec_array(specs))
    tab = np.ascontiguousarray(tables.array())
    ea.check(ea.lib().ms_gen_chunks(d_out.ptr(offset_chunks * CHUNK), arr.ctypes.data, len(arr), tab.ctypes.data,
                                    len(tables.tabs), stream))


def chunk_stats(d_vals, nchunks, d_stats, cost=None, d_tl=None, stream=None):
    c = None if cost is None else np.ascontiguousarray(cost, dtype=np.float32).reshape(-1)
    ea.check(ea.lib().ms_chunk_stats(d_vals if isinstance(d_vals, ctypes.c_void_p) else d_vals.ptr(), nchunks,
                                     d_stats.ptr() if hasattr(d_stats, "ptr") else d_stats,
                                     None if c is None else c.ctypes.data,
                                     None if d_tl is None else (d_tl.ptr() if hasattr(d_tl, "ptr") else d_tl),
                                     stream))


def chunk_features(d_stats, nchunks, d_feat, d_sizes, stream=None):
    p = lambda d: None if d is None else (d.ptr() if hasattr(d, "ptr") else d)  # noqa: E731
    ea.check(ea.lib().ms_chunk_features(p(d_stats), nchunks, p(d_feat), p(d_sizes), stream))


def candidate_format(cand, static_bw=1):
    f = ea.FormatDesc()
    ea.check(ea.lib().ms_candidate_format(cand, static_bw, ctypes.byref(f)))
    return f


class MLP:
    """The device network (ms_mlp_*)."""

    def __init__(self, seed=0x4D4C50):
        ea.init()
        self.h = ctypes.c_void_p()
        ea.check(ea.lib().ms_mlp_create(ctypes.byref(self.h), seed))

    @property
    def params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        ea.check(ea.lib().ms_mlp_get_params(self.h, p.ctypes.data))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == MLP_NPARAM
        ea.check(ea.lib().ms_mlp_set_params(self.h, p.ctypes.data))

    def loss_grad(self, d_x, d_y, n, wout, stream=None):
        w = np.ascontiguousarray(wout, dtype=np.float32)
        g = hip.DeviceArray(MLP_NPARAM, np.float32)
        loss = hip.DeviceArray(1, np.float32)
        ea.check(ea.lib().ms_mlp_loss_grad(self.h, d_x.ptr(), d_y.ptr(), w.ctypes.data, n, g.ptr(), loss.ptr(),
                                           stream))
        return float(loss.numpy()[0]), g.numpy()

    def train(self, d_x, d_y, n, wout, epochs, batch=4096, lr=2e-3, seed=1, stream=None):
        w = np.ascontiguousarray(wout, dtype=np.float32)
        losses = np.zeros(max(epochs, 1), np.float32)
        ea.check(ea.lib().ms_mlp_train(self.h, d_x.ptr(), d_y.ptr(), w.ctypes.data, n, epochs, batch,
                                       ctypes.c_float(lr), seed, losses.ctypes.data, stream))
        return losses[:epochs]

    def predict(self, d_x, n, stream=None):
        d_y = hip.DeviceArray(n * NOUT, np.float32)
        ea.check(ea.lib().ms_mlp_predict(self.h, d_x.ptr(), n, d_y.ptr(), stream))
        return d_y

    def __del__(self):
        if self.h and ea._lib is not None:
            ea._lib.ms_mlp_free(self.h)
            self.h = None


def select_formats(d_pred, d_sizes, n, size_tol=0.02, time_margin=0.15, exact_tol=0.03, rule=0, stream=None):
    """Per chunk: candidate index, and whether the exact smallest was taken
    (rule 0: the exact-size fallback fired; rule 1: no admitted candidate
    was predicted enough faster)."""
    sp = SelectParams(SIZE_OFF, SIZE_SCALE, TIME_OFF, TIME_SCALE, size_tol, time_margin, exact_tol, rule)
    d_c = hip.DeviceArray(n, np.uint8)
    ea.check(ea.lib().ms_select_formats(d_pred.ptr(), d_sizes.ptr(), n, ctypes.byref(sp), d_c.ptr(), stream))
    c = d_c.numpy()
    return (c & 0x7f).astype(np.int64), (c & 0x80) != 0


def targets(sizes, times):
    """Network targets from exact sizes (bytes) and time labels (ns)."""
    ys = (np.log2(np.maximum(sizes.astype(np.float64), 1.0)) - SIZE_OFF) / SIZE_SCALE
    yt = (np.log2(np.maximum(times.astype(np.float64), 1e-3)) - TIME_OFF) / TIME_SCALE
    return np.concatenate([ys, yt], axis=1).astype(np.float32)


# ------------------------------------------------------------------ specs
def _u(seed, n):
    return splitmix64(seed, n)


def training_specs(seed, n, tables):
    """n chunk specs: every generator kind with parameters spread over the
    ranges the selector must handle."""
    r = _u(seed, 6 * n).reshape(n, 6)
    sorted_means = np.array((2, 3, 4, 10, 32, 100, 1000), np.uint64)
    rle_means = np.array((1, 2, 4, 8, 16, 64, 256, 1024), np.uint64)
    for a in sorted_means.tolist() + rle_means.tolist():
        tables(a)
    tab_of = np.zeros(1025, np.uint32)
    for a, k in tables.index.items():
        if a <= 1024:
            tab_of[a] = k
    kind = (r[:, 0] % np.uint64(5)).astype(np.uint32)
    q2, q3, q4 = r[:, 2], r[:, 3], r[:, 4]
    out = np.zeros(n, SPEC_DTYPE)
    out["kind"] = kind
    out["seed"] = r[:, 1]
    # uniform: a ~ U{0..64}
    a = (q2 % np.uint64(65)).astype(np.uint32)
    b = np.zeros(n, np.uint32)
    tab = np.zeros(n, np.uint32)
    base = np.zeros(n, np.uint64)
    m = kind == GEN_DVBP
    lo = (1 + q2 % np.uint64(48)).astype(np.uint32)
    a[m] = lo[m]
    b[m] = np.minimum(64, lo + (q3 % np.uint64(33)).astype(np.uint32))[m]
    shifted = q3 >> (q4 % np.uint64(64))
    m = kind == GEN_SORTED
    sm = sorted_means[(q2 % np.uint64(len(sorted_means))).astype(np.int64)]
    a[m] = sm[m].astype(np.uint32)
    tab[m] = tab_of[sm[m].astype(np.int64)]
    base[m] = shifted[m]
    m = kind == GEN_FOR
    a[m] = (1 + q2 % np.uint64(40)).astype(np.uint32)[m]
    base[m] = shifted[m]
    m = kind == GEN_RLE
    rm = rle_means[(q2 % np.uint64(len(rle_means))).astype(np.int64)]
    a[m] = rm[m].astype(np.uint32)
    b[m] = (1 + q3 % np.uint64(64)).astype(np.uint32)[m]
    tab[m] = tab_of[rm[m].astype(np.int64)]
    out["a"], out["b"], out["table"], out["base"] = a, b, tab, base
    return out


def mixture_specs(seed, col, nchunks, tables):
    """Chunks of one column of the selection workload: each 2^16-value chunk
    one of the format-cascade-sweep distributions (dynamic_vbp bit widths
    U{1..32}, sorted with geometric(0.1) gaps, base U[2^40, 2^41) + 16-bit
    offsets, rle runs of mean 64 over 20-bit values) or uniform b-bit values."""
    r = _u(seed ^ (col * 0x10001 + 0xC5), 4 * nchunks).reshape(nchunks, 4)
    t10, t64 = tables(10), tables(64)
    specs = []
    for i in range(nchunks):
        k = int(r[i, 0] % np.uint64(5))
        s, q = int(r[i, 1]), int(r[i, 2])
        if k == GEN_UNIFORM:
            specs.append((k, 1 + q % 32, 0, 0, s, 0))
        elif k == GEN_DVBP:
            specs.append((k, 1, 32, 0, s, 0))
        elif k == GEN_SORTED:
            specs.append((k, 10, 0, t10, s, q >> 24))
        elif k == GEN_FOR:
            specs.append((k, 16, 0, 0, s, (1 << 40) + (q >> 24)))
        else:
            specs.append((k, 64, 20, t64, s, 0))
    return specs


# ------------------------------------------------------------------ calibration
def _calib_values(cand, bw, nchunks, seed):
    n = nchunks * CHUNK
    r = splitmix64(seed, n)
    u = r >> np.uint64(64 - bw)
    if cand == 2:                      # delta: lane-wise running sums of bw-bit deltas
        with np.errstate(over="ignore"):
            return np.cumsum(u.reshape(-1, 8), axis=0, dtype=np.uint64).reshape(-1)
    if cand == 3:                      # for: base + bw-bit offsets
        return np.uint64(1 << 40) + (u >> np.uint64(1) if bw == 64 else u)
    return u


def calibrate(nchunks=16, copies=64, reps=5, stream=None, log=None):
    """ns per 512-value unit of the batched fused select+sum, per candidate
    and bit width (1..64), and rle's fixed and per-run cost: a MS_NCAND x 65
    table for ms_chunk_stats."""
    cost = np.zeros((NCAND, 65), np.float32)
    res = ea.DeviceU64(copies)

    def time_batch(col, val):
        b = ea.Batch()
        for _ in range(copies):
            b.enqueue(col, 2, val)
        b.run(res)
        e0, e1 = hip.Event(), hip.Event()
        hip.synchronize()
        e0.record()
        for _ in range(reps):
            b.run(res)
        e1.record()
        return e0.elapsed_ms(e1) / reps * 1e6      # ns per run
    empty = ea.upload_values(np.zeros(0, np.uint64))
    t0 = time_batch(empty, 1)
    units = copies * nchunks * CHUNK // 512
    for cand in range(4):
        for bw in range(1, 65):
            x = _calib_values(cand, bw, nchunks, 0xCA1 + 97 * cand + bw)
            f = candidate_format(cand, bw)
            col = ea.morph(f, ea.upload_values(x))
            val = int(np.median(x[::97]))
            cost[cand, bw] = max(time_batch(col, val) - t0, 1e-3) / units
        cost[cand, 0] = cost[cand, 1]
        if log:
            log(f"calibrated {CAND_NAMES[cand]}: ns/unit bw1 {cost[cand, 1]:.3f} bw16 {cost[cand, 16]:.3f} "
                f"bw32 {cost[cand, 32]:.3f} bw64 {cost[cand, 64]:.3f}")
    # rle: per-chunk time = fixed + per_run * runs, least squares over run lengths
    from .datagen import rle_runs
    pts = []
    for mean in (2, 8, 32, 128, 512):
        x = rle_runs(0xCA2 + mean, nchunks * CHUNK, mean_len=mean)
        col = ea.morph(candidate_format(4), ea.upload_values(x))
        runs_per_chunk = col.meta().size_used // 16 / nchunks
        pts.append((runs_per_chunk, (time_batch(col, 1 << 19) - t0) / (copies * nchunks)))
    A = np.array([[1.0, r] for r, _ in pts])
    y = np.array([t for _, t in pts])
    fixed, per_run = np.linalg.lstsq(A, y, rcond=None)[0]
    cost[4, 0], cost[4, 1] = max(fixed, 0.0), max(per_run, 1e-6)
    if log:
        log(f"calibrated rle: fixed {cost[4, 0]:.3f} ns/chunk, {cost[4, 1]:.4f} ns/run")
    return cost


# ------------------------------------------------------------------ labels
def label(specs, tables, cost, batch=2048, stream=None):
    """Generates the chunks of `specs` on the device in batches and returns
    device arrays (features n x 24, targets n x 10) and host exact sizes and
    time labels."""
    specs = spec_array(specs)
    n = len(specs)
    d_feat = hip.DeviceArray(n * NFEAT, np.float32)
    sizes = np.zeros((n, NCAND), np.uint64)
    times = np.zeros((n, NCAND), np.float32)
    d_vals = hip.DeviceArray(batch * CHUNK, np.uint64)
    d_st = hip.DeviceArray(batch * NSTAT, np.uint64)
    d_sz = hip.DeviceArray(batch * NCAND, np.uint64)
    d_tl = hip.DeviceArray(batch * NCAND, np.float32)
    for c0 in range(0, n, batch):
        m = min(batch, n - c0)
        gen_chunks(specs[c0:c0 + m], tables, d_vals, 0, stream)
        chunk_stats(d_vals, m, d_st, cost, d_tl, stream)
        chunk_features(d_st, m, d_feat.ptr(c0 * NFEAT), d_sz, stream)
        sizes[c0:c0 + m] = d_sz.numpy(m * NCAND).reshape(m, NCAND)
        times[c0:c0 + m] = d_tl.numpy(m * NCAND).reshape(m, NCAND)
    d_y = hip.DeviceArray.from_numpy(targets(sizes, times).reshape(-1))
    return d_feat, d_y, sizes, times


def label_measured(specs, tables, batch=1024, reps=3, val=1 << 24, stream=None):
    """Labels the engine generates by compressing and timing: the chunks of
    `specs` are generated on the device, compressed in every candidate
    format (one blockwise_individual_f column per candidate, one chunk per
    block) and run through the persistent kernel's fused select + sum
    (x < val), one batch per candidate, which records the wall-clock time
    of every chunk's work items (ms_batch_view_times).  The times are scaled
    so that each batch's chunk times add up to the batch's run time measured
    by events: ns of device time per chunk.  Returns device features and targets, host exact sizes and
    the measured times."""
    from .formats import blockwise_individual_f
    bwi = blockwise_individual_f(CHUNK)
    specs = spec_array(specs)
    n = len(specs)
    d_feat = hip.DeviceArray(n * NFEAT, np.float32)
    sizes = np.zeros((n, NCAND), np.uint64)
    times = np.zeros((n, NCAND), np.float64)
    d_vals = hip.DeviceArray(batch * CHUNK, np.uint64)
    d_st = hip.DeviceArray(batch * NSTAT, np.uint64)
    d_sz = hip.DeviceArray(batch * NCAND, np.uint64)
    res = ea.DeviceU64(NCAND)
    for c0 in range(0, n, batch):
        m = min(batch, n - c0)
        gen_chunks(specs[c0:c0 + m], tables, d_vals, 0, stream)
        chunk_stats(d_vals, m, d_st, None, None, stream)
        chunk_features(d_st, m, d_feat.ptr(c0 * NFEAT), d_sz, stream)
        st = d_st.numpy(m * NSTAT).reshape(m, NSTAT)
        sizes[c0:c0 + m] = d_sz.numpy(m * NCAND).reshape(m, NCAND)
        whole = ea.from_device(d_vals.ptr(), m * CHUNK)
        for k in range(NCAND):
            # one candidate per batch: the chunks share the device with their
            # own kind only, and the batch's run time, measured by events,
            # is split over its chunks in proportion to their item times
            col = ea.morph_blockwise(bwi, whole, format_codes(np.full(m, k), st[:, ST_MAX_BW]))
            b = ea.Batch()
            b.enqueue(col, 2, val)
            b.run(res)
            hip.synchronize()
            e0, e1 = hip.Event(), hip.Event()
            e0.record()
            for _ in range(reps):
                b.run(res)
            e1.record()
            t_run = e0.elapsed_ms(e1) / reps * 1e6
            t = b.view_times(res, reps)
            times[c0:c0 + m, k] = t * (t_run / max(t.sum(), 1e-9))
            del b, col
        del whole
    d_y = hip.DeviceArray.from_numpy(targets(sizes, times).reshape(-1))
    return d_feat, d_y, sizes, times


# ------------------------------------------------------------------ workload
# blockwise_individual_f format codes of the candidates
# (format_selector_commons.h:25-103): static_vbp_f at the chunk's maximum
# bit width is that bit width; rle_f's pairs are this library's code 71.
_CAND_CODE = (None, 65, 67, 69, 71)


def format_codes(cand, max_bw):
    """Per-chunk candidate picks as blockwise_individual_f codes."""
    return np.array([int(b) if int(k) == 0 else _CAND_CODE[int(k)] for k, b in zip(cand, max_bw)], np.uint8)


def _masked_sum_col(col, val, res):
    ea.select_sum(2, val, col, res)
    return int(res.get()[0])


def column_level_choice(esz, max_bw):
    """Whole-column format for the per-column path: the candidate whose
    column size (sum of the chunks' exact sizes; static_vbp at the column's
    maximum bit width) is smallest."""
    tot = esz.sum(axis=0).astype(np.float64)
    tot[0] = len(esz) * (CHUNK // 8) * float(max_bw)
    k = int(np.argmin(tot))
    return k, int(max_bw)


def run_selection_workload(ncols=1024, chunks_per_col=64, ntrain=32768, epochs=60, seed=0x4D6F7270 + 4,
                           val=1 << 24, steps=10, warmup=2, log=None, cost=None, graph_leg=True, labels="measured",
                           size_weight=4.0, rule=1):
    """The learned-selection workload: ncols columns of chunks_per_col
    chunks (a mixture of the format-cascade-sweep distributions switching
    every 2^16 values).  Trains the selector on ntrain generated chunks
    (labels measured by the engine, or with labels="table" from the
    per-(format, bit width) timing calibration), picks a format per chunk, and stores every column as one
    blockwise_individual_f<2^16> column holding each chunk in its pick.
    Times SELECT SUM(x) WHERE x < val over every column: per-column calls
    on whole columns in their column-level best format (plain, eager and as
    one hipGraph) against one autotuned batch over the learned-pick columns
    and over the exact-best-pick columns.  Every path's per-column sums
    must equal the sums over the uncompressed columns."""
    import time
    from .formats import blockwise_individual_f
    say = log or (lambda *a: None)
    bwi = blockwise_individual_f(CHUNK)
    t0 = time.perf_counter()
    tables = GeoTables()
    specs = training_specs(seed, ntrain, tables)
    if labels == "measured":
        d_x, d_y, _, _ = label_measured(specs, tables, val=val)
    else:
        if cost is None:
            cost = calibrate(log=say)
        d_x, d_y, _, _ = label(specs, tables, cost)
    t_cal = time.perf_counter() - t0
    say(f"selector labels ({labels}): {ntrain} chunks in {t_cal:.1f}s")
    net = MLP(seed=seed)
    t1 = time.perf_counter()
    # Adam with the step size halved after each third of the epochs
    # the size head is weighted up: the exact-size check rejects picks more
    # than 3% above the smallest, while a time error only costs speed
    wout = np.concatenate([np.full(NCAND, size_weight), np.ones(NOUT - NCAND)]).astype(np.float32)
    losses = np.concatenate([net.train(d_x, d_y, ntrain, wout, epochs=e, batch=4096, lr=lr,
                                       seed=seed + k)
                             for k, (e, lr) in enumerate(((epochs // 3, 2e-3), (epochs // 3, 1e-3),
                                                          (epochs - 2 * (epochs // 3), 5e-4)))])
    hip.synchronize()
    t_train = time.perf_counter() - t1
    say(f"selector: {ntrain} labelled chunks, {epochs} epochs in {t_train:.1f}s, loss {losses[0]:.4f} -> "
        f"{losses[-1]:.4f}")
    del d_x, d_y
    n = chunks_per_col
    d_vals = hip.DeviceArray(n * CHUNK, np.uint64)
    d_st = hip.DeviceArray(n * NSTAT, np.uint64)
    d_f = hip.DeviceArray(n * NFEAT, np.float32)
    d_sz = hip.DeviceArray(n * NCAND, np.uint64)
    res1 = ea.DeviceU64()
    learned, exact, plain, want = [], [], [], []
    foot = {"learned": 0, "exact": 0, "plain": 0}
    fallbacks, picks = 0, np.zeros(NCAND, np.int64)
    t2 = time.perf_counter()
    for c in range(ncols):
        gen_chunks(mixture_specs(seed, c, n, tables), tables, d_vals)
        chunk_stats(d_vals, n, d_st)
        chunk_features(d_st, n, d_f, d_sz)
        st = d_st.numpy(n * NSTAT).reshape(n, NSTAT)
        esz = d_sz.numpy(n * NCAND).reshape(n, NCAND)
        pred = net.predict(d_f, n)
        choice, fb = select_formats(pred, d_sz, n, rule=rule)
        ex = esz.argmin(axis=1)
        fallbacks += int(fb.sum())
        picks += np.bincount(choice, minlength=NCAND)
        whole = ea.from_device(d_vals.ptr(), n * CHUNK)
        want.append(_masked_sum_col(whole, val, res1))
        k, mbw = column_level_choice(esz, st[:, ST_MAX_BW].max())
        pc = ea.morph(candidate_format(k, mbw), whole)
        plain.append(pc)
        foot["plain"] += pc.meta().size_used
        lc = ea.morph_blockwise(bwi, whole, format_codes(choice, st[:, ST_MAX_BW]))
        ec = lc if np.array_equal(choice, ex) else ea.morph_blockwise(bwi, whole, format_codes(ex, st[:, ST_MAX_BW]))
        learned.append(lc)
        exact.append(ec)
        foot["learned"] += lc.meta().size_used
        foot["exact"] += ec.meta().size_used
        del whole
    t_prep = time.perf_counter() - t2
    nchunks = ncols * n
    say(f"selection workload: {ncols} columns x {n} chunks prepared in {t_prep:.1f}s; footprint learned "
        f"{foot['learned']} exact {foot['exact']} plain {foot['plain']}; "
        f"{'fallbacks' if rule == 0 else 'exact smallest kept'} {fallbacks}")
    resc = ea.DeviceU64(ncols)
    slots = [resc.at(k) for k in range(ncols)]

    def check_cols(got, tag):
        if got != want:
            bad = [k for k in range(ncols) if got[k] != want[k]]
            raise RuntimeError(f"selection workload: {tag} sums differ at columns {bad[:5]}")

    def timed(fn, reps):
        for _ in range(warmup):
            fn()
        hip.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        return e0.elapsed_ms(e1) / reps

    def plain_fn(stream=None):
        for k, col in enumerate(plain):
            ea.select_sum(2, val, col, slots[k], stream)
    plain_fn()
    check_cols(resc.get().tolist(), "plain")
    ms_plain = timed(plain_fn, max(2, steps // 2))
    ms_graph = None
    if graph_leg:   # the per-column calls recorded once and replayed as one hipGraph
        st_ = hip.Stream()
        plan = ea.Plan()
        with plan.record(st_.handle):
            plain_fn(st_.handle)
        with plan.record(st_.handle):
            plain_fn(st_.handle)
        resc.zero()
        plan.launch(st_.handle)
        st_.synchronize()
        check_cols(resc.get().tolist(), "plain graph")
        ms_graph = timed(lambda: plan.launch(None), steps)
        plan.free()
    out, batches = {}, {}
    for tag, cols in (("learned", learned), ("exact", exact)):
        b = ea.Batch()
        for col in cols:
            b.enqueue(col, 2, val)
        b.run(resc)
        check_cols(resc.get().tolist(), f"batched {tag}")
        cfg = b.tune(resc, reps=3)
        b.run(resc)
        check_cols(resc.get().tolist(), f"batched {tag} tuned")
        batches[tag] = (b, cfg.as_dict())
    # the two batches timed in alternation, the median of 5 rounds each: the
    # device's run-to-run drift (a few percent) hits both alike
    rounds = {tag: [] for tag in batches}
    for _ in range(5):
        for tag, (b, _) in batches.items():
            rounds[tag].append(timed(lambda: b.run(resc), steps))
    for tag, (b, cfg) in batches.items():
        out[tag] = (float(np.median(rounds[tag])), cfg)
    values = nchunks * CHUNK
    ue = lambda ms: round(8 * values / ms / 1e6, 1)  # noqa: E731
    rep = {
        "workload": f"{ncols} uint64 columns of {n} x 2^16 values (a mixture of the format-cascade-sweep "
                    f"distributions switching every 2^16 values), select(x < {val}) + agg_sum per column",
        "layout": "each column one blockwise_individual_f<65536> column, a format code per chunk",
        "labels": labels, "train_chunks": ntrain, "epochs": epochs, "loss_first_epoch": float(losses[0]),
        "loss_last_epoch": float(losses[-1]), "train_seconds": round(t_train, 2),
        "label_seconds": round(t_cal, 1),
        "footprint_bytes": foot, "learned_over_exact": round(foot["learned"] / foot["exact"], 5),
        "rule": rule, ("exact_fallbacks" if rule == 0 else "exact_smallest_kept"): fallbacks, "chunks": nchunks,
        "learned_picks": {CAND_NAMES[k]: int(picks[k]) for k in range(NCAND)},
        "plain_ms": round(ms_plain, 3), "plain_graph_ms": round(ms_graph, 3) if ms_graph else None,
        "batched_learned_ms": round(out["learned"][0], 3), "batched_exact_ms": round(out["exact"][0], 3),
        "uncompr_equiv_GBps": {"plain": ue(ms_plain), "plain_graph": ue(ms_graph) if ms_graph else None,
                               "batched_learned": ue(out["learned"][0]), "batched_exact": ue(out["exact"][0])},
        "compressed_GBps_batched_learned": round(foot["learned"] / out["learned"][0] / 1e6, 1),
        "speedup_batched_learned_vs_plain": round(ms_plain / out["learned"][0], 2),
        "speedup_batched_learned_vs_plain_graph": round(ms_graph / out["learned"][0], 2) if ms_graph else None,
        "batch_config": out["learned"][1],
    }
    say(f"selection workload: plain {ms_plain:.2f} ms, plain graph {ms_graph} ms, batched learned "
        f"{out['learned'][0]:.3f} ms, batched exact {out['exact'][0]:.3f} ms")
    return rep
