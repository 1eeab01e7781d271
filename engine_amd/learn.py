"""Learned per-chunk format selection: label generation, training,
selection (host side of the ms_gen_chunks / ms_chunk_* / ms_mlp_* /
ms_select_formats entry points).

The reference chooses formats from calibration runs (src/calibration/
bw_prof.cpp:71-292: time and size per format and bit width) and an
exhaustive size search (exhaustive_size_format_selector.h:35-141).  Here the
engine
  1. calibrates: times the batched fused select+sum kernel on columns of
     each candidate format at every bit width (ns per 512-value unit), and
     on rle columns of several run lengths (ns per run);
  2. labels chunks of 2^16 values generated on the device: exact size of
     every candidate (closed forms from per-block bit widths) and decode
     time (the calibration table applied to the chunk's bit widths);
  3. trains the FP32 MLP on the device (24 features -> log2 size and log2
     time of the 5 candidates);
  4. selects per chunk: among candidates predicted within `size_tol` of the
     smallest prediction, the fastest predicted; exact sizes judge the pick
     and replace it by the exact smallest when it is more than `exact_tol`
     larger.
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
                                               "time_weight", "exact_tol")]


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


def gen_chunks(specs, tables, d_out, offset_chunks=0, stream=None):
    """Writes len(specs) chunks into d_out (a hip.DeviceArray of uint64)."""
    arr = (ChunkSpec * len(specs))(*[ChunkSpec(*s) for s in specs])
    tab = np.ascontiguousarray(tables.array())
    ea.check(ea.lib().ms_gen_chunks(d_out.ptr(offset_chunks * CHUNK), arr, len(specs), tab.ctypes.data,
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


def select_formats(d_pred, d_sizes, n, size_tol=0.02, time_weight=1e6, exact_tol=0.03, stream=None):
    """Per chunk: candidate index, and whether the exact-size fallback fired."""
    sp = SelectParams(SIZE_OFF, SIZE_SCALE, TIME_OFF, TIME_SCALE, size_tol, time_weight, exact_tol)
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
    sorted_means = (2, 3, 4, 10, 32, 100, 1000)
    rle_means = (1, 2, 4, 8, 16, 64, 256, 1024)
    specs = []
    for i in range(n):
        k = int(r[i, 0] % np.uint64(5))
        s = int(r[i, 1])
        q2, q3, q4 = int(r[i, 2]), int(r[i, 3]), int(r[i, 4])
        if k == GEN_UNIFORM:
            specs.append((k, q2 % 65, 0, 0, s, 0))
        elif k == GEN_DVBP:
            lo = 1 + q2 % 48
            hi = min(64, lo + q3 % 33)
            specs.append((k, lo, hi, 0, s, 0))
        elif k == GEN_SORTED:
            a = sorted_means[q2 % len(sorted_means)]
            base = q3 >> (q4 % 64)
            specs.append((k, a, 0, tables(a), s, base))
        elif k == GEN_FOR:
            bits = 1 + q2 % 40
            base = q3 >> (q4 % 64)
            specs.append((k, bits, 0, 0, s, base))
        else:
            a = rle_means[q2 % len(rle_means)]
            specs.append((k, a, 1 + q3 % 64, tables(a), s, 0))
    return specs


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


def calibrate(nchunks=16, copies=8, reps=5, stream=None, log=None):
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
