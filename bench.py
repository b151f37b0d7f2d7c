"""Benchmark: CSR SpMV GB/s and fraction of MI355X HBM peak, plus the
reduced-precision, learned-selection and batched forms.

Headline workload (one MI355X): the single large matrix -- n = 2^22 rows and
columns, row lengths Zipf(s=1.5) clipped to [1, 10^4] (about 6.4e8 nnz),
sorted distinct uniform int32 columns, N(0,1) values, FP64, alpha=1, beta=0.
One step = one compat SpMV y = alpha*A*x + beta*y over the whole matrix
through the column-panel plan with stream staging (its analysis runs once,
before timing; the L2-staged plan, the learned selector's choice among the
direct kernels and the plans, and its best direct kernel are reported beside
it), inputs resident in HBM.  `value` is
algorithmic bytes per SpMV / time:

  B = nnz*(s_v + s_i) + (rows+1)*s_o + cols*s_x + rows*s_y*(1 + [beta != 0])

Extra lines (same JSON object, "extra"): FP32 storage; SpMM with a 32-column
row-major block (plain, column-panel and ordered forms); CSR->CSC and COO->CSR on the same matrix; the precision
case (n = 2^20, 32 nnz/row, lognormal values: FP64 vs FP32/BF16/FP16
storage, rows/s and error against the FP64 result); the batched case
(10,000 matrices in one hipGraph-replayed persistent launch, SpMV/s).

Every timed output is checked after its timing (the "guard" object): the
headline and FP32 SpMVs on every row against the ordered kernel (the
reference's summation order, itself compared bit for bit with the
reference's own output on the CPU baseline's rows) within the plain-mode
bound 4*len_i*eps*(|A||x|)_i; the plain and ordered SpMM on sampled rows
against a recomputation in the reference's order (within the plain bound and
bit for bit); the conversions' structure.  A failed check sets
"error" in the line and the exit status to 3.  The roofline object also
carries the measured HBM read and copy ceilings (libasc_sparse_probe.so).

Prints one JSON line (the repository's benchmark contract).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "asc-cpp_amd"))

import asc_sparse_hip as H  # noqa: E402
from asc_sparse_hip import gen as G  # noqa: E402
from asc_sparse_hip import selector as S  # noqa: E402

METRIC = "CSR SpMV GB/s & %HBM peak (FP64 plain vs BF16/learned-select); batched SpMV/s"
STAGINGS = {"auto": H.STAGING_AUTO, "l2": H.STAGING_L2, "lds": H.STAGING_LDS,
            "stream": H.STAGING_STREAM}
STAGING_NAMES = {v: k for k, v in STAGINGS.items()}
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(rows, cols, nnz, sv, si, sx, sy, so=8, beta_nonzero=False):
    return nnz * (sv + si) + (rows + 1) * so + cols * sx + rows * sy * (2 if beta_nonzero else 1)


def time_ops(ctx, fn, steps, warmup, flush=None):
    """Per-call HIP events on the context stream (median and mean, ms) and
    the whole loop's events; `flush` (untimed) runs before every timed call."""
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    starts, stops = [], []
    first = ctx.record()
    for _ in range(steps):
        if flush is not None:
            flush()
        starts.append(ctx.record())
        fn()
        stops.append(ctx.record())
    last = ctx.record()
    ctx.synchronize()
    per = [s.elapsed_ms(e) for s, e in zip(starts, stops)]
    return {"loop_ms": first.elapsed_ms(last) / steps, "mean_ms": float(np.mean(per)),
            "median_ms": float(np.median(per))}


def measured_traffic(rows, nnz, alg):
    """HBM bytes per launch from the committed rocprofv3 counter summary of
    this exact workload and kernel (FETCH_SIZE + WRITE_SIZE, calibrated on a
    1 GiB copy), or None when no matching summary exists."""
    path = os.path.join(ROOT, "profiles", "spmv_%s_fp64_zipf_n%d.json" % (alg, rows))
    if not os.path.exists(path):
        return None, None
    prof = json.load(open(path))
    if f"{nnz} nnz" not in prof.get("workload", ""):
        return None, None
    return prof["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)


def host_facts():
    """The GPU host's CPU, for the CPU baseline line."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model,
            "allowed_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(m, x, sample_rows, reps=20):
    """Serial reference compat SpMV (oracle/_ref, the reference's own
    sources compiled here; the oracle's restatement when that is absent) on
    rows [0, sample_rows) of the same matrix, pinned to one host core.
    Returns the line and the reference's y for those rows (the headline
    guard compares the device's ordered kernel with it bit for bit)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as O
    r = min(sample_rows, m.rows)
    nnz = int(m.offsets[r])
    off = np.ascontiguousarray(m.offsets[: r + 1])
    idx = m.indices[:nnz].astype(np.int64)
    val = np.ascontiguousarray(m.values[:nnz])
    y = np.zeros(r)
    kind = "port"
    fn = lambda: O.compat_spmv(off, idx, val, 1.0, x, 0.0, y)  # noqa: E731
    if O.ref_available():
        kind = "reference"
        fn = lambda: O.ref_compat_spmv(off, idx, val, 1.0, x, 0.0, y, cols=m.cols)  # noqa: E731
    facts = host_facts()
    saved = os.sched_getaffinity(0)
    core = min(saved)
    os.sched_setaffinity(0, {core})  # this thread (the one calling the serial loop)
    try:
        fn()
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, saved)
    t = float(np.median(times))
    b = spmv_bytes(r, m.cols, nnz, 8, 8, 8, 8)  # the reference's int64 indices
    line = {"value": round(b / t / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"rows [0, {r}) of the same matrix ({nnz} nnz, int64 indices as in the "
                      f"reference), compat SpMV FP64 pinned to host core {core}, median of "
                      f"{reps} after 1 warm-up ({t * 1e3:.1f} ms per SpMV, spread "
                      f"{min(times) * 1e3:.1f}-{max(times) * 1e3:.1f} ms)",
            "host": facts}
    return line, y


def pinned_median(fn, reps):
    """Median wall time of fn() on one host core (after one warm-up)."""
    saved = os.sched_getaffinity(0)
    core = min(saved)
    os.sched_setaffinity(0, {core})
    try:
        fn()
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, saved)
    return float(np.median(times)), core


def cpu_conversions(m, sample_rows, reps=3):
    """CPU baselines of CSR->CSC, labelled by algorithm: the O(nnz + cols)
    counting-sort restatement (oracle) on rows [0, sample_rows) of the large
    matrix, and the reference's own FromCompressed (oracle/_ref, its
    O(cols * (rows + nnz)) scan) beside the restatement on the 4096 x 4096,
    16-entries-per-row matrix, where the scan finishes.  One host core each;
    int64 indices as in the reference."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as O
    out = {}
    r = min(sample_rows, m.rows)
    nnz = int(m.offsets[r])
    off = np.ascontiguousarray(m.offsets[: r + 1])
    idx = m.indices[:nnz].astype(np.int64)
    val = np.ascontiguousarray(m.values[:nnz])
    t, core = pinned_median(lambda: O.transpose(0, r, m.cols, off, idx, val), reps)
    by = (r + 1) * 8 + 2 * nnz * 16 + (m.cols + 1) * 8
    out["restatement_sample"] = {
        "algorithm": "O(nnz + cols) stable counting sort (the oracle's restatement)",
        "sample": f"rows [0, {r}) of the large matrix ({nnz} nnz) -> CSC with {m.cols} columns",
        "ms": round(t * 1e3, 1), "gbs": round(by / t / 1e9, 3),
        "entries_per_s": round(nnz / t), "core": core, "reps": reps}
    small = G.uniform16(4096)
    so, si, sv = small.offsets, small.indices.astype(np.int64), small.values
    t, _ = pinned_median(lambda: O.transpose(0, small.rows, small.cols, so, si, sv), reps)
    out["restatement_4096"] = {"algorithm": "O(nnz + cols) restatement", "nnz": int(small.nnz),
                               "ms": round(t * 1e3, 3), "entries_per_s": round(small.nnz / t)}
    if O.ref_available():
        t, _ = pinned_median(lambda: O.ref_transpose(0, small.rows, small.cols, so, si, sv), reps)
        out["reference_4096"] = {
            "algorithm": "the reference's FromCompressed, O(cols * (rows + nnz)) "
                         "(compressed.h:1151-1221, compiled from its sources)",
            "nnz": int(small.nnz), "ms": round(t * 1e3, 1), "entries_per_s": round(small.nnz / t)}
    return out


def abs_ax_rows(m, x, rows=None, parts=16):
    """(|A||x|)_i in FP64 with numpy (the guard's error-bound scale), over
    all rows or the rows listed."""
    lens = np.diff(m.offsets)
    if rows is not None:
        out = np.empty(rows.shape[0])
        for k, r in enumerate(rows):
            b, e = int(m.offsets[r]), int(m.offsets[r + 1])
            out[k] = float(np.sum(np.abs(m.values[b:e]) * np.abs(x[m.indices[b:e]])))
        return out
    out = np.zeros(m.rows)
    cuts = np.searchsorted(m.offsets, np.linspace(0, m.nnz, parts + 1)).tolist()
    cuts[0], cuts[-1] = 0, m.rows
    for r0, r1 in zip(cuts[:-1], cuts[1:]):
        if r1 <= r0:
            continue
        b, e = int(m.offsets[r0]), int(m.offsets[r1])
        prod = np.abs(m.values[b:e]) * np.abs(x[m.indices[b:e]])
        nz = np.nonzero(lens[r0:r1])[0]
        if nz.size:
            out[r0 + nz] = np.add.reduceat(prod, m.offsets[r0:r1][nz] - b)
        del prod
    return out


def bound_violations(y, yref, lens, absax, eps):
    """Rows where |y - yref| exceeds the plain-mode bound 4*len*eps*(|A||x|)_i
    (alpha = 1, beta = 0), and the largest error/bound ratio."""
    bound = 4.0 * np.maximum(lens, 1) * eps * absax
    err = np.abs(y.astype(np.float64) - yref.astype(np.float64))
    bad = ~(err <= bound)
    ratio = float(np.max(np.where(bound > 0, err / np.where(bound > 0, bound, 1.0), 0.0)))
    return int(np.count_nonzero(bad)), ratio


def learned_pick(ctx, a):
    """Runs the committed learned selector on the device matrix: returns the
    Selection (a direct kernel, or a column-panel plan and its staging) and
    the network's per-output scores."""
    if not os.path.exists(S.DEFAULT_MODEL):
        return None, None
    model = S.load_model()
    pick, scores = S.select(ctx, model, a)
    return pick, {name: round(v, 4) for name, v in zip(model.get("outputs", S.OUTPUT_NAMES), scores)}


def calibration_copy(ctx):
    """A plain 1 GiB device copy: HBM counters read it as 1 GiB fetched and
    1 GiB written, which scales the counters of the kernels profiled beside it."""
    ca, cb = ctx.alloc(1 << 30), ctx.alloc(1 << 30)
    ctx.fill(ca, 1)
    ctx.copy(cb.ptr, ca.ptr, 1 << 30)
    ctx.synchronize()
    del ca, cb


def bench_large(ctx, args, extra):
    t0 = time.perf_counter()
    m = G.zipf_large(args.rows)
    log(f"generated n={m.rows} nnz={m.nnz} in {time.perf_counter() - t0:.1f}s")
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    x = G.vector(m.cols, 12345)
    xb, yb = ctx.upload(x), ctx.alloc(8 * m.rows)
    xd, yd = H.vector(xb, m.cols, H.F64, writable=False), H.vector(yb, m.rows, H.F64)
    ws = H.Workspace(ctx)
    if args.calibrate:
        calibration_copy(ctx)
    nbytes = spmv_bytes(m.rows, m.cols, m.nnz, 8, 4, 8, 8)
    lens = np.diff(m.offsets)
    if args.alg in ("auto", "panel"):
        # What the learned selector picks: its choice, and the direct kernel
        # the plan cache holds meanwhile, timed ...
        pick, scores = learned_pick(ctx, a)
        direct = pick.direct_algorithm if pick is not None else H.ALG_MERGE
        rd = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=direct, ws=ws), 5, 2)
        chosen = (f"panel_{STAGING_NAMES[pick.staging]}" if pick is not None and
                  pick.algorithm == H.ALG_PANEL else H.ALG_NAMES[direct])
        extra["learned_select"] = {"choice": chosen, "scores": scores,
                                   "fell_back": bool(pick.fell_back) if pick is not None else None,
                                   "direct_algorithm": H.ALG_NAMES[direct],
                                   "direct_gbs": round(nbytes / rd["median_ms"] / 1e6, 1),
                                   "direct_ms": round(rd["median_ms"], 3)}
        log(f"learned pick {chosen} (direct {H.ALG_NAMES[direct]}: {rd['median_ms']:.3f} ms)")
        # ... the column-panel plan with the other staging, for comparison ...
        staging = STAGINGS[args.staging]
        other = H.STAGING_STREAM if staging != H.STAGING_STREAM else H.STAGING_AUTO
        H.csr_spmv_analyze(ctx, a, H.panel_options(staging=other))
        ro = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=H.ALG_PANEL, ws=ws), 5, 2)
        extra["panel_other_staging"] = {"staging": STAGING_NAMES[other],
                                        "gbs": round(nbytes / ro["median_ms"] / 1e6, 1),
                                        "ms": round(ro["median_ms"], 3)}
        log(f"panel plan, {STAGING_NAMES[other]} staging: {ro['median_ms']:.3f} ms")
        # ... and the column-panel plan with the chosen staging, which the
        # timed steps run (analysis once, outside the timed region).
        t0 = time.perf_counter()
        H.csr_spmv_analyze(ctx, a, H.panel_options(staging=staging))
        ctx.synchronize()
        # Default listed-row minimum: 1 for the stream form (every row is
        # streamed), 128 for the L2/LDS forms (shorter rows run beside them).
        lmin = 1 if staging == H.STAGING_STREAM else 128
        extra["panel_analysis"] = {"ms": round((time.perf_counter() - t0) * 1e3, 2),
                                   "staging": STAGING_NAMES[staging],
                                   "long_row_min": lmin,
                                   "long_rows": int(np.sum(lens >= lmin)),
                                   "long_row_nnz_share": round(float(lens[lens >= lmin].sum()) / m.nnz, 4),
                                   "analysis_bytes": int(a.analysis.nbytes)}
        alg = H.ALG_PANEL
    else:
        alg = {"merge": H.ALG_MERGE, "ordered": H.ALG_ORDERED,
               **{f"vec{g}": v for g, v in H.ALG_VECTOR.items()}}[args.alg]
    name = H.ALG_NAMES[alg]
    res = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=alg, ws=ws),
                   args.steps, args.warmup)
    # Enqueue + wait on the host (the reference's second timing figure):
    # call, then synchronize, per SpMV.
    host = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=alg, ws=ws)
        ctx.synchronize()
        host.append(time.perf_counter() - t0)
    res["host_enqueue_wait_ms"] = float(np.median(host)) * 1e3
    log(f"FP64 {name}: {res['loop_ms']:.3f} ms/step, {nbytes / res['loop_ms'] / 1e6:.0f} GB/s")
    # Guard (after timing, outside the timed region): every row of the timed
    # SpMV's output against the ordered kernel's (the reference's summation,
    # bit for bit; checked against the reference's own output on the CPU
    # baseline's rows below) within the plain-mode bound.
    y_head = ctx.download(yb, np.float64, m.rows)
    H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=H.ALG_ORDERED, ws=ws)
    y_ord = ctx.download(yb, np.float64, m.rows)
    t0 = time.perf_counter()
    absax = abs_ax_rows(m, x)
    nbad, ratio = bound_violations(y_head, y_ord, lens, absax, 2.0 ** -53)
    guard = {"headline": {"rows_checked": m.rows, "rows_over_bound": nbad,
                          "max_err_over_bound": round(ratio, 4),
                          "against": "ordered kernel (reference summation order)"}}
    log(f"guard: headline {nbad} rows over the bound (max err/bound {ratio:.3f}, "
        f"{time.perf_counter() - t0:.1f}s)")
    del y_head
    if args.sweep or not args.quick:
        sweep = {}
        for v in H.TUNABLE_ALGS:
            r = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=v, ws=ws), 3, 1)
            sweep[H.ALG_NAMES[v]] = round(nbytes / r["median_ms"] / 1e6, 1)
        extra["fp64_variant_gbs"] = sweep
        log(f"variants: {sweep}")
    if not args.quick:
        a32 = H.convert_values(ctx, a, H.F32)
        xb32, yb32 = ctx.upload(x.astype(np.float32)), ctx.alloc(4 * m.rows)
        x32d, y32d = H.vector(xb32, m.cols, H.F32, writable=False), H.vector(yb32, m.rows, H.F32)
        if alg == H.ALG_PANEL:
            H.csr_spmv_analyze(ctx, a32, H.panel_options(staging=STAGINGS[args.staging]))
        r32 = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a32, x32d, 0.0, y32d, alg=alg, ws=ws), 5, 2)
        b32 = spmv_bytes(m.rows, m.cols, m.nnz, 4, 4, 4, 4)
        extra["fp32_same_matrix"] = {"gbs": round(b32 / r32["median_ms"] / 1e6, 1),
                                     "ms": round(r32["median_ms"], 3), "algorithm": name}
        log(f"FP32 {name}: {r32['median_ms']:.3f} ms")
        y32 = ctx.download(yb32, np.float32, m.rows)
        H.csr_spmv(ctx, 1.0, a32, x32d, 0.0, y32d, alg=H.ALG_ORDERED, ws=ws)
        y32o = ctx.download(yb32, np.float32, m.rows)
        # |A32||x32| <= |A||x| (1 + 2^-24)^2 for the FP32-rounded operands.
        nbad, ratio = bound_violations(y32, y32o, lens, absax * (1 + 2.0 ** -22), 2.0 ** -24)
        guard["fp32"] = {"rows_checked": m.rows, "rows_over_bound": nbad,
                         "max_err_over_bound": round(ratio, 4)}
        del a32, xb32, yb32, y32, y32o
        bench_spmm(ctx, m, a, lens, extra, guard)
        bench_conversions(ctx, m, a, extra, guard)
        bench_shuffled_finalize(ctx, m, extra, guard)
    return m, alg, res, nbytes, y_ord, guard, x


# SpMM panel form: a stream-staged column-panel analysis with panels of 16384
# columns (B's panel slice, 16384 rows x 256 bytes = 4 MiB, one XCD's L2) and
# rows of at least 4096 entries listed (68% of the Zipf matrix's entries);
# measured on that matrix at k = 32: 8192/16384-column panels 20.3/17.5 ms,
# listed-row minimum 2048/4096 with 16384 columns 18.3/17.5 ms.
SPMM_PANEL_COLUMNS = 16384
SPMM_PANEL_LONG_MIN = 4096


def bench_spmm(ctx, m, a, lens, extra, guard, steps=3):
    """SpMM C += A*B on the large matrix, B row-major 2^22 x 32 (U[-1,1]): the
    plain, panel and ordered forms timed, sampled rows checked, and the B-row
    gather ceiling measured beside them."""
    k = 32
    bmat = ctx.alloc(8 * m.cols * k)
    bh = G.vector(m.cols * k, 777)
    ctx.copy(bmat.ptr, bh.ctypes.data, bh.nbytes)
    cmat = ctx.alloc(8 * m.rows * k)
    ctx.fill(cmat, 0)
    ctx.synchronize()
    bd = H.dense(bmat, m.cols, k, k, H.ROW_MAJOR, H.F64, writable=False)
    cd = H.dense(cmat, m.rows, k, k, H.ROW_MAJOR, H.F64)
    sws = H.Workspace(ctx)
    bs = m.nnz * 12 + (m.rows + 1) * 8 + m.cols * k * 8 + 2 * m.rows * k * 8
    spmm_ms = {}
    # Guard: sampled rows (the longest among them) recomputed in the
    # reference's order, C(r,:) += (1*v)*B(c,:) entry by entry, after one
    # more call of each form from zero: the ordered form bit for bit, the
    # plain and panel forms within 4*(len+1)*eps*(|A||B|)(r,:).
    bhost = bh.reshape(m.cols, k)
    pick = np.concatenate([np.random.default_rng(5).integers(0, m.rows, 256),
                           np.argsort(lens)[-8:]])
    serial, magnitude = {}, {}
    for r in pick:
        acc, mag = np.zeros(k), np.zeros(k)
        for p in range(int(m.offsets[r]), int(m.offsets[r + 1])):
            acc = acc + (1.0 * m.values[p]) * bhost[m.indices[p]]
            mag = mag + abs(m.values[p]) * np.abs(bhost[m.indices[p]])
        serial[r], magnitude[r] = acc, mag
    spmm_guard = {"rows_checked": int(pick.size)}

    def run_form(form, alg):
        rs = time_ops(ctx, lambda: H.standard_spmm(ctx, H.OP_N, 1.0, a, bd, cd, sws, alg=alg), 3, 1)
        spmm_ms[form] = rs["median_ms"]
        log(f"SpMM k=32 {form}: {rs['median_ms']:.2f} ms")
        ctx.fill(cmat, 0)
        H.standard_spmm(ctx, H.OP_N, 1.0, a, bd, cd, sws, alg=alg)
        cgot = ctx.download(cmat, np.float64, m.rows * k).reshape(m.rows, k)
        if form.startswith("ordered"):
            spmm_guard[f"{form}_rows_bit_exact"] = sum(
                int(np.array_equal(serial[r].view(np.uint64), cgot[r].view(np.uint64)))
                for r in pick)
        else:
            spmm_guard[f"{form}_rows_within_bound"] = sum(
                int(np.all(np.abs(cgot[r] - serial[r]) <=
                           4 * (lens[r] + 1) * 2.0 ** -53 * magnitude[r])) for r in pick)
        return cgot[pick]

    # The plain and ordered forms with no analysis of the matrix (AUTO would
    # take the panel form on an analysis listing only long rows) ...
    H.plan_clear(ctx)
    if getattr(a, "analysis", None) is not None:
        a.analysis.free()
        a.analysis = None
    run_form("plain", H.ALG_AUTO)
    run_form("ordered", H.ALG_ORDERED)
    # ... then the panel form's analysis (once, outside the timed region) and
    # the panel form.
    t0 = time.perf_counter()
    H.csr_spmv_analyze(ctx, a, H.panel_options(panel_columns=SPMM_PANEL_COLUMNS,
                                                long_row_min=SPMM_PANEL_LONG_MIN,
                                                staging=H.STAGING_STREAM))
    ctx.synchronize()
    panel_analysis_ms = (time.perf_counter() - t0) * 1e3
    got_panel = run_form("panel", H.ALG_PANEL)
    # The ordered form on the same analysis: the panels in order, the same bits.
    run_form("ordered_panel", H.ALG_ORDERED)
    # AUTO now takes the panel form: the same bits on the sampled rows.
    ctx.fill(cmat, 0)
    H.standard_spmm(ctx, H.OP_N, 1.0, a, bd, cd, sws, alg=H.ALG_AUTO)
    got_auto = ctx.download(cmat, np.float64, m.rows * k).reshape(m.rows, k)[pick]
    spmm_guard["auto_takes_panel_form"] = bool(np.array_equal(got_auto.view(np.uint64),
                                                               got_panel.view(np.uint64)))
    gathered = m.nnz * k * 8  # every stored entry gathers one 256-byte row of B
    ceiling = bench_row_gather_ceiling(ctx)
    extra["spmm_k32_row_major"] = {
        "gbs": round(bs / spmm_ms["plain"] / 1e6, 1), "ms": round(spmm_ms["plain"], 3),
        "form": "plain (FMA, 256 MiB column passes)",
        "panel_ms": round(spmm_ms["panel"], 3),
        "panel_gbs": round(bs / spmm_ms["panel"] / 1e6, 1),
        "panel_b_row_gather_gbs": round(gathered / spmm_ms["panel"] / 1e6, 1),
        "panel_form": {"panel_columns": SPMM_PANEL_COLUMNS, "long_row_min": SPMM_PANEL_LONG_MIN,
                       "listed_rows": int(np.sum(lens >= SPMM_PANEL_LONG_MIN)),
                       "listed_nnz_share": round(float(lens[lens >= SPMM_PANEL_LONG_MIN].sum())
                                                 / m.nnz, 4),
                       "analysis_ms": round(panel_analysis_ms, 1)},
        "ordered_gbs": round(bs / spmm_ms["ordered"] / 1e6, 1),
        "ordered_ms": round(spmm_ms["ordered"], 3),
        "ordered_panel_ms": round(spmm_ms["ordered_panel"], 3),
        "b_row_gathers_gb": round(gathered / 1e9, 1),
        "b_row_gather_gbs": round(gathered / spmm_ms["plain"] / 1e6, 1),
        "b_row_gather_ceiling": ceiling,
        "frac_of_gather_ceiling_256MiB": round(
            gathered / spmm_ms["plain"] / 1e6 / ceiling["table_256MiB_gbs"], 3)}
    guard["spmm_k32"] = spmm_guard
    H.plan_clear(ctx)
    a.analysis.free()
    a.analysis = None
    del bmat, cmat, bh, bhost


def bench_conversions(ctx, m, a, extra, guard, steps=2):
    """CSR->CSC and COO->CSR on the large matrix, structure checked.  Each
    timed call writes into the output buffers of the warm-up call (the C
    ABI's caller-provided outputs: no allocation inside the timed region)."""
    cws = H.Workspace(ctx)
    holder = {}

    def to_csc():
        holder["csc"] = H.convert_compressed(ctx, a, cws, into=holder.get("csc"))
    rc = time_ops(ctx, to_csc, steps, 1)
    bc = 2 * m.nnz * 12 + (m.rows + 1) * 8 + (m.cols + 1) * 8
    extra["csr_to_csc"] = {"gbs": round(bc / rc["median_ms"] / 1e6, 1), "ms": round(rc["median_ms"], 3)}
    # Guard: the column counts (the full transpose is checked bit for bit
    # by the GPU tests at this size).
    off_c = ctx.download(H._ptr_buf(ctx, holder["csc"].desc.outer_offsets), np.int64, m.cols + 1)
    guard["csr_to_csc"] = {"column_counts_match": bool(
        off_c[0] == 0 and np.array_equal(np.diff(off_c), np.bincount(m.indices, minlength=m.cols)))}
    holder.clear()
    coo = H.convert_to_coordinate(ctx, a)

    def to_csr():
        holder["csr"] = H.convert_coordinate(ctx, coo, H.CSR, 4, ws=cws, into=holder.get("csr"))
    rr = time_ops(ctx, to_csr, steps + 1, 1)
    bcoo = m.nnz * (16 + 8) + m.nnz * (4 + 8) + (m.rows + 1) * 8
    extra["coo_to_csr"] = {"gbs": round(bcoo / rr["median_ms"] / 1e6, 1), "ms": round(rr["median_ms"], 3)}
    log(f"CSR->CSC {rc['median_ms']:.2f} ms, COO->CSR {rr['median_ms']:.2f} ms")
    off_r = ctx.download(H._ptr_buf(ctx, holder["csr"].desc.outer_offsets), np.int64, m.rows + 1)
    guard["coo_to_csr"] = {"offsets_match": bool(np.array_equal(off_r, m.offsets))}
    holder.clear()
    del coo, cws


def bench_shuffled_finalize(ctx, m, extra, guard):
    """COO->CSR through the sort path: the same matrix's entries as raw
    triplets in a scrambled order (entry i placed at (a*i + b) mod nnz, a
    coprime with nnz, which moves every entry far from its neighbours),
    finalized on the device (stable sort by (row, col), duplicates summed,
    zeros kept).  Algorithmic bytes: triplets read (16 + 8 per entry) and the
    canonical result written (16 + 8 per entry)."""
    import math
    n = m.nnz
    t0 = time.perf_counter()
    a_mul = int(n * 0.6180339887)
    while math.gcd(a_mul, n) != 1:
        a_mul += 1
    pos = (np.arange(n, dtype=np.int64) * a_mul + 12345) % n
    src = np.empty(n, dtype=np.int64)
    src[pos] = np.arange(n, dtype=np.int64)  # the entry placed at each position
    del pos
    coords = np.empty((n, 2), dtype=np.int64)
    coords[:, 0] = np.repeat(np.arange(m.rows, dtype=np.int32), np.diff(m.offsets))[src]
    coords[:, 1] = m.indices[src]
    vals = m.values[src]
    del src
    t = H.triplets(ctx, m.rows, m.cols, coords, vals, H.F64)
    del coords, vals
    log(f"shuffled COO of {n} entries built in {time.perf_counter() - t0:.1f}s")
    ws = H.Workspace(ctx)
    holder = {}

    def fin():
        holder["out"] = H.coo_finalize(ctx, t, H.DUPLICATE_SUM, H.ZERO_KEEP, ws)
    res = time_ops(ctx, fin, 3, 1)
    kept = int(holder["out"].desc.nnz)
    by = n * 24 + kept * 24
    extra["coo_finalize_shuffled_same_matrix"] = {
        "entries": n, "unique": kept, "ms": round(res["median_ms"], 3),
        "gbs": round(by / res["median_ms"] / 1e6, 1),
        "entries_per_s": round(n / (res["median_ms"] * 1e-3))}
    csr = H.convert_coordinate(ctx, holder["out"], H.CSR, 4)
    off_r = ctx.download(H._ptr_buf(ctx, csr.desc.outer_offsets), np.int64, m.rows + 1)
    guard["coo_finalize_shuffled"] = {"unique_equals_nnz": kept == n,
                                      "offsets_match": bool(np.array_equal(off_r, m.offsets))}
    log(f"COO finalize of the shuffled matrix: {res['median_ms']:.2f} ms")
    del holder, csr, t, ws


def bench_precision(ctx, extra):
    """n = 2^20, 32 nnz/row: FP64 vs FP32/BF16/FP16 storage (FP32 x, y, FP32
    accumulation plain or compensated), each through its fastest form (direct
    kernels and, for plain accumulation, the column-panel plans; analysis
    outside the timed region, as for the headline).  MALL is flushed before
    each timed call (the BF16 working set, 222 MB, would otherwise stay
    cached).  Every result's error against the ordered FP64 SpMV is reported."""
    m = G.precision_matrix(1 << 20)
    n = m.rows
    a64 = H.clone_compressed(ctx, H.CSR, n, n, m.offsets, m.indices, m.values, H.F64, 4)
    aabs = H.clone_compressed(ctx, H.CSR, n, n, m.offsets, m.indices, np.abs(m.values), H.F64, 4)
    x64 = G.vector(n, 4242)
    ws = H.Workspace(ctx)
    flush = ctx.alloc(1 << 30)
    xb64, xabs = ctx.upload(x64), ctx.upload(np.abs(x64))
    yref, yabs = ctx.alloc(8 * n), ctx.alloc(8 * n)
    xd64 = H.vector(xb64, n, H.F64, writable=False)
    # FP64 reference on the GPU in ordered mode (bit-exact with the serial reference).
    H.csr_spmv(ctx, 1.0, a64, xd64, 0.0, H.vector(yref, n, H.F64), alg=H.ALG_ORDERED, ws=ws)
    H.csr_spmv(ctx, 1.0, aabs, H.vector(xabs, n, H.F64, writable=False), 0.0,
               H.vector(yabs, n, H.F64), alg=H.ALG_ORDERED, ws=ws)
    y64 = ctx.download(yref, np.float64, n)
    norm = float(np.max(ctx.download(yabs, np.float64, n)))
    out = {}
    yb64 = ctx.alloc(8 * n)
    yd64 = H.vector(yb64, n, H.F64)

    # Column-panel plans at their defaults, and the L2/LDS plans with every
    # row listed (the default lists rows of at least 128 entries, so on this
    # 32-entry-per-row matrix it leaves every row to the direct kernel).
    plans = ((H.STAGING_L2, 0, "panel_l2"), (H.STAGING_L2, 1, "panel_l2_all_rows"),
             (H.STAGING_LDS, 0, "panel_lds"), (H.STAGING_LDS, 1, "panel_lds_all_rows"),
             (H.STAGING_STREAM, 0, "panel_stream"))

    def timed(a, xd, yd, alg, scale=1.0, acc=H.ACC_PLAIN, steps=3, warmup=1):
        return time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0 / scale, a, xd, 0.0, yd, alg=alg, acc=acc,
                                                ws=ws),
                        steps, warmup, flush=lambda: ctx.fill(flush, 0))["median_ms"]

    def cold_pick(a, xd, yd, scale=1.0, acc=H.ACC_PLAIN):
        """The fastest form under the timed conditions (MALL flushed before
        each call): every direct variant and, for plain accumulation, the
        three column-panel plans (each analysed once, untimed).  Leaves the
        winning plan analysed; returns (algorithm, name)."""
        best, best_ms = (H.ALG_MERGE, "merge"), None
        for alg in H.TUNABLE_ALGS:
            t = timed(a, xd, yd, alg, scale, acc)
            if best_ms is None or t < best_ms:
                best, best_ms = (alg, H.ALG_NAMES[alg]), t
        if acc != H.ACC_PLAIN:
            return best
        best_plan = None
        for staging, lmin, name in plans:
            try:
                H.csr_spmv_analyze(ctx, a, H.panel_options(staging=staging, long_row_min=lmin))
            except H.AscHipError:
                continue
            t = timed(a, xd, yd, H.ALG_PANEL, scale)
            if t < best_ms:
                best, best_ms, best_plan = (H.ALG_PANEL, name), t, (staging, lmin)
        if best_plan is not None:
            H.csr_spmv_analyze(ctx, a, H.panel_options(staging=best_plan[0], long_row_min=best_plan[1]))
        return best

    def spmv_gbs(sv, sx, ms):
        by = spmv_bytes(n, n, m.nnz, sv, 4, sx, sx)
        return round(by / ms / 1e6, 1), round(by / ms / 1e6 / HBM_PEAK_GBS, 4)

    best64 = cold_pick(a64, xd64, yd64)
    r = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a64, xd64, 0.0, yd64, alg=best64[0], ws=ws),
                 7, 2, flush=lambda: ctx.fill(flush, 0))
    fp64_rows = n / (r["median_ms"] * 1e-3)
    y = ctx.download(yb64, np.float64, n)
    gbs, frac = spmv_gbs(8, 8, r["median_ms"])
    out["fp64"] = {"rows_per_s": fp64_rows, "ms": round(r["median_ms"], 4), "algorithm": best64[1],
                   "gbs": gbs, "frac_of_hbm_peak": frac,
                   "rel_error_vs_ordered_fp64": float(np.max(np.abs(y - y64)) / norm)}
    if best64[0] == H.ALG_PANEL:
        a64.analysis.free()
        a64.analysis = None
        H.plan_clear(ctx)
    xb32 = ctx.upload(x64.astype(np.float32))
    x32d = H.vector(xb32, n, H.F32, writable=False)
    vmax = float(np.max(np.abs(m.values)))
    for kind, kname in ((H.F32, "fp32"), (H.BF16, "bf16"), (H.F16, "fp16")):
        scale = 1.0
        if kind == H.F16:
            while vmax * scale > 32768.0:
                scale *= 0.5
        ak = H.convert_values(ctx, a64, kind, scale)
        yb = ctx.alloc(4 * n)
        yd = H.vector(yb, n, H.F32)
        for acc, aname in ((H.ACC_PLAIN, "plain"), (H.ACC_TWOSUM, "twosum")):
            best = cold_pick(ak, x32d, yd, scale, acc)
            rr = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0 / scale, ak, x32d, 0.0, yd, alg=best[0],
                                                  acc=acc, ws=ws), 7, 2,
                          flush=lambda: ctx.fill(flush, 0))
            y = ctx.download(yb, np.float32, n).astype(np.float64)
            err = float(np.max(np.abs(y - y64)) / norm)
            rows_s = n / (rr["median_ms"] * 1e-3)
            gbs, frac = spmv_gbs(H.ELEM_BYTES[kind], 4, rr["median_ms"])
            out[f"{kname}_{aname}"] = {"rows_per_s": rows_s, "ms": round(rr["median_ms"], 4),
                                       "speedup_vs_fp64": round(rows_s / fp64_rows, 3),
                                       "rel_error_vs_fp64": err, "algorithm": best[1],
                                       "gbs": gbs, "frac_of_hbm_peak": frac,
                                       "value_scale": scale}
            if best[0] == H.ALG_PANEL:
                ak.analysis.free()
                ak.analysis = None
                H.plan_clear(ctx)
        del ak, yb
    log("precision: " + ", ".join(f"{k} {v.get('speedup_vs_fp64', 1.0)}x err "
                                  f"{v.get('rel_error_vs_fp64', 0):.2e}" for k, v in out.items()))
    extra["precision_n2^20_32nnz"] = out


def bench_finalize(ctx, extra, n=1 << 26, dim=1 << 22):
    """COO finalize (CoordinateBuilder::Finalize on the device): n raw
    triplets over a dim x dim shape in random order, 1/8 of them repeating an
    earlier coordinate, FP64, duplicates summed, zeros kept.  Algorithmic
    bytes: triplets read (16 + 8 per entry) and the canonical result written
    (16 + 8 per surviving entry)."""
    rng = np.random.default_rng(0x5EEDA5C5)
    r = rng.integers(0, dim, n)
    c = rng.integers(0, dim, n)
    rep = rng.random(n) < 0.125
    src = rng.integers(0, n, n)
    r = np.where(rep, r[src], r)
    c = np.where(rep, c[src], c)
    coords = np.stack([r, c], 1).reshape(-1)
    vals = rng.standard_normal(n)
    del r, c, rep, src
    t = H.triplets(ctx, dim, dim, coords, vals, H.F64)
    del coords, vals
    ws = H.Workspace(ctx)
    holder = {}

    def fin():
        holder["out"] = H.coo_finalize(ctx, t, H.DUPLICATE_SUM, H.ZERO_KEEP, ws)
    res = time_ops(ctx, fin, 3, 1)
    kept = int(holder["out"].desc.nnz)
    by = n * 24 + kept * 24
    extra["coo_finalize"] = {"entries": n, "unique": kept, "ms": round(res["median_ms"], 3),
                             "gbs": round(by / res["median_ms"] / 1e6, 1),
                             "entries_per_s": round(n / (res["median_ms"] * 1e-3))}
    log(f"COO finalize of {n} triplets ({kept} unique): {res['median_ms']:.2f} ms")


def guard_failures(guard):
    """Names of the output checks that failed."""
    bad = []
    for k, v in guard.items():
        if v.get("rows_over_bound", 0) != 0:
            bad.append(k)
        for count in ("rows_bit_exact", "ordered_rows_bit_exact", "ordered_panel_rows_bit_exact",
                      "plain_rows_within_bound", "panel_rows_within_bound"):
            if count in v and v[count] != v["rows_checked"]:
                bad.append(k)
        if any(val is False for val in v.values()):
            bad.append(k)
    return sorted(set(bad))


def bench_hbm_ceiling(ctx, nbytes=4 << 30):
    """Measured HBM ceilings: a plain nontemporal 16-byte read of 4 GiB and a
    copy of 2 GiB (read + written bytes), from libasc_sparse_probe.so."""
    from asc_sparse_hip import probe as P
    src = ctx.alloc(nbytes)
    ctx.fill(src, 1)
    part = ctx.alloc(4 * P.STREAM_PROBE_BLOCKS)
    r = time_ops(ctx, lambda: P.stream_read(ctx, src, nbytes, part), 10, 3)
    half = nbytes // 2
    dst = ctx.alloc(half)
    c = time_ops(ctx, lambda: P.stream_copy(ctx, src, dst, half), 10, 3)
    out = {"read_gbs": round(nbytes / r["median_ms"] / 1e6, 1),
           "copy_gbs": round(2 * half / c["median_ms"] / 1e6, 1),
           "source": "libasc_sparse_probe.so stream probes, 4 GiB read / 2 GiB copy"}
    log(f"HBM ceiling: read {out['read_gbs']} GB/s, copy {out['copy_gbs']} GB/s")
    del src, dst, part
    return out


def bench_gather_ceiling(ctx, extra, bytes_per_entry, n=1 << 26):
    """The random-gather ceiling behind the SpMV roofline: n streamed
    (int32 index, FP64 value) pairs, each gathering x at a uniformly random
    position, for x of 256 KiB (L2-resident), 2 MiB (one column panel) and
    32 MiB (the headline's x, MALL-resident).  `spmv_ceiling_gbs` is that
    rate times the headline's algorithmic bytes per nonzero."""
    from asc_sparse_hip import probe as P
    rng = np.random.default_rng(G.SEED_BASE + 13)
    bv = ctx.upload(np.ones(n))
    out = ctx.alloc(8 * P.GATHER_PROBE_BLOCKS)
    res = {}
    for xn, tag in ((1 << 15, "x_256KiB_l2"), (1 << 18, "x_2MiB_panel"), (1 << 22, "x_32MiB_mall")):
        bi = ctx.upload(rng.integers(0, xn, n).astype(np.int32))
        bx = ctx.upload(G.vector(xn, 3))
        r = time_ops(ctx, lambda: P.gather_probe(ctx, bi, bv, n, bx, xn, out), 5, 2)
        rate = n / (r["median_ms"] * 1e-3)
        res[tag] = {"ggathers_per_s": round(rate / 1e9, 1),
                    "spmv_ceiling_gbs": round(rate * bytes_per_entry / 1e9, 1)}
        del bi, bx
    extra["gather_ceiling"] = res
    log("gather ceiling: " + ", ".join(f"{k} {v['ggathers_per_s']} G/s" for k, v in res.items()))


def bench_row_gather_ceiling(ctx, n=1 << 26):
    """The B-row gather ceiling behind the SpMM roofline: n random 256-byte
    row gathers (32 FP64, an 8-lane group per row as the plain SpMM kernel
    gathers) from tables of 2 MiB (L2-resident), 256 MiB (one column pass of
    the plain SpMM, MALL-sized) and 1 GiB (the whole 2^22 x 32 B), from
    libasc_sparse_probe.so; rates in gathered bytes per second."""
    from asc_sparse_hip import probe as P
    rng = np.random.default_rng(G.SEED_BASE + 14)
    out = ctx.alloc(8 * P.ROW_GATHER_PROBE_BLOCKS)
    res = {}
    for mib in (2, 256, 1024):
        rows = (mib << 20) // 256
        bi = ctx.upload(rng.integers(0, rows, n).astype(np.int32))
        bt = ctx.alloc(mib << 20)
        ctx.fill(bt, 0)
        r = time_ops(ctx, lambda: P.row_gather_probe(ctx, bi, n, bt, rows, out), 5, 2)
        res[f"table_{mib}MiB_gbs"] = round(n * 256 / r["median_ms"] / 1e6, 1)
        del bi, bt
    res["source"] = "libasc_sparse_probe.so row gather probe, 2^26 random rows per table"
    log("B-row gather ceiling: " + ", ".join(f"{k} {v}" for k, v in res.items() if k != "source"))
    return res


def bench_dense_rows(ctx, extra, rows=1 << 14, cols=1 << 22, per_row=20_000):
    """Column-panel SpMV on dense rows: rows x cols with per_row uniformly
    random columns each (FP64, int32 columns), where the default analysis
    stages 160 KiB x panels in LDS; the direct vec64 kernel, the L2-staged
    plan and the stream-staged plan are timed beside it.  Same algorithmic
    bytes as the headline."""
    m = G.make(G.UNIFORM_FIXED, rows, cols, seed=G.SEED_BASE + 9, values=G.VALUES_NORMAL,
               per_row=per_row)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    nnz = m.nnz
    del m
    xb, yb = ctx.upload(G.vector(cols, 5)), ctx.alloc(8 * rows)
    xd, yd = H.vector(xb, cols, H.F64, writable=False), H.vector(yb, rows, H.F64)
    ws = H.Workspace(ctx)
    by = spmv_bytes(rows, cols, nnz, 8, 4, 8, 8)
    out = {"rows": rows, "cols": cols, "nnz": nnz}
    r = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=H.ALG_VECTOR[64], ws=ws), 5, 2)
    out["direct_vec64_gbs"] = round(by / r["median_ms"] / 1e6, 1)
    for tag, opts in (("panel_l2_gbs", H.panel_options(staging=H.STAGING_L2)),
                      ("panel_auto_lds_gbs", None),
                      ("panel_stream_gbs", H.panel_options(staging=H.STAGING_STREAM))):
        H.csr_spmv_analyze(ctx, a, opts)
        r = time_ops(ctx, lambda: H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=H.ALG_PANEL, ws=ws), 10, 2)
        out[tag] = round(by / r["median_ms"] / 1e6, 1)
        out[tag.replace("_gbs", "_ms")] = round(r["median_ms"], 3)
    H.plan_clear(ctx)
    extra["dense_rows_2^14x2^22_20k"] = out
    log(f"dense rows: direct {out['direct_vec64_gbs']} GB/s, L2 panels {out['panel_l2_gbs']}, "
        f"LDS panels {out['panel_auto_lds_gbs']}, stream panels {out['panel_stream_gbs']} GB/s")


def bench_triangular(ctx, extra, n=1 << 18, per_row=8, band=2000):
    """Level-scheduled triangular solve (FP64, one right-hand side, bit-exact
    with the serial solver): a lower-triangular n x n matrix whose row i has
    a nonzero diagonal and up to per_row entries at distinct offsets spread
    over the band below it (offsets k*band/per_row + U[1, band/per_row]).
    Reports the analysis and the solve."""
    rng = np.random.default_rng(G.SEED_BASE + 11)
    step = band // per_row
    d = np.arange(per_row) * step + rng.integers(1, step + 1, (n, per_row))
    i = np.arange(n)[:, None]
    cols = i - d[:, ::-1]  # ascending columns
    keep = cols >= 0
    counts = keep.sum(1) + 1
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    full = np.concatenate([np.where(keep, cols, -1), np.arange(n)[:, None]], 1)
    indices = full[full >= 0].astype(np.int64)
    vals = np.where(full >= 0, np.concatenate([rng.uniform(-0.5, 0.5, (n, per_row)),
                                               rng.uniform(1.0, 2.0, (n, 1))], 1), np.nan)
    values = vals[full >= 0]
    a = H.clone_compressed(ctx, H.CSR, n, n, offsets, indices, values, H.F64)
    t0 = time.perf_counter()
    plan = H.triangular_analyze(ctx, H.OP_N, H.LOWER, H.NON_UNIT, a)
    analysis_ms = (time.perf_counter() - t0) * 1e3
    b = G.vector(n, 12)
    bb = ctx.upload(b)
    rhs = H.dense(bb, n, 1, 1, H.ROW_MAJOR, H.F64)

    def solve():
        ctx.copy(bb.ptr, b.ctypes.data, b.nbytes)
        H.triangular_solve(ctx, H.OP_N, 1.0, a, rhs, plan)
    r = time_ops(ctx, solve, 5, 1)
    extra["triangular_lower_band"] = {
        "rows": n, "nnz": int(offsets[-1]), "levels": int(plan.info.levels),
        "widest_level": int(plan.info.max_level_rows), "analysis_ms": round(analysis_ms, 2),
        "solve_ms_incl_rhs_copy": round(r["median_ms"], 3),
        "rows_per_s": round(n / (r["median_ms"] * 1e-3))}
    log(f"triangular: {n} rows, {plan.info.levels} levels, analysis {analysis_ms:.1f} ms, "
        f"solve {r['median_ms']:.3f} ms")


def bench_batched(ctx, extra, count):
    t0 = time.perf_counter()
    specs = G.batched_specs(count)
    mats = [G.make_spec(s) for s in specs]
    p = G.pack(mats)
    del mats
    n = int(p.rows.sum())
    nnz = int(p.nnz.sum())
    log(f"batched: {count} matrices, {n} rows, {nnz} nnz generated in {time.perf_counter() - t0:.1f}s")
    res = {}
    for kind, kname in ((H.F64, "fp64"), (H.F32, "fp32")):
        entries, bufs = H.upload_packed(ctx, p, G.vector(n, 1), np.zeros(n), kind, 1.0, 0.0)
        b = H.Batch(ctx, entries, keep=bufs)
        lanes, _ = b.tune(reps=2)
        b.capture()
        r = time_ops(ctx, b.replay, 10, 2)
        es = 8 if kind == H.F64 else 4
        by = nnz * (es + 4) + (n + count) * 8 + n * es * 2
        res[kname] = {"spmv_per_s": count / (r["median_ms"] * 1e-3),
                      "ms_per_batch": round(r["median_ms"], 4),
                      "gbs": round(by / r["median_ms"] / 1e6, 1), "class_lanes": lanes,
                      "work_items": int(b.plan.items)}
        log(f"batched {kname}: {r['median_ms']:.3f} ms per launch of {count} SpMVs "
            f"({res[kname]['spmv_per_s']:.3g} SpMV/s, {res[kname]['gbs']} GB/s)")
        del b, entries, bufs
    res["matrices"], res["rows"], res["nnz"] = count, n, nnz
    extra["batched_10k"] = res


def bench_section(ctx, args):
    """One section (SpMM or the conversions on the large matrix, one timed
    call per form; or the precision or batched case), for counter runs
    (rocprofv3 --pmc) of its kernels."""
    extra, guard = {}, {}
    if args.calibrate:
        calibration_copy(ctx)
    if args.only == "precision":
        bench_precision(ctx, extra)
    elif args.only == "batched":
        bench_batched(ctx, extra, args.batch_count)
    elif args.only == "finalize":
        t0 = time.perf_counter()
        m = G.zipf_large(args.rows)
        log(f"generated n={m.rows} nnz={m.nnz} in {time.perf_counter() - t0:.1f}s")
        bench_shuffled_finalize(ctx, m, extra, guard)
        bench_finalize(ctx, extra)
    else:
        t0 = time.perf_counter()
        m = G.zipf_large(args.rows)
        log(f"generated n={m.rows} nnz={m.nnz} in {time.perf_counter() - t0:.1f}s")
        a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
        if args.only == "spmm":
            bench_spmm(ctx, m, a, np.diff(m.offsets), extra, guard, steps=1)
        else:
            bench_conversions(ctx, m, a, extra, guard, steps=1)
    bad = guard_failures(guard)
    out = {"section": args.only, "extra": extra, "guard": guard}
    if bad:
        out["error"] = "output check failed: " + ", ".join(bad)
    print(json.dumps(out), flush=True)
    if bad:
        sys.exit(3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=1 << 22)
    ap.add_argument("--alg", default="auto",
                    help="auto (column-panel plan, learned pick reported), merge, ordered, vecN")
    ap.add_argument("--staging", default="stream", choices=sorted(STAGINGS),
                    help="x-panel staging of the column-panel plan the timed steps run")
    ap.add_argument("--cpu-sample-rows", type=int, default=1 << 18)
    ap.add_argument("--cpu-reps", type=int, default=20)
    ap.add_argument("--batch-count", type=int, default=10_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline line only")
    ap.add_argument("--sweep", action="store_true", help="time every SpMV variant")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one 1 GiB device copy (counter calibration)")
    ap.add_argument("--only", choices=("spmm", "convert", "finalize", "precision", "batched"),
                    help="profiling aid: run one section (the large-matrix ones with one timed "
                         "call per form) and print its own JSON line instead of the benchmark line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if H.device_count() < 1:
        print(json.dumps({"error": "no HIP device visible"}))
        sys.exit(1)
    ctx = H.Context(0)
    extra = {}
    if args.only:
        bench_section(ctx, args)
        return
    m, alg, res, nbytes, y_ord, guard, x = bench_large(ctx, args, extra)
    name = H.ALG_NAMES[alg]
    gbs = nbytes / (res["loop_ms"] * 1e-3) / 1e9
    kgbs = nbytes / (res["mean_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu, y_cpu = cpu_baseline(m, x, args.cpu_sample_rows, args.cpu_reps)
        log(f"cpu baseline {cpu['value']} GB/s ({time.perf_counter() - t0:.1f}s)")
        if not args.quick:
            t0 = time.perf_counter()
            extra["cpu_conversions"] = cpu_conversions(m, args.cpu_sample_rows)
            log(f"cpu conversions ({time.perf_counter() - t0:.1f}s): " +
                ", ".join(f"{k} {v['ms']} ms" for k, v in extra["cpu_conversions"].items()))
        # The device's ordered kernel against the reference's own output.
        guard["ordered_vs_cpu_reference"] = {
            "rows": int(y_cpu.shape[0]),
            "bit_exact": bool(np.array_equal(y_ord[: y_cpu.shape[0]].view(np.uint64),
                                             y_cpu.view(np.uint64)))}
    rows, nnz = m.rows, m.nnz
    del m, y_ord
    ceiling = bench_hbm_ceiling(ctx)
    if not args.quick:
        bench_gather_ceiling(ctx, extra, nbytes / nnz)
        bench_precision(ctx, extra)
        bench_dense_rows(ctx, extra)
        bench_triangular(ctx, extra)
        bench_finalize(ctx, extra)
        bench_batched(ctx, extra, args.batch_count)
    plan_tag = name if alg != H.ALG_PANEL else ("stream" if args.staging == "stream" else "panel")
    traffic, traffic_src = measured_traffic(rows, nnz, plan_tag)
    out = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["loop_ms"], 4),
        "host_enqueue_wait_ms": round(res["host_enqueue_wait_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single large matrix: n=2^22 square, Zipf(1.5) row lengths "
                               "clipped to [1,1e4], int32 columns, int64 offsets, N(0,1) FP64 "
                               "values, x U[-1,1], alpha=1, beta=0",
                   "rows": rows, "cols": rows, "nnz": nnz, "algorithm": name,
                   "selection": (f"column-panel analysis, {args.staging} staging"
                                 if alg == H.ALG_PANEL else "forced"),
                   "bytes_per_spmv": nbytes, "hbm_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(kgbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(kgbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "measured_ceiling": ceiling,
                     "frac_of_measured_read": round(kgbs / ceiling["read_gbs"], 4),
                     "frac_of_measured_copy": round(kgbs / ceiling["copy_gbs"], 4)},
        "cpu_baseline": cpu,
        "guard": guard,
        "extra": extra,
    }
    failed = guard_failures(guard)
    if failed:
        out["error"] = "output check failed: " + ", ".join(failed)
    print(json.dumps(out), flush=True)
    ctx.close()
    if failed:
        sys.exit(3)


if __name__ == "__main__":
    main()
