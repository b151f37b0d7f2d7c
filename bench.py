"""Benchmark: CSR SpMV GB/s and fraction of MI355X HBM peak.

Default workload (one MI355X): the single large matrix -- n = 2^22 rows and
columns, row lengths Zipf(s=1.5) clipped to [1, 10^4] (about 6.4e8 nnz),
sorted distinct uniform int32 columns, N(0,1) values, FP64, alpha=1, beta=0.
One step = one compat SpMV y = alpha*A*x + beta*y over the whole matrix,
inputs resident in HBM.  `value` = algorithmic bytes per SpMV / time:

  B = nnz*(s_v + s_i) + (rows+1)*s_o + cols*s_x + rows*s_y*(1 + [beta != 0])

Prints one JSON line (see the repository's benchmark contract).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "asc-cpp_amd"))

import asc_sparse_hip as H  # noqa: E402
from asc_sparse_hip import gen as G  # noqa: E402

METRIC = "CSR SpMV GB/s & %HBM peak (FP64 plain vs BF16/learned-select); batched SpMV/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(rows, cols, nnz, sv, si, sx, sy, so=8, beta_nonzero=False):
    return nnz * (sv + si) + (rows + 1) * so + cols * sx + rows * sy * (2 if beta_nonzero else 1)


def time_spmv(ctx, a, xd, yd, alg, steps, warmup, acc=H.ACC_PLAIN):
    ws = H.Workspace(ctx)
    for _ in range(warmup):
        H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=alg, acc=acc, ws=ws)
    ctx.synchronize()
    # Whole-loop timing (enqueue + run), and per-call events on the same stream.
    starts, stops = [], []
    t0 = time.perf_counter()
    first = ctx.record()
    for _ in range(steps):
        starts.append(ctx.record())
        H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=alg, acc=acc, ws=ws)
        stops.append(ctx.record())
    last = ctx.record()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    per_call = [s.elapsed_ms(e) for s, e in zip(starts, stops)]
    total_ms = first.elapsed_ms(last)
    return {"ms_per_step": total_ms / steps, "kernel_ms_mean": float(np.mean(per_call)),
            "kernel_ms_median": float(np.median(per_call)), "wall_s": wall}


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


def cpu_baseline(m, sample_rows, reps=3):
    """Serial reference compat SpMV on the first `sample_rows` rows."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as O
    r = min(sample_rows, m.rows)
    nnz = int(m.offsets[r])
    off = np.ascontiguousarray(m.offsets[: r + 1])
    idx = m.indices[:nnz].astype(np.int64)
    val = np.ascontiguousarray(m.values[:nnz])
    x = G.vector(m.cols, 12345)
    y = np.zeros(r)
    kind = "port"
    fn = lambda: O.compat_spmv(off, idx, val, 1.0, x, 0.0, y)  # noqa: E731
    ref = getattr(O, "ref_compat_spmv", None)
    if ref is not None and O.ref_available():
        kind = "reference"
        fn = lambda: ref(off, idx, val, 1.0, x, 0.0, y)  # noqa: E731
    fn()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    b = spmv_bytes(r, m.cols, nnz, 8, 8, 8, 8)  # the reference's int64 indices
    return {"value": b / t / 1e9, "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"rows [0, {r}) of the same matrix ({nnz} nnz, int64 indices as in the "
                      f"reference), compat SpMV FP64, median of {reps} after 1 warmup, "
                      f"{t * 1e3:.1f} ms per SpMV"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=1 << 22)
    ap.add_argument("--alg", default="merge")
    ap.add_argument("--cpu-sample-rows", type=int, default=1 << 19)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--sweep", action="store_true", help="time every SpMV variant")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if H.device_count() < 1:
        print(json.dumps({"error": "no HIP device visible"}))
        sys.exit(1)
    algs = {"merge": H.ALG_MERGE, "ordered": H.ALG_ORDERED}
    algs.update({f"vec{g}": v for g, v in H.ALG_VECTOR.items()})
    alg = algs[args.alg]

    t0 = time.perf_counter()
    m = G.zipf_large(args.rows)
    log(f"generated n={m.rows} nnz={m.nnz} in {time.perf_counter() - t0:.1f}s")
    ctx = H.Context(0)
    t0 = time.perf_counter()
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    x = G.vector(m.cols, 12345)
    xb = ctx.upload(x)
    yb = ctx.alloc(8 * m.rows)
    xd = H.vector(xb, m.cols, H.F64, writable=False)
    yd = H.vector(yb, m.rows, H.F64)
    log(f"staged on device in {time.perf_counter() - t0:.1f}s")

    if args.calibrate:
        # A plain copy of known size: HBM counters read ~1 GiB fetched and
        # ~1 GiB written for it, which scales the SpMV kernel's counters.
        ca, cb = ctx.alloc(1 << 30), ctx.alloc(1 << 30)
        ctx.fill(ca, 1)
        ctx.copy(cb.ptr, ca.ptr, 1 << 30)
        ctx.synchronize()
        del ca, cb
    res = time_spmv(ctx, a, xd, yd, alg, args.steps, args.warmup)
    nbytes = spmv_bytes(m.rows, m.cols, m.nnz, 8, 4, 8, 8)
    gbs = nbytes / (res["ms_per_step"] * 1e-3) / 1e9
    kgbs = nbytes / (res["kernel_ms_mean"] * 1e-3) / 1e9
    log(f"FP64 {args.alg}: {res['ms_per_step']:.3f} ms/step, {gbs:.0f} GB/s "
        f"(per-call events {res['kernel_ms_mean']:.3f} ms, {kgbs:.0f} GB/s)")

    extra = {}
    if args.sweep:
        sweep = {}
        for name, v in algs.items():
            r = time_spmv(ctx, a, xd, yd, v, max(3, args.steps // 4), 2)
            sweep[name] = round(nbytes / (r["kernel_ms_mean"] * 1e-3) / 1e9, 1)
            log(f"  sweep {name}: {r['kernel_ms_mean']:.3f} ms  {sweep[name]} GB/s")
        extra["fp64_variant_gbs"] = sweep
    if not args.no_extra:
        # FP32-stored values of the same matrix (FP32 x, y).
        a32 = H.convert_values(ctx, a, H.F32)
        xb32 = ctx.upload(x.astype(np.float32))
        yb32 = ctx.alloc(4 * m.rows)
        r32 = time_spmv(ctx, a32, H.vector(xb32, m.cols, H.F32, writable=False),
                        H.vector(yb32, m.rows, H.F32), alg, args.steps, args.warmup)
        b32 = spmv_bytes(m.rows, m.cols, m.nnz, 4, 4, 4, 4)
        extra["fp32_gbs"] = round(b32 / (r32["kernel_ms_mean"] * 1e-3) / 1e9, 1)
        extra["fp32_rows_per_s"] = m.rows / (r32["kernel_ms_mean"] * 1e-3)
        extra["fp64_rows_per_s"] = m.rows / (res["kernel_ms_mean"] * 1e-3)
        log(f"FP32 {args.alg}: {r32['kernel_ms_mean']:.3f} ms, {extra['fp32_gbs']} GB/s")
        del a32, xb32, yb32

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(m, args.cpu_sample_rows)
        log(f"cpu baseline {cpu['value']:.2f} GB/s ({time.perf_counter() - t0:.1f}s)")

    traffic, traffic_src = measured_traffic(m.rows, m.nnz, args.alg)
    out = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single large matrix: n=2^22 square, Zipf(1.5) row lengths "
                               "clipped to [1,1e4], int32 columns, int64 offsets, N(0,1) FP64 "
                               "values, x U[-1,1], alpha=1, beta=0",
                   "rows": m.rows, "cols": m.cols, "nnz": m.nnz, "algorithm": args.alg,
                   "bytes_per_spmv": nbytes, "hbm_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(kgbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(kgbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src},
        "cpu_baseline": cpu,
        "extra": extra,
        "host": platform.node(),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
