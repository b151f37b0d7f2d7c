"""Learned SpMV-variant selector: data collection, GPU training, evaluation.

The selector maps the 16 GPU-computed sparsity features of a CSR matrix
(asc_sparse_hip_csr_features) through an MLP 16 -> 64 -> 64 -> 11 to the
predicted log(cost / best cost) of each variant: the eight direct kernels
(merge-path and vector-per-row with 1..64 lanes; cost = SpMV time) and the
three column-panel plans (L2-, LDS- and stream-staged x panels at their
default widths; cost = SpMV time + analysis time / SELECT_AMORTIZE).
Training data are exhaustive autotuner timings measured on the GPU: the
C ABI's autotuner over the direct kernels (asc_sparse_hip_csr_spmv_autotune)
and `time_plans` below over the plans (each analysed, then timed); the
network is trained by the library's own forward, gradient and Adam kernels.

Workload: seeded matrices from six generators (uniform-random, Zipf
power-law, banded, 8x8 block-diagonal, stencil, random bipartite), rows
log-uniform in [2^12, 2^22], FP64 values, int32 columns.

Usage on a machine with an MI355X:
    python -m asc_sparse_hip.selector collect --split train --start 0 --count 500 --out f.npz
    python -m asc_sparse_hip.selector train --data a.npz b.npz --out model.json
    python -m asc_sparse_hip.selector evaluate --model model.json --data held.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import threading
import time

import numpy as np

from . import (ALG_MERGE, ALG_NAMES, ALG_PANEL, CSR, F32, F64, SELECT_AMORTIZE, STAGING_L2,
               STAGING_LDS, STAGING_STREAM, TUNABLE_ALGS, AscHipError, Context, Selection,
               Selector, Workspace, autotune, check, clone_compressed, csr_spmv, csr_spmv_analyze,
               features, mlp_sizes, panel_options, plan_clear, vector)
from . import gen as G

FAMILIES = ("uniform", "zipf", "banded", "blockdiag8", "stencil", "bipartite")
PLANS = ((STAGING_L2, "panel_l2"), (STAGING_LDS, "panel_lds"), (STAGING_STREAM, "panel_stream"))
DIRECT_K = len(TUNABLE_ALGS)  # merge, vec1 .. vec64
K = DIRECT_K + len(PLANS)     # and the three column-panel plans
OUTPUT_NAMES = [ALG_NAMES[a] for a in TUNABLE_ALGS] + [name for _, name in PLANS]
MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models")
DEFAULT_MODEL = os.path.join(MODEL_DIR, "selector_mi355x.json")
SPLIT_SEED = {"train": G.SEED_BASE + 4, "heldout": G.SEED_BASE + 5}
ROWS_LO, ROWS_HI = 1 << 12, 1 << 22


def spec(split: str, i: int) -> dict:
    """Deterministic description of matrix i of a split."""
    rng = np.random.default_rng([SPLIT_SEED[split], i])
    rows = int(round(math.exp(rng.uniform(math.log(ROWS_LO), math.log(ROWS_HI)))))
    fam = FAMILIES[i % len(FAMILIES)]
    seed = int(rng.integers(0, 2 ** 48))
    if fam == "uniform":
        per = int(round(math.exp(rng.uniform(math.log(2), math.log(64)))))
        return dict(name=fam, family=G.UNIFORM_VAR, rows=rows, per_row=per, seed=seed)
    if fam == "zipf":
        s = float(rng.uniform(1.3, 2.5))
        return dict(name=fam, family=G.ZIPF, rows=rows, zipf_s=s, zipf_max=min(rows, 10_000), seed=seed)
    if fam == "banded":
        return dict(name=fam, family=G.BANDED, rows=rows, band=int(rng.integers(2, 65)),
                    fill=float(rng.uniform(0.3, 1.0)), seed=seed)
    if fam == "blockdiag8":
        return dict(name=fam, family=G.BLOCKDIAG8, rows=rows, seed=seed)
    if fam == "stencil":
        if rng.random() < 0.5:
            k = max(2, int(round(rows ** 0.5)))
            return dict(name="stencil2d", family=G.STENCIL2D, rows=k * k, grid=k, seed=seed)
        k = max(2, int(round(rows ** (1 / 3))))
        return dict(name="stencil3d", family=G.STENCIL3D, rows=k ** 3, grid=k, seed=seed)
    per = int(round(math.exp(rng.uniform(math.log(2), math.log(32)))))
    return dict(name=fam, family=G.BIPARTITE, rows=rows, per_row=per, seed=seed)


def make(sp: dict) -> G.Csr:
    kw = {k: v for k, v in sp.items() if k not in ("name", "family", "rows", "seed")}
    values = G.VALUES_LAPLACIAN if sp["family"] in (G.STENCIL2D, G.STENCIL3D) else G.VALUES_UNIFORM
    return G.make(sp["family"], sp["rows"], sp["rows"], seed=sp["seed"], values=values, **kw)


# ---- data collection ---------------------------------------------------------------

def time_plans(ctx: Context, a, xd, yd, reps: int, ws: Workspace):
    """(SpMV ms, analysis ms) of each column-panel plan in PLANS: analysis
    timed on the host around the (synchronizing) analyze call, the SpMV by
    events over `reps` launches after a warm-up.  A plan the library refuses
    for the matrix costs infinity."""
    spmv = np.full(len(PLANS), np.inf, np.float32)
    analysis = np.full(len(PLANS), np.inf, np.float32)
    for j, (staging, _) in enumerate(PLANS):
        try:
            ctx.synchronize()
            t0 = time.perf_counter()
            csr_spmv_analyze(ctx, a, panel_options(staging=staging))
            ctx.synchronize()
            analysis[j] = (time.perf_counter() - t0) * 1e3
            csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=ALG_PANEL, ws=ws)
            e0 = ctx.record()
            for _ in range(reps):
                csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=ALG_PANEL, ws=ws)
            e1 = ctx.record()
            ctx.synchronize()
            spmv[j] = e0.elapsed_ms(e1) / reps
        except AscHipError:
            pass
        finally:
            ctx.synchronize()
            if getattr(a, "analysis", None) is not None:
                a.analysis.free()
                a.analysis = None
            plan_clear(ctx)
    return spmv, analysis


def costs(d: dict) -> np.ndarray:
    """Per-variant cost (ms) of a dataset: the direct kernels' SpMV times, then
    each plan's SpMV time plus its analysis spread over SELECT_AMORTIZE SpMVs."""
    return np.concatenate([d["times"], d["plan_spmv"] + d["plan_analysis"] / SELECT_AMORTIZE],
                          axis=1).astype(np.float64)


def collect(ctx: Context, split: str, start: int, count: int, reps: int = 3, log=print):
    """Features and exhaustive timings of matrices [start, start+count)."""
    feats = np.zeros((count, 16), np.float32)
    times = np.zeros((count, DIRECT_K), np.float32)
    plan_spmv = np.zeros((count, len(PLANS)), np.float32)
    plan_analysis = np.zeros((count, len(PLANS)), np.float32)
    nnz = np.zeros(count, np.int64)
    names = []
    ws = Workspace(ctx)
    box = {}

    def produce(i):
        box[i] = make(spec(split, start + i))

    th = threading.Thread(target=produce, args=(0,))
    th.start()
    t0 = time.time()
    for i in range(count):
        th.join()
        m = box.pop(i)
        if i + 1 < count:  # generate the next matrix while this one is timed
            th = threading.Thread(target=produce, args=(i + 1,))
            th.start()
        a = clone_compressed(ctx, CSR, m.rows, m.cols, m.offsets, m.indices, m.values, F64, 4)
        xb = ctx.upload(G.vector(m.cols, 5))
        yb = ctx.alloc(8 * m.rows)
        xd = vector(xb, m.cols, F64, writable=False)
        yd = vector(yb, m.rows, F64)
        feats[i] = features(ctx, a, ws)
        _, t = autotune(ctx, a, xd, yd, reps, ws)
        times[i] = [t[alg] for alg in TUNABLE_ALGS]
        plan_clear(ctx)
        plan_spmv[i], plan_analysis[i] = time_plans(ctx, a, xd, yd, reps, ws)
        nnz[i] = m.nnz
        names.append(spec(split, start + i)["name"])
        del a, xb, yb, m
        if (i + 1) % 25 == 0 or i + 1 == count:
            log(f"[selector] {split} {start + i + 1}/{start + count}  {time.time() - t0:.0f}s", flush=True)
    return dict(features=feats, times=times, plan_spmv=plan_spmv, plan_analysis=plan_analysis,
                nnz=nnz, names=np.array(names), index=np.arange(start, start + count))


# ---- training on the GPU -------------------------------------------------------------

# A refused plan's cost, in the training targets: log-regret of this much.
TARGET_CAP = 4.0


def targets(cost: np.ndarray) -> np.ndarray:
    """log(cost / best cost) per variant, capped at TARGET_CAP (refused plans
    and very slow variants alike)."""
    t = np.log(cost / cost.min(axis=1, keepdims=True))
    return np.minimum(t, TARGET_CAP).astype(np.float32)


def init_params(k: int = K, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((16, 64), (64, 64), (64, k)):
        bound = math.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-bound, bound, (fan_out, fan_in)).astype(np.float32).ravel())
        parts.append(np.zeros(fan_out, np.float32))
    return np.concatenate(parts)


class DeviceMLP:
    """Training state on the device; every step runs the library's kernels."""

    def __init__(self, ctx: Context, params: np.ndarray, batch: int, k: int = K):
        self.ctx, self.k, self.batch = ctx, k, batch
        n, scratch = mlp_sizes(k, batch)
        assert params.shape[0] == n
        self.n = n
        self.p = ctx.upload(params.astype(np.float32))
        self.g = ctx.alloc(4 * n)
        self.m = ctx.alloc(4 * n)
        self.v = ctx.alloc(4 * n)
        ctx.fill(self.m, 0)
        ctx.fill(self.v, 0)
        self.scratch = ctx.alloc(scratch)
        self.loss = ctx.alloc(4)
        self.step_no = 0

    def gradient(self, xb, tb):
        check(self.ctx.lib.asc_sparse_hip_mlp_gradient(
            self.ctx.handle, self.k, self.p.ptr, self.batch, xb.ptr, tb.ptr, self.g.ptr,
            self.loss.ptr, self.scratch.ptr, self.scratch.nbytes))

    def step(self, xb, tb, lr=1e-3):
        self.gradient(xb, tb)
        self.step_no += 1
        check(self.ctx.lib.asc_sparse_hip_mlp_adam(self.ctx.handle, self.p.ptr, self.g.ptr, self.m.ptr,
                                                   self.v.ptr, self.n, lr, 0.9, 0.999, 1e-8,
                                                   self.step_no))

    def forward(self, xb, batch):
        _, scratch = mlp_sizes(self.k, batch)
        sb = self.ctx.alloc(scratch)
        yb = self.ctx.alloc(4 * batch * self.k)
        check(self.ctx.lib.asc_sparse_hip_mlp_forward(self.ctx.handle, self.k, self.p.ptr, batch,
                                                      xb.ptr, yb.ptr, sb.ptr, sb.nbytes))
        return self.ctx.download(yb, np.float32, batch * self.k).reshape(batch, self.k)

    def params(self):
        return self.ctx.download(self.p, np.float32, self.n)

    def last_loss(self):
        return float(self.ctx.download(self.loss, np.float32, 1)[0])


def normalize_on_device(ctx: Context, feats: np.ndarray, mean: np.ndarray, std: np.ndarray):
    n = feats.shape[0]
    xb = ctx.upload(np.ascontiguousarray(feats, np.float32))
    mb, sb = ctx.upload(mean.astype(np.float32)), ctx.upload(std.astype(np.float32))
    out = ctx.alloc(4 * 16 * n)
    check(ctx.lib.asc_sparse_hip_normalize_features(ctx.handle, xb.ptr, mb.ptr, sb.ptr, n, out.ptr))
    return out


def choose(pred: np.ndarray, norm: np.ndarray, margin: float, clip: float) -> np.ndarray:
    """Variant index per row, with the merge-path fallback when unsure."""
    best = pred.argmin(axis=1)
    unsure = (pred[:, 0] - pred[np.arange(pred.shape[0]), best] < margin) | \
        (np.abs(norm) > clip).any(axis=1)
    return np.where(unsure, 0, best)


def regret(cost: np.ndarray, pick: np.ndarray) -> float:
    """Geometric-mean regret against the exhaustive autotuner's best (a pick
    the library would refuse counts as a 10^6-fold regret)."""
    r = cost[np.arange(cost.shape[0]), pick] / cost.min(axis=1)
    return float(np.exp(np.mean(np.log(np.minimum(r, 1e6)))) - 1.0)


def train(ctx: Context, feats: np.ndarray, cost: np.ndarray, steps: int = 4000,
          lr: float = 2e-3, seed: int = 0, log=print) -> dict:
    """Trains on per-variant costs [n, k] (k = 8 direct, or 11 with the plans)."""
    k = cost.shape[1]
    mean = feats.mean(axis=0)
    std = feats.std(axis=0)
    std[std < 1e-6] = 1.0
    n = feats.shape[0]
    xb = normalize_on_device(ctx, feats, mean, std)
    tb = ctx.upload(targets(cost))
    net = DeviceMLP(ctx, init_params(k, seed), n, k)
    curve = []
    for s in range(steps):
        net.step(xb, tb, lr)
        if s % 500 == 0 or s == steps - 1:
            curve.append((s, net.last_loss()))
            log(f"[selector] step {s} loss {curve[-1][1]:.5f}", flush=True)
    norm = ctx.download(xb, np.float32, n * 16).reshape(n, 16)
    pred = net.forward(xb, n)
    # Fallback thresholds picked on the training set.
    best = (math.inf, 0.0, 6.0)
    for margin in (0.0, 0.02, 0.05, 0.1, 0.2):
        for clip in (4.0, 6.0, 10.0, 1e9):
            r = regret(cost, choose(pred, norm, margin, clip))
            if r < best[0] - 1e-9:
                best = (r, margin, clip)
    return dict(k=k, outputs=OUTPUT_NAMES[:k], amortize=SELECT_AMORTIZE, mean=mean.tolist(),
                std=std.tolist(), params=net.params().tolist(), margin=best[1], clip=best[2],
                train=dict(samples=n, steps=steps, lr=lr, loss_curve=curve, regret=best[0]))


def predict(ctx: Context, model: dict, feats: np.ndarray):
    n = feats.shape[0]
    xb = normalize_on_device(ctx, feats, np.array(model["mean"]), np.array(model["std"]))
    net = DeviceMLP(ctx, np.array(model["params"], np.float32), n, model["k"])
    pred = net.forward(xb, n)
    norm = ctx.download(xb, np.float32, n * 16).reshape(n, 16)
    return pred, norm


def evaluate(ctx: Context, model: dict, feats: np.ndarray, cost: np.ndarray) -> dict:
    """Regrets (geometric mean over matrices, against the exhaustive autotuner
    over every variant in `cost`) of the selector and of the methods without
    it: merge-path alone (its fallback), the best single variant, and the
    exhaustive autotuner over the direct kernels only."""
    pred, norm = predict(ctx, model, feats)
    pick = choose(pred, norm, model["margin"], model["clip"])
    raw = pred.argmin(axis=1)
    ratio = np.minimum(cost / cost.min(axis=1, keepdims=True), 1e6)
    single = int(np.argmin(np.exp(np.mean(np.log(ratio), axis=0))))
    direct_best = cost[:, :DIRECT_K].argmin(axis=1)
    return dict(regret_selector=regret(cost, pick), regret_argmin=regret(cost, raw),
                regret_merge_only=regret(cost, np.zeros(len(cost), int)),
                regret_best_single=regret(cost, np.full(len(cost), single)),
                best_single=OUTPUT_NAMES[single],
                regret_direct_autotuner=regret(cost, direct_best),
                fallback_rate=float(np.mean(pick == 0)),
                plan_pick_rate=float(np.mean(pick >= DIRECT_K)), samples=int(len(cost)),
                outputs=int(cost.shape[1]))


def device_selector(ctx: Context, model: dict):
    """A C-ABI selector (model blob on the device) for asc_sparse_hip_select."""
    blob = np.concatenate([np.array(model["mean"], np.float32), np.array(model["std"], np.float32),
                           np.array(model["params"], np.float32)])
    buf = ctx.upload(blob)
    sel = Selector(buf.ptr, model["k"], model["margin"], min(model["clip"], 3.0e38), 0)
    return sel, buf


def select(ctx: Context, model: dict, a):
    """Runs asc_sparse_hip_select with `model` on the device matrix `a`;
    returns (Selection, scores)."""
    sel, keep = device_selector(ctx, model)
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_select_workspace_size(ctx.handle, C.byref(sel), C.byref(need)))
    ws = ctx.alloc(need.value)
    out = Selection()
    scores = (C.c_float * model["k"])()
    check(ctx.lib.asc_sparse_hip_select(ctx.handle, C.byref(a.desc), C.byref(sel), ws.ptr,
                                        ws.nbytes, C.byref(out), scores))
    del keep, ws
    return out, [float(v) for v in scores]


def load_model(path: str = DEFAULT_MODEL) -> dict:
    return json.load(open(path))


def load_data(paths) -> dict:
    parts = [np.load(p) for p in paths]
    keys = ("features", "times", "plan_spmv", "plan_analysis", "nnz", "names", "index")
    return {k: np.concatenate([p[k] for p in parts]) for k in keys}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--split", default="train")
    c.add_argument("--start", type=int, default=0)
    c.add_argument("--count", type=int, default=100)
    c.add_argument("--reps", type=int, default=3)
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--steps", type=int, default=4000)
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--data", nargs="+", required=True)
    args = ap.parse_args(argv)
    ctx = Context(0)
    if args.cmd == "collect":
        d = collect(ctx, args.split, args.start, args.count, args.reps)
        np.savez(args.out, **d)
    elif args.cmd == "train":
        d = load_data(args.data)
        model = train(ctx, d["features"], costs(d), args.steps)
        json.dump(model, open(args.out, "w"))
        print(json.dumps({"train_regret": model["train"]["regret"], "margin": model["margin"],
                          "clip": model["clip"]}))
    else:
        d = load_data(args.data)
        print(json.dumps(evaluate(ctx, load_model(args.model), d["features"], costs(d))))
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
