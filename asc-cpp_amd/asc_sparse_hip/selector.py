"""Learned SpMV-variant selector: data collection, GPU training, evaluation.

The selector maps the 16 GPU-computed sparsity features of a CSR matrix
(asc_sparse_hip_csr_features) through an MLP 16 -> 64 -> 64 -> 8 to the
predicted log(time / best time) of each tunable variant (merge-path and
vector-per-row with 1..64 lanes).  Training data are exhaustive autotuner
timings (asc_sparse_hip_csr_spmv_autotune) measured on the GPU; the network
is trained by the library's own forward, gradient and Adam kernels.

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

from . import (ALG_MERGE, CSR, F32, F64, TUNABLE_ALGS, Context, Selector, Workspace, autotune,
               check, clone_compressed, features, mlp_sizes, vector)
from . import gen as G

FAMILIES = ("uniform", "zipf", "banded", "blockdiag8", "stencil", "bipartite")
K = len(TUNABLE_ALGS)  # 8 variants: merge, vec1 .. vec64
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
        return dict(name=fam, family=G.ZIPF, rows=rows, zipf_s=s, zipf_max=min(rows, 4096), seed=seed)
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

def collect(ctx: Context, split: str, start: int, count: int, reps: int = 3, log=print):
    """Features and exhaustive timings of matrices [start, start+count)."""
    feats = np.zeros((count, 16), np.float32)
    times = np.zeros((count, K), np.float32)
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
        nnz[i] = m.nnz
        names.append(spec(split, start + i)["name"])
        del a, xb, yb, m
        if (i + 1) % 25 == 0 or i + 1 == count:
            log(f"[selector] {split} {start + i + 1}/{start + count}  {time.time() - t0:.0f}s", flush=True)
    return dict(features=feats, times=times, nnz=nnz, names=np.array(names),
                index=np.arange(start, start + count))


# ---- training on the GPU -------------------------------------------------------------

def targets(times: np.ndarray) -> np.ndarray:
    """log(time / best time) per variant."""
    return np.log(times / times.min(axis=1, keepdims=True)).astype(np.float32)


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


def regret(times: np.ndarray, pick: np.ndarray) -> float:
    """Geometric-mean regret against the exhaustive autotuner's best."""
    r = times[np.arange(times.shape[0]), pick] / times.min(axis=1)
    return float(np.exp(np.mean(np.log(r))) - 1.0)


def train(ctx: Context, feats: np.ndarray, times: np.ndarray, steps: int = 4000,
          lr: float = 2e-3, seed: int = 0, log=print) -> dict:
    mean = feats.mean(axis=0)
    std = feats.std(axis=0)
    std[std < 1e-6] = 1.0
    n = feats.shape[0]
    xb = normalize_on_device(ctx, feats, mean, std)
    tb = ctx.upload(targets(times))
    net = DeviceMLP(ctx, init_params(K, seed), n)
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
            r = regret(times, choose(pred, norm, margin, clip))
            if r < best[0] - 1e-9:
                best = (r, margin, clip)
    return dict(k=K, algorithms=list(TUNABLE_ALGS), mean=mean.tolist(), std=std.tolist(),
                params=net.params().tolist(), margin=best[1], clip=best[2],
                train=dict(samples=n, steps=steps, lr=lr, loss_curve=curve, regret=best[0]))


def predict(ctx: Context, model: dict, feats: np.ndarray):
    n = feats.shape[0]
    xb = normalize_on_device(ctx, feats, np.array(model["mean"]), np.array(model["std"]))
    net = DeviceMLP(ctx, np.array(model["params"], np.float32), n)
    pred = net.forward(xb, n)
    norm = ctx.download(xb, np.float32, n * 16).reshape(n, 16)
    return pred, norm


def evaluate(ctx: Context, model: dict, feats: np.ndarray, times: np.ndarray) -> dict:
    pred, norm = predict(ctx, model, feats)
    pick = choose(pred, norm, model["margin"], model["clip"])
    raw = pred.argmin(axis=1)
    # Without the network: always merge-path (the fallback), or the best single variant.
    single = np.argmin(np.exp(np.mean(np.log(times / times.min(axis=1, keepdims=True)), axis=0)))
    return dict(regret_selector=regret(times, pick), regret_argmin=regret(times, raw),
                regret_merge_only=regret(times, np.zeros(len(times), int)),
                regret_best_single=regret(times, np.full(len(times), single)),
                best_single=TUNABLE_ALGS[int(single)], fallback_rate=float(np.mean(pick == 0)),
                samples=int(len(times)))


def device_selector(ctx: Context, model: dict):
    """A C-ABI selector (model blob on the device) for asc_sparse_hip_select."""
    blob = np.concatenate([np.array(model["mean"], np.float32), np.array(model["std"], np.float32),
                           np.array(model["params"], np.float32)])
    buf = ctx.upload(blob)
    sel = Selector(buf.ptr, model["k"], model["margin"], min(model["clip"], 3.0e38), 0)
    return sel, buf


def load_model(path: str = DEFAULT_MODEL) -> dict:
    return json.load(open(path))


def load_data(paths) -> dict:
    parts = [np.load(p) for p in paths]
    return {k: np.concatenate([p[k] for p in parts]) for k in ("features", "times", "nnz", "names", "index")}


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
        model = train(ctx, d["features"], d["times"], args.steps)
        json.dump(model, open(args.out, "w"))
        print(json.dumps({"train_regret": model["train"]["regret"], "margin": model["margin"],
                          "clip": model["clip"]}))
    else:
        d = load_data(args.data)
        print(json.dumps(evaluate(ctx, load_model(args.model), d["features"], d["times"])))
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
