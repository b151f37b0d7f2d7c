"""FP64 numpy restatement of the selector MLP and the sparsity features --
test infrastructure that checks the GPU kernels (features.hip, mlp.hip).

MLP: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3, loss =
mean over batch*k of (y - t)^2; parameter layout W1[64][16] b1 W2[64][64] b2
W3[k][64] b3.
"""
import numpy as np


def split(p, k):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((64, 16), (64,), (64, 64), (64,), (k, 64), (k,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def forward(p, x, k):
    w1, b1, w2, b2, w3, b3 = split(p, k)
    h1 = np.maximum(x @ w1.T + b1, 0.0)
    h2 = np.maximum(h1 @ w2.T + b2, 0.0)
    return h1, h2, h2 @ w3.T + b3


def loss_and_grad(p, x, t, k):
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    w1, b1, w2, b2, w3, b3 = split(p, k)
    h1, h2, y = forward(p, x, k)
    n = y.size
    loss = np.sum((y - t) ** 2) / n
    dy = 2.0 * (y - t) / n
    gw3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ w3) * (h2 > 0)
    gw2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ w2) * (h1 > 0)
    gw1 = dh1.T @ x
    gb1 = dh1.sum(0)
    g = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])
    return loss, g


def features(offsets, indices, rows, cols):
    """The 16 selector features (definitions in csrc/kernels/features.hip)."""
    offsets = np.asarray(offsets, np.int64)
    idx = np.asarray(indices, np.int64)
    lens = np.diff(offsets).astype(np.float64)
    nnz = int(offsets[-1])
    R = max(rows, 1)
    N = max(nnz, 1)
    mean = lens.sum() / R
    var = (lens ** 2).sum() / R - mean ** 2
    f = np.zeros(16)
    f[0], f[1], f[2] = np.log2(rows + 1.0), np.log2(cols + 1.0), np.log2(nnz + 1.0)
    f[3] = np.log2(1.0 + mean)
    f[4] = np.sqrt(max(var, 0.0)) / mean if mean > 0 else 0.0
    f[5] = np.log2(1.0 + (lens.max() if rows else 0.0))
    f[6] = np.sum(lens == 0) / R
    for j, (lo, hi) in enumerate(((1, 4), (4, 16), (16, 64), (64, 256), (256, np.inf))):
        f[7 + j] = np.sum((lens >= lo) & (lens < hi)) / R
    f[12] = lens[lens >= 256].sum() / N
    row_of = np.repeat(np.arange(rows), np.diff(offsets))
    diag = (row_of + 0.5) * (cols / rows if rows else 0.0)
    d = np.abs(idx - diag)
    f[13] = (d / cols).sum() / N if cols else 0.0
    f[14] = np.sum(d < 4096.0) / N
    first = np.zeros(nnz, bool)
    first[offsets[:-1][lens > 0]] = True
    gaps = np.diff(idx.astype(np.float64), prepend=0.0)[~first]
    f[15] = np.log2(1.0 + gaps).mean() if gaps.size else 0.0
    return f
