"""CPU restatement of the learned refinement proposer (test infrastructure
only): element features, the network's forward pass, its loss and gradient
in float64.  The GPU kernels are checked against it; it is pinned by finite
differences and closed-form cases in tests/test_proposer_pins.py.

Layout (include/adg.h): 12 inputs, two hidden layers of 64 ReLU units, one
logistic output; parameters row-major W1[64][12], b1[64], W2[64][64], b2[64],
W3[64], b3.  The loss is the weighted logistic loss
    L = sum_b w_b [log(1 + e^{z_b}) - y_b z_b] / sum_b w_b.
"""
import numpy as np

F, H = 12, 64
NPARAMS = H * F + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x):
    """logits z, probabilities, hidden activations"""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, F)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    z = h2 @ W3 + b3
    return z, 1.0 / (1.0 + np.exp(-z)), h1, h2


def loss_and_grad(p, x, y, w=None):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, F)
    y = np.asarray(y, np.float64)
    w = np.ones(len(y)) if w is None else np.asarray(w, np.float64)
    z, prob, h1, h2 = forward(p, x)
    ws = w.sum()
    loss = np.sum(w * (np.logaddexp(0.0, z) - y * z)) / ws
    dz = w * (prob - y) / ws
    gW3 = h2.T @ dz
    gb3 = dz.sum()
    dh2 = np.outer(dz, W3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g


def features(level, suppt, u, eps, nmax):
    """Features and leaf flags of every element of one instance (arrays in the
    library's element order), as include/adg.h defines them."""
    level = np.asarray(level)
    suppt = np.asarray(suppt)
    E, D = level.shape
    K = len(u) // E
    norm = np.sqrt(np.sum(np.asarray(u, np.float64).reshape(E, K) ** 2, axis=1))
    where = {(tuple(l), tuple(s)): e for e, (l, s) in enumerate(zip(level.tolist(), suppt.tolist()))}

    def f(nrm):
        return float(np.clip(np.log10(max(nrm, 1e-30) / eps) * 0.25, -4.0, 4.0))

    feats = np.zeros((E, F))
    leaf = np.zeros(E, np.int32)
    for e in range(E):
        l, s = list(level[e]), list(suppt[e])
        feats[e, 0] = f(norm[e])
        have = total = 0
        for d in range(D):
            nl = nr = norm[e]
            if l[d] >= 2:
                cnt = 2 ** (l[d] - 1)
                i = (s[d] - 1) // 2
                vals = []
                for j in ((i - 1) % cnt, (i + 1) % cnt):
                    s2 = list(s)
                    s2[d] = 2 * j + 1
                    k = where.get((tuple(l), tuple(s2)))
                    vals.append(norm[k] if k is not None else 0.0)
                nl, nr = vals
            feats[e, 1 + 2 * d] = f(nl)
            feats[e, 2 + 2 * d] = f(nr)
            kids = []
            if l[d] == 0:
                kids = [(1, 1)]
            elif l[d] < nmax:
                kids = [(l[d] + 1, 2 * s[d] - 1), (l[d] + 1, 2 * s[d] + 1)]
            total += len(kids)
            for lc, sc in kids:
                l2, s2 = list(l), list(s)
                l2[d], s2[d] = lc, sc
                have += (tuple(l2), tuple(s2)) in where
        feats[e, 1 + 2 * D] = sum(l) / (D * nmax)
        feats[e, 2 + 2 * D] = max(l) / nmax
        leaf[e] = 1 if have < total else 0
    return feats, leaf
