"""FP64 numpy restatement of the learned preconditioner network (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference. This module is the oracle its HIP
forward and backward kernels are checked against, and it is itself pinned by central
finite differences (tests/test_netref.py).

Network (per vertical column c, features = the column's residual levels, width Kp):
    A1 = relu(X  W1 + b1)                      Kp  -> H
    A2 = relu(A1 W2 + b2)                      H   -> H
    A3 = relu([A2 | Agg A2] W3 + b3)           2H  -> H   (graph convolution)
    A4 = relu([A3 | Agg A3] W4 + b4)           2H  -> H   (graph convolution)
    A5 = relu(A4 W5 + b5)                      H   -> H
    Y  = A5 W6 + b6                            H   -> Kp
Agg = mean over the column's horizontal neighbours (rows of one sample only).
Loss = sum over valid entries (first L of Kp) of (Y - T)^2 / count.
With bf16=True, every GEMM operand (activations and weights) is rounded to bfloat16
(round to nearest even) before an exact FP64 product, as the MFMA path consumes them.
"""
from __future__ import annotations

import numpy as np

LAYERS = ("W1", "W2", "W3", "W4", "W5", "W6")


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounding = 0x7FFF + ((u >> 16) & 1)
    u = ((u + rounding) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shapes(kp: int, h: int = 128):
    return {"W1": (kp, h), "W2": (h, h), "W3": (2 * h, h), "W4": (2 * h, h), "W5": (h, h), "W6": (h, kp)}


def init_params(kp: int, h: int = 128, seed: int = 0, scale_out: float = 0.1):
    rng = np.random.default_rng(seed)
    p = {}
    for name, (k, n) in shapes(kp, h).items():
        std = np.sqrt(2.0 / k) * (scale_out if name == "W6" else 1.0)
        p[name] = rng.normal(0.0, std, size=(k, n))
        p["b" + name[1:]] = np.zeros(n)
    return p


def aggregate(A, nbr_ptr, nbr_idx, ncol):
    """Mean over horizontal neighbours, per sample block of ncol rows."""
    M = A.shape[0]
    out = np.zeros_like(A)
    deg = np.diff(nbr_ptr)
    for s0 in range(0, M, ncol):
        blk = A[s0:s0 + ncol]
        acc = np.zeros_like(blk)
        for c in range(ncol):
            acc[c] = blk[nbr_idx[nbr_ptr[c]:nbr_ptr[c + 1]]].sum(axis=0)
        out[s0:s0 + ncol] = acc / deg[:, None]
    return out


def aggregate_t(G, nbr_ptr, nbr_idx, ncol):
    """Transpose of `aggregate` (adjacency is symmetric: sum over neighbours of G/deg)."""
    return _aggregate_t(G, nbr_ptr, nbr_idx, ncol, np.diff(nbr_ptr).astype(np.float64))


def _aggregate_t(G, nbr_ptr, nbr_idx, ncol, deg):
    M = G.shape[0]
    out = np.zeros_like(G)
    for s0 in range(0, M, ncol):
        g = G[s0:s0 + ncol] / deg[:, None]
        acc = np.zeros_like(g)
        for c in range(ncol):
            acc[c] = g[nbr_idx[nbr_ptr[c]:nbr_ptr[c + 1]]].sum(axis=0)
        out[s0:s0 + ncol] = acc
    return out


def forward(p, X, graph, bf16=False):
    nbr_ptr, nbr_idx, ncol = graph
    r = bf16_round if bf16 else (lambda a: a)

    def lin(A, W, b):
        return r(A) @ r(W) + b

    c = {}
    c["X"] = X
    c["A1"] = np.maximum(lin(X, p["W1"], p["b1"]), 0.0)
    c["A2"] = np.maximum(lin(c["A1"], p["W2"], p["b2"]), 0.0)
    c["C2"] = np.concatenate([c["A2"], aggregate(c["A2"], nbr_ptr, nbr_idx, ncol)], axis=1)
    c["A3"] = np.maximum(lin(c["C2"], p["W3"], p["b3"]), 0.0)
    c["C3"] = np.concatenate([c["A3"], aggregate(c["A3"], nbr_ptr, nbr_idx, ncol)], axis=1)
    c["A4"] = np.maximum(lin(c["C3"], p["W4"], p["b4"]), 0.0)
    c["A5"] = np.maximum(lin(c["A4"], p["W5"], p["b5"]), 0.0)
    c["Y"] = lin(c["A5"], p["W6"], p["b6"])
    return c


def loss_and_grad(p, X, T, L, graph, bf16=False):
    """Returns loss, gradients (dict like p) and the forward cache."""
    nbr_ptr, nbr_idx, ncol = graph
    r = bf16_round if bf16 else (lambda a: a)
    c = forward(p, X, graph, bf16)
    M, kp = c["Y"].shape
    mask = np.zeros(kp)
    mask[:L] = 1.0
    count = M * L
    diff = (c["Y"] - T) * mask
    loss = float((diff * diff).sum() / count)
    g = {}
    dZ = 2.0 * diff / count                        # output layer, no activation

    def back(dZ, A, Wname):
        g[Wname] = r(A).T @ r(dZ)
        g["b" + Wname[1:]] = dZ.sum(axis=0)
        return r(dZ) @ r(p[Wname]).T

    dA5 = back(dZ, c["A5"], "W6")
    dA4 = back(dA5 * (c["A5"] > 0), c["A4"], "W5")
    dC3 = back(dA4 * (c["A4"] > 0), c["C3"], "W4")
    h = c["A3"].shape[1]
    dA3 = dC3[:, :h] + _aggregate_t(dC3[:, h:], nbr_ptr, nbr_idx, ncol, np.diff(nbr_ptr).astype(np.float64))
    dC2 = back(dA3 * (c["A3"] > 0), c["C2"], "W3")
    dA2 = dC2[:, :h] + _aggregate_t(dC2[:, h:], nbr_ptr, nbr_idx, ncol, np.diff(nbr_ptr).astype(np.float64))
    dA1 = back(dA2 * (c["A2"] > 0), c["A1"], "W2")
    back(dA1 * (c["A1"] > 0), c["X"], "W1")
    return loss, g, c


def column_graph(mesh):
    """Horizontal neighbour lists of the columns (from the 2-D triangulation)."""
    tri = mesh.tri
    nh = mesh.nh
    pairs = set()
    for a, b, c in tri.T:
        for u, v in ((a, b), (b, c), (a, c)):
            pairs.add((min(u, v), max(u, v)))
    nb = [[] for _ in range(nh)]
    for u, v in sorted(pairs):
        nb[u].append(v)
        nb[v].append(u)
    ptr = np.zeros(nh + 1, dtype=np.int32)
    for c in range(nh):
        nb[c].sort()
        ptr[c + 1] = ptr[c] + len(nb[c])
    idx = np.array([v for c in range(nh) for v in nb[c]], dtype=np.int32)
    return ptr, idx, nh
