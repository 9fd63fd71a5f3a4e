"""The network oracle (oracle/netref.py) pinned by central finite differences and by the
adjoint identity of the graph aggregation."""
import numpy as np

import netref


def tiny_graph():
    # 5 columns in a ring plus one chord
    nb = {0: [1, 4, 2], 1: [0, 2], 2: [1, 3, 0], 3: [2, 4], 4: [3, 0]}
    ptr = np.zeros(6, np.int32)
    idx = []
    for c in range(5):
        idx += sorted(nb[c])
        ptr[c + 1] = len(idx)
    return ptr, np.array(idx, np.int32), 5


def test_aggregate_adjoint():
    g = tiny_graph()
    rng = np.random.default_rng(0)
    A = rng.normal(size=(10, 3))   # two samples
    G = rng.normal(size=(10, 3))
    lhs = (netref.aggregate(A, *g) * G).sum()
    rhs = (A * netref.aggregate_t(G, *g)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gradients_match_finite_differences():
    g = tiny_graph()
    kp, L, h = 8, 6, 8
    p = netref.init_params(kp, h, seed=3, scale_out=1.0)
    rng = np.random.default_rng(1)
    for k in p:
        p[k] = p[k] + 0.05 * rng.normal(size=p[k].shape)   # nonzero biases, generic signs
    X = rng.normal(size=(10, kp)); X[:, L:] = 0
    T = rng.normal(size=(10, kp))
    loss, grad, _ = netref.loss_and_grad(p, X, T, L, g)
    eps = 1e-6
    for name in p:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = netref.loss_and_grad(p, X, T, L, g)[0]
            flat[idx] = old - eps
            lm = netref.loss_and_grad(p, X, T, L, g)[0]
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = grad[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)), (name, idx, fd, an)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0, 1e-3])
    r = netref.bf16_round(x)
    # bf16 keeps 8 significant bits: relative rounding error <= 2^-8
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    assert r[0] == 1.0 and r[1] == 1.0   # 1 + 2^-8 is a tie -> even
    assert r[2] == 1.0078125             # rounds up
