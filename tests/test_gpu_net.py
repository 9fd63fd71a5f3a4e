"""Learned preconditioner on the GPU: MFMA forward and backward against the FP64 oracle
(oracle/netref.py, bf16-rounded operands), the transpose of laplc against the dense
transpose of the CPU oracle's laplc, the physics-informed training loss and its gradient,
and what the trained network changes in the flexible GCR solve."""
import numpy as np
import pytest

import netref
from common import mesh_for, rel_l2, small_cfg, tiny_cfg
from oracle import Oracle
from src_amd import nft

pytestmark = pytest.mark.gpu
LINE = dict(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)


def unflatten(flat, kp, h):
    out, off = {}, 0
    for l, (k, n) in enumerate(netref.shapes(kp, h).values(), start=1):
        out[f"W{l}"] = flat[off:off + k * n].reshape(k, n).astype(np.float64); off += k * n
        out[f"b{l}"] = flat[off:off + n].astype(np.float64); off += n
    return out


@pytest.fixture(scope="module")
def small_net():
    m = mesh_for("small")
    s = nft.Solver(m, nft.make_params(small_cfg(**LINE)))
    s.net_init(hidden=128, seed=5, max_batch=4)
    yield m, s
    s.close()


def _batch(info, ns, seed):
    rng = np.random.default_rng(seed)
    kp, L = info["kp"], info["levels"]
    X = rng.normal(size=(ns * info["ncol"], kp)).astype(np.float32)
    X[:, L:] = 0
    T = (0.3 * rng.normal(size=X.shape)).astype(np.float32)
    T[:, L:] = 0
    return X, T


def test_forward_matches_oracle(small_net):
    m, s = small_net
    info = s.net_info()
    flat = s.net_params()
    # perturb the biases so every ReLU sees both signs
    rng = np.random.default_rng(2)
    p = unflatten(flat, info["kp"], 128)
    for l in range(1, 7):
        p[f"b{l}"] = 0.05 * rng.normal(size=p[f"b{l}"].shape)
    s.net_params(np.concatenate([np.concatenate([p[f"W{l}"].ravel(), p[f"b{l}"]]) for l in range(1, 7)]))
    X, _ = _batch(info, 2, 3)
    Y, _, _ = s.net_eval(X)
    graph = netref.column_graph(m)
    ref = netref.forward(p, X.astype(np.float64), graph, bf16=True)["Y"]
    err = rel_l2(Y[:, :info["levels"]], ref[:, :info["levels"]])
    # the same network without bf16 rounding of operands: a misplaced rounding point in any
    # layer moves the result by about this much, far more than FP32 accumulation does
    ref64 = netref.forward(p, X.astype(np.float64), graph, bf16=False)["Y"]
    err64 = rel_l2(Y[:, :info["levels"]], ref64[:, :info["levels"]])
    print("forward rel L2", err, "against the unrounded network", err64)
    assert err <= 1e-3
    assert err < 0.2 * err64


def test_gradients_match_oracle(small_net):
    m, s = small_net
    info = s.net_info()
    p = unflatten(s.net_params(), info["kp"], 128)
    X, T = _batch(info, 2, 4)
    Y, grad, loss = s.net_eval(X, T)
    graph = netref.column_graph(m)
    lref, gref, _ = netref.loss_and_grad(p, X.astype(np.float64), T.astype(np.float64), info["levels"], graph,
                                         bf16=True)
    assert abs(loss - lref) <= 1e-3 * lref
    g = unflatten(grad, info["kp"], 128)
    for name in gref:
        e = rel_l2(g[name], gref[name])
        print(name, "grad rel L2", e)
        assert e <= 3e-2, name


@pytest.mark.parametrize("impfl", [0, 1])
def test_laplc_transpose_matches_dense_oracle(impfl):
    """L^T y from the GPU against the transpose of the dense matrix assembled column by
    column from the CPU oracle's laplc (tiny2 mesh: 378 nodes, boundary maps of every kind)."""
    m = mesh_for("tiny2")
    cfg = tiny_cfg()
    o = Oracle(m, cfg)
    n = m.npoin
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = o.laplc(e, impfl)
        e[j] = 0.0
    o.close()
    s = nft.Solver(m, nft.make_params(cfg))
    rng = np.random.default_rng(7 + impfl)
    for _ in range(3):
        y = rng.normal(size=n)
        got = s.laplc_t(y, impfl)
        err = rel_l2(got, A.T @ y)
        print("L^T rel L2", err)
        assert err <= 1e-13
    s.close()


@pytest.mark.parametrize("impfl", [0, 1])
def test_laplc_transpose_adjoint(impfl):
    """<L x, y> = <x, L^T y> on the 40,000-node mesh."""
    m = mesh_for("small")
    s = nft.Solver(m, nft.make_params(small_cfg(**LINE)))
    rng = np.random.default_rng(3)
    for _ in range(3):
        x, y = rng.normal(size=m.npoin), rng.normal(size=m.npoin)
        a, b = np.dot(s.laplc(x, impfl), y), np.dot(x, s.laplc_t(y, impfl))
        print("adjoint", a, b)
        assert abs(a - b) <= 1e-12 * (np.linalg.norm(s.laplc(x, impfl)) * np.linalg.norm(y))
    s.close()


@pytest.fixture(scope="module")
def small_pairs():
    m = mesh_for("small")
    cfg = small_cfg(**LINE)
    s = nft.Solver(m, nft.make_params(cfg))
    s.net_init(hidden=128, seed=9, max_batch=4)
    s.state_init()
    s.dump_pairs(max_pairs=24, stride=5)
    s.run_steps(2)
    s.dump_pairs(0)
    yield m, cfg, s
    s.close()


def test_pair_loss_matches_oracle(small_pairs):
    """With a zero network z = q, so the loss is the mean of |L q - r|^2/|r|^2 over the
    pairs: compared with the CPU oracle's laplc applied to the stored pairs."""
    m, cfg, s = small_pairs
    npairs = s.pair_count()
    assert npairs >= 8
    keep = s.net_params()
    s.net_params(np.zeros_like(keep))
    got = s.pair_loss(0, 8)
    s.net_params(keep)
    r, q, sc, r2 = s.pairs_get(0, 8)
    assert np.allclose(sc, np.abs(r).max(axis=1), rtol=0, atol=0)
    assert np.allclose(r2, (r * r).sum(axis=1), rtol=1e-12)
    o = Oracle(m, cfg)
    ref = np.mean([np.sum((o.laplc(q[k], 1) - r[k]) ** 2) / r2[k] for k in range(8)])
    o.close()
    print("line-only loss", got, ref)
    assert abs(got - ref) <= 1e-10 * ref


@pytest.mark.parametrize("objective,linear", [(0, 0), (1, 0), (0, 1), (2, 0)])
def test_pair_loss_gradient(small_pairs, objective, linear):
    """Central difference of the loss along a random direction of the output bias b6
    against the directional derivative from the gradient: checks dY = s L^T(dl/dv)/B (error
    objective: dY = s dl/dz/B) and its column sums. The residual and error objectives are
    exactly quadratic in b6 (central differences are exact up to rounding); the angle
    objective is smooth, with an O(eps^2) difference error."""
    m, cfg, s = small_pairs
    info = s.net_info()
    flat = s.net_params()
    s.net_options(objective, linear)
    base = s.net_params()
    nb6 = info["kp"]
    off = flat.size - nb6
    l0, g = s.pair_loss(0, 4, grad=True)
    rng = np.random.default_rng(1)
    v = np.zeros(nb6, np.float32)
    v[:info["levels"]] = rng.normal(size=info["levels"])
    eps = 2e-3 if objective == 1 else 1e-2
    lp, lm = [], []
    for sign, out in ((1, lp), (-1, lm)):
        f = base.copy()
        f[off:] += np.float32(sign * eps) * v
        s.net_params(f)
        out.append(s.pair_loss(0, 4))
    s.net_options(0, 0)
    s.net_params(flat)
    fd = (lp[0] - lm[0]) / (2 * eps)
    an = float(np.dot(g[off:].astype(np.float64), v))
    print("loss", l0, "directional derivative", fd, an)
    assert abs(fd - an) <= 1e-3 * abs(fd) + 1e-9


def test_adam_update_matches_restatement(small_pairs):
    """Two Adam steps of nft_net_train on one pair (batch 1, so every minibatch is pair 0)
    against a float32 restatement of the update (b1 = 0.9, b2 = 0.999, eps = 1e-8, bias
    corrections 1 - b^t; no FMA contraction), from the gradient nft_net_pair_loss returns
    for the same pair and weights: weights, first and second moments equal bit for bit."""
    m, cfg, s = small_pairs
    saved = s.net_state()
    P = s.net_info()["nparams"]
    f32 = np.float32
    b1, b2, eps, lr = f32(0.9), f32(0.999), f32(1e-8), f32(1e-3)
    try:
        state = saved.copy()
        state[P:] = 0.0          # moments and step count start at zero
        s.net_state(state)
        w, mom, vel = state[:P].copy(), state[P:2 * P].copy(), state[2 * P:3 * P].copy()
        for t in (1, 2):
            _, g = s.pair_loss(0, 1, grad=True)
            s.net_train(steps=1, batch=1, lr=float(lr), seed=5, use_pairs=1)
            got = s.net_state()
            mom = b1 * mom + (f32(1) - b1) * g
            vel = b2 * vel + (f32(1) - b2) * g * g
            c1 = f32(1) - b1 ** f32(t)
            c2 = f32(1) - b2 ** f32(t)
            w = w - lr * (mom / c1) / (np.sqrt(vel / c2) + eps)
            assert got[3 * P] == t
            for name, a, b in (("m", got[P:2 * P], mom), ("v", got[2 * P:3 * P], vel), ("w", got[:P], w)):
                bad = np.flatnonzero(a != b)
                print(f"step {t} {name}: {bad.size} of {P} differ")
                assert bad.size == 0, (t, name, a[bad[:5]], b[bad[:5]])
    finally:
        s.net_state(saved)


def test_learned_preconditioner_end_to_end():
    """Train on (r, M_line^-1 r) pairs stored by plain line-preconditioned solves on the
    40,000-node mesh (residual objective), check that the loss falls on held-out pairs,
    then run the same steps with the learned correction: every solve still reaches 1e-8,
    the fields agree with the plain run to the solver tolerance, and the iteration counts
    and solve times of both runs are reported (on this problem the learned correction
    lowers the one-step residual but not the GCR iteration count)."""
    m = mesh_for("small")
    cfg = small_cfg(lord=4, **LINE)
    plain = nft.Solver(m, nft.make_params(cfg))
    plain.net_init(hidden=128, seed=1, max_batch=8)
    plain.dump_pairs(max_pairs=400, stride=2)
    plain.state_init()
    plain.run_steps(6)
    npairs = plain.pair_count()
    assert npairs > 100
    held = min(32, npairs // 4)
    before = plain.pair_loss(npairs - held, held)
    l0, l1 = plain.net_train(steps=600, batch=8, lr=1e-3, seed=3, use_pairs=npairs - held)
    after = plain.pair_loss(npairs - held, held)
    print(f"pairs {npairs}, training loss {l0:.4e} -> {l1:.4e}; held-out {held} pairs {before:.4e} -> {after:.4e}")
    assert l1 < l0 and after < before
    params = plain.net_params()
    plain.close()

    def run(learned):
        s = nft.Solver(m, nft.make_params(cfg))
        if learned:
            s.net_init(hidden=128, seed=1, max_batch=1)
            s.net_params(params)
            s.set_learned(1, stall=0.99)
        s.state_init()
        st = s.run_steps(6)
        fb, li = s.set_learned(-1)
        unk = s.state("unk")
        s.close()
        return st, unk, (fb, li)

    sp, up, _ = run(False)
    sl, ul, fb = run(True)
    gain = sp.gcr_iters / max(sl.gcr_iters, 1)
    print(f"plain: {sp.gcr_iters} iterations in {sp.gcr_solves} solves, {sp.ms_gcr:.1f} ms; "
          f"learned: {sl.gcr_iters} iterations, {sl.ms_gcr:.1f} ms, (fallbacks, learned iterations) {fb}; "
          f"iteration ratio {gain:.3f}, time ratio {sp.ms_gcr / max(sl.ms_gcr, 1e-9):.3f}")
    assert sl.gcr_unconverged == 0 and sp.gcr_unconverged == 0
    for k in range(4):
        assert rel_l2(ul[k], up[k]) <= 1e-5


def test_garbage_network_falls_back_to_line_preconditioner():
    """Failure injection: network weights replaced by large random values. The learned
    correction then ruins the first directions; the solve must drop it (fallbacks > 0)
    and still reach the tolerance with the line preconditioner, ending at the plain
    run's solution to the solver tolerance."""
    m = mesh_for("small")
    cfg = small_cfg(**LINE)

    def run(garbage):
        s = nft.Solver(m, nft.make_params(cfg))
        if garbage:
            s.net_init(hidden=128, seed=2, max_batch=1)
            rng = np.random.default_rng(0)
            s.net_params((50.0 * rng.normal(size=s.net_info()["nparams"])).astype(np.float32))
            s.set_learned(1, stall=0.99)
        s.state_init()
        st = s.run_steps(2)
        fb = s.set_learned(-1)[0]
        unk = s.state("unk")
        s.close()
        return st, fb, unk

    sp, _, up = run(False)
    sg, fb, ug = run(True)
    print("fallbacks", fb, "iterations", sp.gcr_iters, sg.gcr_iters)
    assert fb > 0
    assert sg.gcr_unconverged == 0
    for k in range(4):
        assert rel_l2(ug[k], up[k]) <= 1e-5


def test_forward_64_level_columns_matches_oracle():
    """64-level columns (Kp = 64): every layer runs the LDS-tiled MFMA kernel; forward and
    gradients against the FP64 restatement as above."""
    m = mesh_for("wave64")
    s = nft.Solver(m, nft.make_params(tiny_cfg(**LINE)))
    s.net_init(hidden=128, seed=7, max_batch=2)
    info = s.net_info()
    assert info["kp"] == 64
    flat = s.net_params()
    rng = np.random.default_rng(4)
    p = unflatten(flat, info["kp"], 128)
    for l in range(1, 7):
        p[f"b{l}"] = 0.05 * rng.normal(size=p[f"b{l}"].shape)
    s.net_params(np.concatenate([np.concatenate([p[f"W{l}"].ravel(), p[f"b{l}"]]) for l in range(1, 7)]))
    X, T = _batch(info, 2, 5)
    Y, grad, loss = s.net_eval(X, T)
    graph = netref.column_graph(m)
    ref = netref.forward(p, X.astype(np.float64), graph, bf16=True)["Y"]
    err = rel_l2(Y[:, :info["levels"]], ref[:, :info["levels"]])
    print("forward rel L2", err)
    assert err <= 1e-2
    s.close()
