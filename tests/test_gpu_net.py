"""Learned preconditioner on the GPU: MFMA forward and backward against the FP64 oracle
(oracle/netref.py, bf16-rounded operands), training on pairs the plain solver dumps, and
what the trained network changes in the flexible GCR solve."""
import numpy as np
import pytest

import netref
from common import mesh_for, rel_l2, small_cfg
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
    print("forward rel L2", err)
    assert err <= 1e-2


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


def test_learned_preconditioner_end_to_end():
    """Train on (residual, correction) pairs dumped by plain line-preconditioned solves on
    the 40,000-node mesh, then run the same steps with the learned correction: every solve
    still reaches 1e-8, the fields agree with the plain run to the solver tolerance, and
    the iteration counts of both runs are reported."""
    m = mesh_for("small")
    cfg = small_cfg(lord=4, **LINE)
    plain = nft.Solver(m, nft.make_params(cfg))
    plain.net_init(hidden=128, seed=1, max_batch=8)
    plain.dump_pairs(max_pairs=600, max_stage=300)
    plain.state_init()
    st_dump = plain.run_steps(6)
    npairs = plain.pair_count()
    assert npairs > 100
    l0, l1 = plain.net_train(steps=600, batch=8, lr=2e-3, seed=3)
    print(f"pairs {npairs}, training loss {l0:.4e} -> {l1:.4e}")
    assert l1 < l0
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
        fb = s.set_learned(-1)
        unk = s.state("unk")
        s.close()
        return st, unk, fb

    sp, up, _ = run(False)
    sl, ul, fb = run(True)
    gain = sp.gcr_iters / max(sl.gcr_iters, 1)
    print(f"plain: {sp.gcr_iters} iterations in {sp.gcr_solves} solves, {sp.ms_gcr:.1f} ms; "
          f"learned: {sl.gcr_iters} iterations, {sl.ms_gcr:.1f} ms, fallbacks {fb}; "
          f"iteration ratio {gain:.3f}, time ratio {sp.ms_gcr / max(sl.ms_gcr, 1e-9):.3f}")
    assert sl.gcr_unconverged == 0 and sp.gcr_unconverged == 0
    for k in range(4):
        assert rel_l2(ul[k], up[k]) <= 1e-5
