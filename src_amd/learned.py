"""Training and evaluating the learned preconditioner over several meshes (host side of the
learned path; no reference counterpart).

The network (a column MLP with two horizontal graph convolutions, bf16 MFMA; libnft_hip's
nft_net_*) is trained on (residual, line-preconditioned correction) pairs that the plain
FP64 solver stores during its own GCR solves (nft_dump_pairs). The training set comes from
meshes of one triangulation with seeded hill heights (the 40,000-node mesh: 2,000 jittered
columns x 20 levels): each hill height gives one solver context holding its own pairs, and
the training visits the contexts in turn, carrying the whole training state (weights and
Adam moments, nft_net_state) from one to the next, so every pair is trained against its own
mesh's operator. The trained weights then go into a context of the workload it is applied
to, where the flexible GCR adds the correction and falls back to the line preconditioner
when an iteration stalls (nft_set_learned).
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import mesh as M
from . import nft


def hill_heights(count, seed=2024, lo=250.0, hi=1500.0):
    """Seeded hill heights, uniform in [lo, hi] m."""
    return np.random.default_rng(seed).uniform(lo, hi, size=count)


@dataclasses.dataclass
class PairStore:
    h0: float
    mesh: M.Mesh
    solver: nft.Solver
    pairs: int
    solves: int
    iters: int


def training_flow():
    """The flow and solver the training pairs come from: the 40,000-node workload's flow
    (U = 10 m/s, N = 0.01 1/s, dt = 10 s) with flexible GCR(4) to 1e-8 relative residual and
    the band line preconditioner, in FP64."""
    c = M.CONFIGS["small"]
    return M.RunConfig("training", c.mesh, u0=c.u0, bvf=c.bvf, dt=c.dt, lord=4, eps0=1e-8, exit_mode=1, itr=400,
                       precon=3, itrp1=1, omega=1.0, sponge_depth=3000.0, sponge_tau=300.0)


def workload(name):
    """The learned preconditioner's workloads with the same solver settings: "small" (the
    40,000-node mesh and flow, h0 = 500 m, N = 0.01) and "learned" (the same mesh with a
    1 km hill, N = 0.005)."""
    c = M.CONFIGS[name]
    return M.RunConfig(name, c.mesh, u0=c.u0, bvf=c.bvf, dt=c.dt, lord=4, eps0=1e-8, exit_mode=1, itr=400,
                       precon=3, itrp1=1, omega=1.0, sponge_depth=3000.0, sponge_tau=300.0)


def build_pair_stores(flow: M.RunConfig, heights, pairs_per_mesh, hidden=128, seed=1, max_batch=8,
                      stride=1, max_steps=200, log=None):
    """One FP64 context per hill height (the flow and solver of `flow`, the 40,000-node mesh
    with that hill), with pairs stored from its initial projection and time steps until it
    holds pairs_per_mesh (or max_steps steps have run)."""
    stores = []
    for h0 in heights:
        t0 = time.perf_counter()
        m = M.generate(M.small_hill_spec(h0))
        cfg = dataclasses.replace(flow, mesh=m.spec)
        s = nft.Solver(m, nft.make_params(cfg, precision="fp64"))
        s.net_init(hidden=hidden, seed=seed, max_batch=max_batch)
        s.dump_pairs(pairs_per_mesh, stride)
        s.state_init()
        solves, iters = 1, s.gcr_stats().niter
        steps = 0
        while s.pair_count() < pairs_per_mesh and steps < max_steps:
            st = s.run_steps(1)
            solves += st.gcr_solves
            iters += st.gcr_iters
            steps += 1
        s.dump_pairs(0)
        stores.append(PairStore(float(h0), m, s, s.pair_count(), solves, iters))
        if log:
            log(f"[pairs] h0 {h0:7.1f} m: {s.pair_count()} pairs from {solves} solves ({iters} iterations, "
                f"{steps} steps) in {time.perf_counter() - t0:.1f} s")
    return stores


def train_round_robin(stores, rounds, steps_per_visit, batch=4, lr=1e-3, seed=7, log=None):
    """Adam over the stores' pairs: each round visits every store once (in a seeded order)
    for steps_per_visit steps on its own pairs, the training state moving with it.
    Returns the trained state and the per-round mean of the last minibatch losses."""
    rng = np.random.default_rng(seed)
    state = stores[0].solver.net_state()
    history = []
    for r in range(rounds):
        losses = []
        for k in rng.permutation(len(stores)):
            s = stores[k].solver
            s.net_state(state)
            l0, l1 = s.net_train(steps=steps_per_visit, batch=batch, lr=lr, seed=int(rng.integers(1 << 30)))
            losses.append(l1)
            state = s.net_state()
        history.append(float(np.mean(losses)))
        if log:
            log(f"[train] round {r + 1}/{rounds}: mean last-minibatch loss {history[-1]:.4f}")
    return state, history


def held_out_loss(store, state):
    """Mean training objective over all pairs of a store (weights from `state`)."""
    s = store.solver
    s.net_state(state)
    return s.pair_loss(0, store.pairs)


def run_workload(cfg: M.RunConfig, mesh, steps, precision, weights=None, hidden=128, max_batch=8, stall=0.99,
                 warmup=1):
    """`steps` time steps of cfg after `warmup` untimed ones; with weights, the learned
    correction is on (line-preconditioner fallback on stalls). Returns the run statistics."""
    s = nft.Solver(mesh, nft.make_params(cfg, precision=precision))
    if weights is not None:
        s.net_init(hidden=hidden, seed=1, max_batch=max_batch)
        s.net_params(weights)
    s.state_init()
    if warmup:
        s.run_steps(warmup)
    if weights is not None:
        s.set_learned(1, stall=stall)
    st = s.run_steps(steps)
    fb, li = s.set_learned(-1) if weights is not None else (0, 0)
    unk = s.state("unk")
    s.close()
    return dict(value=round(1000.0 / (st.ms_total / steps), 3), unit="timesteps/s",
                iters_per_solve=round(st.gcr_iters / max(st.gcr_solves, 1), 2),
                solve_ms=round(st.ms_gcr / max(st.gcr_solves, 1), 3), steps=steps,
                unconverged=st.gcr_unconverged, fallbacks=fb, learned_iterations=li,
                precision=precision, finite=bool(np.isfinite(unk).all())), unk


def study(heights=20, pairs_per_mesh=1000, rounds=10, steps_per_visit=50, eval_steps=10, objective=0,
          held_h0=900.0, log=None):
    """The learned preconditioner end to end: pairs from FP64 solves of the training flow on
    `heights` seeded hill heights (250-1500 m), round-robin training, the held-out loss on a
    mesh with a hill of held_h0, then the workloads "small" (mixed precision: the 40,000-node
    mixed-precision workload) and "learned" (1 km hill, N = 0.005) with the band line alone
    (FP64 and mixed) and with the learned correction (mixed). Returns a summary dict."""
    t0 = time.perf_counter()
    flow = training_flow()
    stores = build_pair_stores(flow, hill_heights(heights), pairs_per_mesh, log=log)
    held = build_pair_stores(flow, [held_h0], min(pairs_per_mesh, 300), log=log)[0]
    for st in stores + [held]:
        st.solver.net_options(objective, 0)
    t_pairs = time.perf_counter() - t0
    before = held_out_loss(held, stores[0].solver.net_state())
    t0 = time.perf_counter()
    state, hist = train_round_robin(stores, rounds, steps_per_visit, log=log)
    t_train = time.perf_counter() - t0
    after = held_out_loss(held, state)
    nparams = held.solver.net_info()["nparams"]
    weights = state[:nparams].copy()
    npairs = sum(st.pairs for st in stores)
    for st in stores + [held]:
        st.solver.close()
    out = dict(training=dict(meshes=heights, pairs=npairs, held_out_h0=held_h0, adam_steps=rounds * heights * steps_per_visit,
                             batch=4, objective=["residual", "angle", "error"][objective], pair_seconds=round(t_pairs, 1),
                             train_seconds=round(t_train, 1), held_out_loss=[round(before, 5), round(after, 5)],
                             loss_per_round=[round(h, 5) for h in hist]),
               workloads={})
    for name in ("small", "learned"):
        cfg = workload(name)
        m = M.generate(cfg.mesh)
        runs = {}
        base_unk = None
        for key, prec, w in (("line_fp64", "fp64", None), ("line_mixed", "mixed", None),
                             ("learned_mixed", "mixed", weights)):
            r, unk = run_workload(cfg, m, eval_steps, prec, weights=w)
            if base_unk is None:
                base_unk = unk
            else:
                r["field_rel_l2_vs_line_fp64"] = [float(np.linalg.norm(unk[k] - base_unk[k]) / np.linalg.norm(base_unk[k]))
                                                  for k in range(4)]
            runs[key] = r
            if log:
                log(f"[{name}] {key}: {r}")
        runs["learned_mixed_vs_line_fp64"] = round(runs["learned_mixed"]["value"] / runs["line_fp64"]["value"], 4)
        runs["learned_iteration_gain"] = round(runs["line_mixed"]["iters_per_solve"] /
                                               runs["learned_mixed"]["iters_per_solve"], 4)
        out["workloads"][name] = runs
    return out
