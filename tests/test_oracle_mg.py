"""The oracle's multigrid branch of precon (mg_flag=1, lippsLFroude_implgwopks3D_unstr3.f
:618-748) against the flang-compiled reference run on generated nested meshes
(tests/golden/tiny_mg.npz), and the corrected transfers (transfer=1, no reference
counterpart) against their definitions written independently with numpy."""
import os

import numpy as np
import pytest

from common import MG_SWEEPS, mesh_checksum, mesh_for, mg_levels_for, seeded_fields, tiny_cfg
from oracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PARITY = dict(precon=1, sm_flag=1, itrp1=2, omega=0.5)


def load(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="module")
def gold():
    g = load("tiny_mg.npz")
    m = mesh_for("tiny")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    levels = mg_levels_for("tiny")
    for k, (lm, t) in enumerate(levels):
        assert np.array_equal(g["level_checksums"][k], mesh_checksum(lm)), f"level {k + 1} mesh differs"
        assert np.array_equal(g[f"match{k + 1}"], t)
    sw = g["sweeps"].tolist()
    assert sw == [MG_SWEEPS[k] for k in ("itrp2", "itrp_c1", "itrp_c2", "itrp_cc")]
    return g, m


def seeded(m):
    # the vectors of make_golden.seeded_vectors (seed 7)
    rng = np.random.default_rng(7)
    n = m.npoin
    x, y, z = m.coord
    v = np.zeros((8, n))
    v[0] = 50.0 * np.sin(x / 700.0) * np.cos(y / 900.0) + 10.0 * z / 1000.0 + rng.normal(0, 1.0, n)
    v[1] = 10.0 + rng.normal(0, 1.0, n)
    v[2] = rng.normal(0, 1.0, n)
    v[3] = rng.normal(0, 0.5, n)
    v[4] = rng.normal(0, 1e-3, n)
    v[5] = 10.0 + rng.normal(0, 0.5, n)
    v[6] = rng.normal(0, 0.5, n)
    v[7] = rng.normal(0, 0.2, n)
    return v


def mg_oracle(m, levels, transfer=0, **kw):
    o = Oracle(m, tiny_cfg(), **PARITY, **kw)
    o.set_multigrid(levels, transfer=transfer, **MG_SWEEPS)
    return o


def test_pieces_bitwise(gold):
    g, m = gold
    v = seeded(m)
    levels = mg_levels_for("tiny", as_read=True)
    n1, n2 = levels[0][0].npoin, levels[1][0].npoin
    o = mg_oracle(m, levels)
    assert np.array_equal(o.level_field(1, "d1"), g["d2"])
    assert np.array_equal(o.level_field(2, "d1"), g["d3"])
    assert [o.level_beta(1), o.level_beta(2)] == g["beta23"].tolist()
    assert np.array_equal(o.restrict(1, v[0]), g["restrict_1"])
    assert np.array_equal(o.restrict(2, v[1][:n1]), g["restrict_2"])
    assert np.array_equal(o.prolong(1, v[2][:n1]), g["prolong_1"])
    assert np.array_equal(o.prolong(2, v[3][:n2]), g["prolong_2"])
    assert np.array_equal(o.level_laplc(1, v[0][:n1], 1), g["laplc_c_1"])
    assert np.array_equal(o.level_laplc(1, v[0][:n1], 0), g["laplc_c_0"])
    assert np.array_equal(o.level_laplc(2, v[0][:n2], 1), g["laplc_cc_1"])
    assert np.array_equal(o.precon(v[4], 1), g["precon"])
    p, outs = o.gcrk(np.zeros(m.npoin), v[1], v[2], v[3], 1)
    assert np.array_equal(p, g["gcrk_p"])
    assert np.array_equal(np.concatenate(outs), g["gcrk_u"])
    st = o.stats()
    assert [st.niter, st.eer, st.eem] == g["gcrk_stats"].tolist()


def test_time_steps_bitwise(gold):
    g, m = gold
    levels = mg_levels_for("tiny", as_read=True)
    for nt in (1, 10):
        o = mg_oracle(m, levels, exit_mode=0)
        o.init()
        its = [o.stats().niter]
        for _ in range(nt):
            o.step()
            its.append(o.stats().niter)
        for k in ("unk", "frc", "uuu", "unkdash", "prpot"):
            assert np.array_equal(o.state(k), g[f"s{nt}_{k}"]), (nt, k)
        assert its == g[f"s{nt}_niter"].tolist()


def test_reader_quirk_and_index_quirk_are_what_the_fixture_pins(gold):
    """The fixture distinguishes the plausible alternatives: the second coarse mesh as
    written (not as the reference reads it) and the neighbour-list transfers both change
    the results."""
    g, m = gold
    v = seeded(m)
    n2 = mg_levels_for("tiny")[1][0].npoin
    o = mg_oracle(m, mg_levels_for("tiny"))
    assert not np.array_equal(o.level_laplc(2, v[0][:n2], 1), g["laplc_cc_1"])
    o = mg_oracle(m, mg_levels_for("tiny", as_read=True), transfer=1)
    assert not np.array_equal(o.restrict(1, v[0]), g["restrict_1"])
    assert not np.array_equal(o.prolong(1, v[2][:mg_levels_for("tiny")[0][0].npoin]), g["prolong_1"])


def neighbours(mesh):
    """Neighbour lists in neigh_definition's order (interior edges, then boundary-face
    edges), built with numpy."""
    n = mesh.npoin
    nb = [[] for _ in range(n)]
    for arr in (mesh.iedge, mesh.kbface):
        for a, b in zip(arr[0] - 1, arr[1] - 1):
            nb[a].append(b)
            nb[b].append(a)
    return nb


@pytest.mark.parametrize("key", ["tiny", "tiny2"])
def test_corrected_transfers_match_their_definition(key):
    m = mesh_for(key)
    levels = mg_levels_for(key)
    (m1, t1), (m2, t2) = levels
    o = Oracle(m, tiny_cfg(), **PARITY)
    o.set_multigrid(levels, transfer=1, **MG_SWEEPS)
    f = seeded_fields(m)["p"]
    nb = neighbours(m)
    # restriction: the mean over the neighbours of the edge midpoint values
    want = np.array([0.5 * f[ip] + 0.5 * np.mean(f[nb[ip]]) for ip in t1 - 1])
    np.testing.assert_allclose(o.restrict(1, f), want, rtol=1e-13, atol=1e-12)
    # prolongation: injection, then the mean of the matched neighbours (0 if none)
    coarse = seeded_fields(m1, seed=3)["p"]
    matched = np.full(m.npoin, -1)
    matched[t1 - 1] = np.arange(m1.npoin)
    want = np.zeros(m.npoin)
    for ip in range(m.npoin):
        if matched[ip] >= 0:
            want[ip] = coarse[matched[ip]]
        else:
            ks = [matched[j] for j in nb[ip] if matched[j] >= 0]
            want[ip] = np.mean(coarse[ks]) if ks else 0.0
    got = o.prolong(1, coarse)
    assert np.array_equal(got[t1 - 1], coarse)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-12)
    # constants pass through both transfers (1.0 sums exactly)
    assert np.array_equal(o.restrict(2, np.ones(m1.npoin)), np.ones(m2.npoin))
    one = o.prolong(2, np.ones(m2.npoin))
    assert set(np.unique(one).tolist()) <= {0.0, 1.0}


def test_corrected_cycle_reduces_iterations_on_small_mesh():
    """With the neighbour-list transfers the V-cycle is a working coarse-grid correction:
    GCR with it needs fewer iterations than with the fine smoother alone."""
    from common import small_cfg
    m = mesh_for("small")
    cfg = small_cfg(exit_mode=1, eps0=1e-6, itr=200)
    f = seeded_fields(m)
    base = Oracle(m, cfg, precon=1, sm_flag=1, itrp1=2, omega=0.5)
    base.gcrk(np.zeros(m.npoin), f["u"], f["v"], f["w"], 1)
    n0 = base.stats().niter
    o = Oracle(m, cfg, precon=1, sm_flag=1, itrp1=2, omega=0.5)
    o.set_multigrid(mg_levels_for("small"), transfer=1, itrp2=2, itrp_c1=4, itrp_c2=4, itrp_cc=8)
    o.gcrk(np.zeros(m.npoin), f["u"], f["v"], f["w"], 1)
    n1 = o.stats().niter
    assert n1 < n0, (n1, n0)
