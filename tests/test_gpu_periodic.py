"""The doubly periodic box (the reference's iper=2) on the GPU, through the C ABI
(nft_init_periodic): against the oracle bit for bit, and against the reference's own
outputs (tiny_per.npz).

Reference: cyclicr/cyclicr2/cyclicr4 (lippsLFroude_implgwopks3D_unstr3.f:4377-4446),
mpdatamper (:3617-3941), the periodic branches of derivative (:4015-4019), bcv
(:1603-1612) and courant (:2049), mpsolver's vol closure (:99-103) and mpdatamper call
(:314-319). Meshes: "periodic" (8 x 7 columns x 5 levels, 280 nodes, the fixture's) and
"periodic2" (40 x 33 columns x 12 levels, 15,840 nodes: 840 coincident sets over several
wavefronts of set-kernel threads).
"""
import os

import numpy as np
import pytest

from common import (diag_close, mesh_checksum, mesh_for, periodic_cfg, periodic_field, seeded_fields,
                    unique_nodes)
from oracle import Oracle
from src_amd import mesh as M
from src_amd import nft

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def solver(m, cfg, **kw):
    return nft.Solver(m, nft.make_params(cfg, **kw))


@pytest.fixture(scope="module", params=[("periodic", 1), ("periodic", 0), ("periodic2", 1), ("periodic2", 0)],
                ids=["periodic-columnar", "periodic-sell64", "periodic2-columnar", "periodic2-sell64"])
def pair(request):
    key, columnar = request.param
    print(f"case {key} columnar={columnar}", flush=True)
    m = mesh_for(key)
    cfg = periodic_cfg()
    s = solver(m, cfg)
    assert s.periodic
    assert s.set_incidence(columnar)
    o = Oracle(m, cfg)
    yield m, cfg, s, o
    s.close()


def test_periodic_setup_fields_bitwise(pair):
    """d1 and the line factors come from the closed vol."""
    m, cfg, s, o = pair
    for k in ("alpha", "d1", "coefw", "line_a", "line_inv", "line_c"):
        assert np.array_equal(s.field(k), o.field(k)), k


def test_periodic_operators_bitwise(pair):
    m, cfg, s, o = pair
    f = seeded_fields(m)
    for idx in (1, 2, 3):
        assert np.array_equal(s.derivative(f["p"], idx), o.derivative(f["p"], idx)), idx
    for impfl in (0, 1):
        assert np.array_equal(s.laplc(f["p"], impfl), o.laplc(f["p"], impfl)), impfl
        a, b = s.prfr(f["p"], f["u"], f["v"], f["w"], impfl), o.prfr(f["p"], f["u"], f["v"], f["w"], impfl)
        for k in range(3):
            assert np.array_equal(a[k], b[k]), ("prfr", impfl, k)
    for flag in (0.0, 1.0):
        a, b = s.bcv(f["u"], f["v"], f["w"], flag), o.bcv(f["u"], f["v"], f["w"], flag)
        for k in range(3):
            assert np.array_equal(a[k], b[k]), ("bcv", flag, k)
    for iflg in (-1, 1):
        assert np.array_equal(s.div(f["u"], f["v"], f["w"], iflg), o.div(f["u"], f["v"], f["w"], iflg)), iflg
    assert np.array_equal(s.precon(f["r"], 1), o.precon(f["r"], 1))


def test_periodic_mpdatam_bitwise(pair):
    m, cfg, s, o = pair
    f = seeded_fields(m)
    for iflag, name in ((1, "u"), (3, "w"), (4, "p")):
        assert np.array_equal(s.mpdatam(f["uuu"], f[name], iflag), o.mpdatam(f["uuu"], f[name], iflag)), iflag


def test_periodic_gcrk_bitwise_device_order(pair):
    """A GCR(3) solve with point relaxation (the reference's preconditioner): the same
    reduction tree as the oracle's dot_order=1, so bit for bit."""
    m, cfg, s, o = pair
    f = seeded_fields(m)
    o1 = Oracle(m, cfg, dot_order=1)
    p0 = np.zeros(m.npoin)
    pg, ug, st, rc = s.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    po, uo = o1.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    assert st.niter == o1.stats().niter
    assert np.array_equal(pg, po)
    for k in range(3):
        assert np.array_equal(ug[k], uo[k]), k
    o1.close()


def test_periodic_against_reference_vectors():
    """The GPU on the fixture's inputs equals the reference's own outputs for every operator
    (fixed-order arithmetic). The GCR solve sums its inner products as a tree here and
    sequentially in the reference, and this one stagnates at its 90-iteration cap (eer
    0.026), which carries the order's rounding into the pressure: 1e-6 relative L2 (measured
    1.6e-8; the GPU equals the oracle's tree order bit for bit, and the oracle's sequential
    order equals the reference)."""
    path = os.path.join(GOLD, "tiny_per.npz")
    if not os.path.exists(path):
        pytest.skip("tiny_per.npz not generated")
    g = np.load(path, allow_pickle=False)
    m = mesh_for("periodic")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    s = solver(m, periodic_cfg())
    for k in ("alpha", "d1", "coefw"):
        assert np.array_equal(s.field(k), g[k]), k
    assert s.field("beta")[0] == g["beta"][0]
    v = g["vec_in"]
    for k in (1, 2, 3):
        assert np.array_equal(s.derivative(v[0], k), g[f"derivative_{k}"]), k
    for k in (0, 1):
        assert np.array_equal(s.laplc(v[0], k), g[f"laplc_{k}"]), k
        assert np.array_equal(np.concatenate(s.bcv(v[1], v[2], v[3], float(k))), g[f"bcv_{k}"]), k
        assert np.array_equal(np.concatenate(s.prfr(v[0], v[1], v[2], v[3], k)), g[f"prfr_{k}"]), k
    assert np.array_equal(s.div(v[1], v[2], v[3], -1), g["div"])
    assert np.array_equal(s.precon(v[4], 1), g["precon"])
    uuu = np.ascontiguousarray(v[5:8])
    assert np.array_equal(s.mpdatam(uuu, v[0], 1), g["mpdatam_1"])
    assert np.array_equal(s.mpdatam(uuu, v[0], 4), g["mpdatam_4"])
    p, u, st, rc = s.gcrk(np.zeros(m.npoin), v[1], v[2], v[3], 1)
    assert st.niter == int(g["gcrk_stats"][0])
    assert np.linalg.norm(p - g["gcrk_p"]) <= 1e-6 * np.linalg.norm(g["gcrk_p"])
    s.close()


def test_periodic_steps_against_reference():
    """10 steps of the program against the reference's 10-step state, from the reference
    build whose solves converge (no preconditioner, relative exit, 49-107 iterations per
    solve). Inner products are tree sums here and sequential in the reference: 1e-10
    relative L2 per field (FP64), the same iteration counts. (In the fixture's capped runs
    the point-relaxation GCR stagnates at its 90-iteration cap with eer near 1, and whether
    a stagnating solve touches the tolerance depends on the summation order; there the GPU
    is held bit for bit to the oracle's tree order, and the oracle's sequential order to the
    reference.)"""
    path = os.path.join(GOLD, "tiny_per.npz")
    if not os.path.exists(path):
        pytest.skip("tiny_per.npz not generated")
    g = np.load(path, allow_pickle=False)
    tag = "conv_"
    m = mesh_for("periodic")
    s = solver(m, periodic_cfg(converged=True))
    s.state_init()
    its = [s.gcr_stats().niter]
    for _ in range(10):
        s.run_steps(1)
        its.append(s.gcr_stats().niter)
    ref_its = g[f"{tag}s10_niter"].tolist()
    print("iterations", its, ref_its)
    assert its == ref_its
    for name in ("unk", "uuu", "prpot"):
        a, b = s.state(name), g[f"{tag}s10_{name}"]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err <= 1e-10, (name, err)
    s.close()


@pytest.mark.parametrize("key,kw", [("periodic", dict()), ("periodic2", dict(precon=0, exit_mode=1, eps0=1e-8, itr=400)),
                                    ("periodic2", dict(precon=2, itrp1=1, omega=1.0))],
                         ids=["periodic-point", "periodic2-none-converging", "periodic2-line"])
def test_periodic_steps_bitwise_device_order(key, kw):
    """Initialisation and 3 steps (mpdatamper x4, the solves, projections, forcing) equal
    the oracle's in dot_order=1 bit for bit, and so do the diagnostics' extrema (sums to
    1e-10)."""
    m = mesh_for(key)
    cfg = periodic_cfg(**kw)
    s, o = solver(m, cfg), Oracle(m, cfg, dot_order=1)
    s.state_init()
    o.init()
    its, ito = [s.gcr_stats().niter], [o.stats().niter]
    for _ in range(3):
        s.run_steps(1)
        o.step()
        its.append(s.gcr_stats().niter)
        ito.append(o.stats().niter)
    print("iterations", its, ito)
    assert its == ito
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    for flag in (1, 2, 3):
        d = s.diagnostics(flag)
        mine = np.concatenate([np.array(d.div), np.array(d.cour), [d.crnic], d.u, d.v, d.w, d.t])
        diag_close(mine, o.diagnostics(flag))
    s.close()
    o.close()


def test_periodic_band_preconditioner_bitwise():
    """precon 3 in the periodic box: the probe colouring reaches across the closures (twin
    columns), and the probed band and its solve equal the oracle's, whose band equals
    laplc's dense matrix (test_oracle_periodic)."""
    m = mesh_for("periodic2")
    cfg = periodic_cfg(precon=3, itrp1=1, omega=1.0)
    s, o = solver(m, cfg), Oracle(m, cfg, dot_order=1)
    for name in ("band_m2", "band_m1", "band_0", "band_p1", "band_p2"):
        assert np.array_equal(s.field(name), o.field(name)), name
    r = seeded_fields(m)["r"]
    assert np.array_equal(s.precon(r, 1), o.precon(r, 1))
    s.close()


def test_periodic_mpdatam_conserves_mass():
    """No boundary flux anywhere: a periodic field (copies equal, nonzero on every side)
    keeps its copies equal, and sum(vol rho psi) over one node per set is unchanged to
    1e-12 of sum(vol rho |psi|)."""
    m = mesh_for("periodic2")
    s = solver(m, periodic_cfg())
    f = seeded_fields(m)
    psi = periodic_field(m, f["p"])
    uuu = np.stack([periodic_field(m, a) for a in f["uuu"]])
    out = s.mpdatam(uuu, psi, 4)
    for sets in (m.coin, m.coin2, m.coin4):
        assert (out[sets - 1] == out[sets[0] - 1]).all()
    keep = unique_nodes(m)
    o = Oracle(m, periodic_cfg())
    vr = o.field("vol")[keep] * s.field("rh0")[keep]
    before, after = (vr * psi[keep]).sum(), (vr * out[keep]).sum()
    assert abs(after - before) <= 1e-12 * (vr * np.abs(psi[keep])).sum()
    assert np.abs(out - psi).max() > 1e-3
    s.close()


def test_periodic_refuses_what_it_does_not_restate():
    m = mesh_for("periodic")
    with pytest.raises(nft.NftError, match="FP64"):
        solver(m, periodic_cfg(), precision="mixed")
    s = solver(m, periodic_cfg())
    with pytest.raises(nft.NftError, match="periodic"):
        s.net_init(hidden=32, seed=1, max_batch=1)
    with pytest.raises(nft.NftError, match="periodic"):
        s.laplc_t(np.zeros(m.npoin), 1)
    s.close()


def test_periodic_init_rejects_bad_sets():
    """nft_init_periodic checks its sets before anything reaches a kernel: a node in two
    sets, a node id out of range."""
    bad = M.generate(M.periodic_spec())     # a mesh of its own (the cached one stays intact)
    coin = bad.coin.copy()
    bad.coin[1, 0] = bad.coin2[0, 0]        # also in a y-side pair
    with pytest.raises(nft.NftError, match="more than one"):
        solver(bad, periodic_cfg())
    bad.coin = coin.copy()
    bad.coin[0, 3] = bad.npoin + 1
    with pytest.raises(nft.NftError, match="out of range"):
        solver(bad, periodic_cfg())
