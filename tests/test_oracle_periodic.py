"""The oracle's doubly periodic box (the reference's iper=2) against the reference itself:
tiny_per.npz holds the reference's own outputs (tests/golden/make_golden.py --periodic) on
the 280-node periodic mesh. Also the properties the periodic closures fix: coincident nodes
hold one value, and MPDATA conserves mass with no boundary flux at all.

Reference: cyclicr/cyclicr2/cyclicr4 (lippsLFroude_implgwopks3D_unstr3.f:4377-4446),
mpdatamper (:3617-3941), derivative's periodic branch (:4015-4019), bcv's (:1603-1612),
courant's (:2049), mpsolver's vol closure (:99-103) and mpdatamper call (:314-319)."""
import os

import numpy as np
import pytest

from common import mesh_checksum, mesh_for, periodic_cfg, periodic_field, seeded_fields, unique_nodes
from oracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def per():
    path = os.path.join(GOLD, "tiny_per.npz")
    if not os.path.exists(path):
        pytest.skip("tiny_per.npz not generated")
    g = np.load(path, allow_pickle=False)
    m = mesh_for("periodic")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m)), "regenerated mesh differs from the fixture's"
    assert g["coin_checksum"].tolist() == [int(a.sum()) for a in (m.coin, m.coin2, m.coin4)]
    return g, m


def test_periodic_setup_fields_bitwise(per):
    """alpha, d1 (compute_d over the closed vol), coefw and beta (its min over vol/h with the
    closed vol)."""
    g, m = per
    o = Oracle(m, periodic_cfg())
    for k in ("alpha", "d1", "coefw"):
        assert np.array_equal(o.field(k), g[k]), k
    assert o.beta == g["beta"][0]


def test_periodic_operators_bitwise(per):
    g, m = per
    o = Oracle(m, periodic_cfg())
    v = g["vec_in"]
    for k in (1, 2, 3):
        assert np.array_equal(o.derivative(v[0], k), g[f"derivative_{k}"]), k
    for k in (0, 1):
        assert np.array_equal(o.laplc(v[0], k), g[f"laplc_{k}"]), k
        assert np.array_equal(np.concatenate(o.bcv(v[1], v[2], v[3], k)), g[f"bcv_{k}"]), k
        assert np.array_equal(np.concatenate(o.prfr(v[0], v[1], v[2], v[3], k)), g[f"prfr_{k}"]), k
    assert np.array_equal(o.div(v[1], v[2], v[3], -1), g["div"])
    assert np.array_equal(o.precon(v[4], 1), g["precon"])
    uuu = np.ascontiguousarray(v[5:8])
    # the harness calls mpdatamper for u (iflag 1) and theta' (iflag 4)
    assert np.array_equal(o.mpdatam(uuu, v[0], 1), g["mpdatam_1"])
    assert np.array_equal(o.mpdatam(uuu, v[0], 4), g["mpdatam_4"])
    p, outs = o.gcrk(np.zeros(m.npoin), v[1], v[2], v[3], 1)
    assert np.array_equal(p, g["gcrk_p"])
    assert np.array_equal(np.concatenate(outs), g["gcrk_u"])
    st = o.stats()
    assert [st.niter, st.eer, st.eem] == g["gcrk_stats"].tolist()


@pytest.mark.parametrize("tag,converged", [("", False), ("conv_", True)], ids=["parity-build", "converging-build"])
def test_periodic_steps_bitwise(per, tag, converged):
    """1 and 10 steps of the program: the build of the other fixtures (its GCR at its
    90-iteration cap) and the converging build (no preconditioner, relative exit)."""
    g, m = per
    for nt in (1, 10):
        o = Oracle(m, periodic_cfg(converged=converged))
        o.init()
        its = [o.stats().niter]
        for _ in range(nt):
            o.step()
            its.append(o.stats().niter)
        assert its == g[f"{tag}s{nt}_niter"].tolist(), (nt, its)
        for k in ("unk", "frc", "uuu", "unkdash", "prpot"):
            assert np.array_equal(o.state(k), g[f"{tag}s{nt}_{k}"]), (nt, k)


def test_periodic_closures_make_copies_equal():
    """Every output of derivative is summed over the coincident sets, so the nodes of a set
    hold one value whatever the input; a closed sum of a pair equals the sum of the two
    open (closed-box form without boundary terms) values."""
    m = mesh_for("periodic2")
    o = Oracle(m, periodic_cfg())
    f = seeded_fields(m)
    for k in (1, 2, 3):
        d = o.derivative(f["p"], k)
        for s in (m.coin, m.coin2, m.coin4):
            assert (d[s - 1] == d[s[0] - 1]).all(), k
    # against the raw edge sums
    raw = np.zeros(m.npoin)
    i1, i2 = m.iedge[0] - 1, m.iedge[1] - 1
    u = (f["p"][i1] + f["p"][i2]) * 0.5
    np.add.at(raw, i1, m.sn[0] * u)
    np.add.at(raw, i2, -m.sn[0] * u)
    j1, j2 = m.kbface[0] - 1, m.kbface[1] - 1
    u = (f["p"][j1] + f["p"][j2]) * 0.5
    np.add.at(raw, j1, m.sb[0] * u)
    np.add.at(raw, j2, -m.sb[0] * u)
    d = o.derivative(f["p"], 1)
    pair = m.coin[:, 0] - 1
    assert d[pair[0]] == pytest.approx(raw[pair[0]] + raw[pair[1]], rel=1e-12)
    quad = m.coin4[:, 0] - 1
    assert d[quad[0]] == pytest.approx(raw[quad].sum(), rel=1e-12)
    inner = unique_nodes(m)
    for s in (m.coin, m.coin2, m.coin4):
        inner[s - 1] = False
    assert np.abs(d[inner] - raw[inner]).max() <= 1e-9 * np.abs(raw).max()


def test_periodic_mpdatam_conserves_mass_of_any_periodic_field():
    """No boundary flux anywhere in the periodic box: a field whose coincident copies agree
    (nonzero on every side) keeps its copies bitwise equal through mpdatamper, and
    sum(vol rho psi) over one node per set is unchanged up to rounding."""
    m = mesh_for("periodic2")
    o = Oracle(m, periodic_cfg())
    f = seeded_fields(m)
    psi = periodic_field(m, f["p"])
    uuu = np.ascontiguousarray(np.stack([periodic_field(m, a) for a in f["uuu"]]))
    out = o.mpdatam(uuu, psi, 4)
    for s in (m.coin, m.coin2, m.coin4):
        assert (out[s - 1] == out[s[0] - 1]).all()
    keep = unique_nodes(m)
    vr = o.field("vol")[keep] * o.field("rh0")[keep]
    before, after = (vr * psi[keep]).sum(), (vr * out[keep]).sum()
    scale = (vr * np.abs(psi[keep])).sum()
    assert abs(after - before) <= 1e-12 * scale, (before, after)
    assert np.abs(out - psi).max() > 1e-3   # the field did move


def test_periodic_fp32_forms_are_refused():
    m = mesh_for("periodic")
    with pytest.raises(ValueError):
        Oracle(m, periodic_cfg(), precision=1)
