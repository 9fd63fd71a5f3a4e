"""Multigrid branch of precon on the GPU (mg_flag=1, lippsLFroude_implgwopks3D_unstr3.f
:618-748) through the C ABI against the CPU oracle, whose restatement equals the flang
reference bit for bit (tests/test_oracle_mg.py).

Bar: bit-exact for the transfers, the coarse laplc, the set-up values and precon (all
fixed-order gathers and pointwise loops); GCR solves and time steps bit-exact against the
oracle in the GPU's inner-product order (dot_order=1), iteration counts included."""
import numpy as np
import pytest

from common import MG_SWEEPS, mesh_for, mg_levels_for, seeded_fields, small_cfg, tiny_cfg
from oracle import Oracle
from src_amd import nft

pytestmark = pytest.mark.gpu

PARITY = dict(precon=1, itrp1=2, omega=0.5)
LINE_MG = dict(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
HEAVY = dict(itrp2=1, itrp_c1=4, itrp_c2=4, itrp_cc=8)
# the line preconditioner's omega = 1 makes the coarse point Jacobi diverge; relax them at 0.5
HEAVY_C = dict(HEAVY, omega_c=0.5)


def cfg_for(key, **kw):
    return (tiny_cfg if key.startswith("tiny") else small_cfg)(**kw)


def pair(key, kw, transfer, sweeps=MG_SWEEPS, as_read=True, dot_order=0):
    m = mesh_for(key)
    levels = mg_levels_for(key, as_read=as_read)
    cfg = cfg_for(key, **kw)
    s = nft.Solver(m, nft.make_params(cfg))
    s.set_multigrid(levels, transfer=transfer, **sweeps)
    o = Oracle(m, cfg, dot_order=dot_order)
    o.set_multigrid(levels, transfer=transfer, **sweeps)
    return m, levels, s, o


@pytest.mark.parametrize("key", ["tiny", "tiny2", "small"])
@pytest.mark.parametrize("transfer", [0, 1])
def test_pieces_bitwise(key, transfer):
    print(f"case {key} transfer={transfer}", flush=True)
    m, levels, s, o = pair(key, PARITY, transfer)
    n1, n2 = levels[0][0].npoin, levels[1][0].npoin
    for lv in (1, 2):
        assert np.array_equal(s.multigrid_field(lv, "d"), o.level_field(lv, "d1")), lv
        assert s.multigrid_field(lv, "beta")[0] == o.level_beta(lv), lv
    f = seeded_fields(m)
    assert np.array_equal(s.multigrid_op("restrict", 1, f["p"]), o.restrict(1, f["p"]))
    assert np.array_equal(s.multigrid_op("restrict", 2, f["u"][:n1]), o.restrict(2, f["u"][:n1]))
    c1 = seeded_fields(levels[0][0], seed=5)["p"]
    c2 = seeded_fields(levels[1][0], seed=6)["p"]
    assert np.array_equal(s.multigrid_op("prolong", 1, c1), o.prolong(1, c1))
    assert np.array_equal(s.multigrid_op("prolong", 2, c2), o.prolong(2, c2))
    for impfl in (0, 1):
        assert np.array_equal(s.multigrid_op("laplc", 1, c1, impfl), o.level_laplc(1, c1, impfl))
        assert np.array_equal(s.multigrid_op("laplc", 2, c2, impfl), o.level_laplc(2, c2, impfl))
    s.close()


@pytest.mark.parametrize("key", ["tiny", "small"])
@pytest.mark.parametrize("kw,transfer,sweeps", [(PARITY, 0, MG_SWEEPS), (PARITY, 1, HEAVY),
                                                (LINE_MG, 1, HEAVY_C), (LINE_MG, 0, MG_SWEEPS),
                                                (dict(LINE_MG, precon=3), 1, HEAVY_C)],
                         ids=["point-ref", "point-nbr", "line-nbr-omega_c", "line-ref", "band-nbr-omega_c"])
def test_precon_bitwise(key, kw, transfer, sweeps):
    print(f"case {key} {kw} transfer={transfer}", flush=True)
    m, levels, s, o = pair(key, kw, transfer, sweeps)
    f = seeded_fields(m)
    assert np.array_equal(s.precon(f["r"], 1), o.precon(f["r"], 1))
    s.close()


@pytest.mark.parametrize("key,kw,transfer,sweeps", [
    ("tiny", PARITY, 0, MG_SWEEPS),
    ("small", dict(PARITY, exit_mode=1, eps0=1e-8, itr=200), 0, MG_SWEEPS),
    ("small", LINE_MG, 1, HEAVY),
    ("small", LINE_MG, 1, HEAVY_C),
])
def test_gcrk_bitwise_device_order(key, kw, transfer, sweeps):
    print(f"case {key} {kw} transfer={transfer}", flush=True)
    m, levels, s, o = pair(key, kw, transfer, sweeps, dot_order=1)
    f = seeded_fields(m)
    p0 = np.zeros(m.npoin)
    pg, ug, st, rc = s.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    po, uo = o.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    so = o.stats()
    print(f"niter gpu {st.niter} oracle {so.niter}")
    assert st.niter == so.niter
    assert np.array_equal(pg, po)
    for k in range(3):
        assert np.array_equal(ug[k], uo[k])
    s.close()


@pytest.mark.parametrize("key,kw,transfer,sweeps,nsteps", [
    ("tiny", dict(), 0, MG_SWEEPS, 10),
    ("small", LINE_MG, 1, HEAVY_C, 3),
])
def test_steps_bitwise_device_order(key, kw, transfer, sweeps, nsteps):
    print(f"case {key} {kw} transfer={transfer} steps={nsteps}", flush=True)
    m, levels, s, o = pair(key, kw, transfer, sweeps, dot_order=1)
    s.state_init()
    o.init()
    its, ito = [s.gcr_stats().niter], [o.stats().niter]
    for _ in range(nsteps):
        s.run_steps(1)
        o.step()
        its.append(s.gcr_stats().niter)
        ito.append(o.stats().niter)
    print("iterations", its)
    assert its == ito
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    s.close()


def test_switching_off_restores_the_fine_smoother():
    m = mesh_for("tiny")
    cfg = tiny_cfg(**PARITY)
    s = nft.Solver(m, nft.make_params(cfg))
    f = seeded_fields(m)
    plain = s.precon(f["r"], 1)
    s.set_multigrid(mg_levels_for("tiny"), transfer=1, **MG_SWEEPS)
    with_mg = s.precon(f["r"], 1)
    assert not np.array_equal(plain, with_mg)
    s.set_multigrid(None, mg_flag=0)
    assert np.array_equal(s.precon(f["r"], 1), plain)
    s.set_multigrid(None, mg_flag=1, **MG_SWEEPS, transfer=1)
    assert np.array_equal(s.precon(f["r"], 1), with_mg)
    s.close()


def test_bad_levels_are_rejected():
    m = mesh_for("tiny")
    s = nft.Solver(m, nft.make_params(tiny_cfg(**PARITY)))
    (m1, t1), (m2, t2) = mg_levels_for("tiny")
    bad = t1.copy(); bad[0] = m.npoin + 1
    with pytest.raises(nft.NftError, match="match entry out of range"):
        s.set_multigrid([(m1, bad), (m2, t2)])
    dup = t1.copy(); dup[1] = dup[0]
    with pytest.raises(nft.NftError, match="two coarse nodes match"):
        s.set_multigrid([(m1, dup), (m2, t2)])
    with pytest.raises(nft.NftError, match="no coarse levels set"):
        s.set_multigrid(None, mg_flag=1)
    s.close()
