"""Full-size parity: the production mesh (65,536 columns x 64 levels, 4,194,304 nodes) and
the stress mesh (131,072 columns x 128 stretched levels, 16,777,216 nodes), at the sizes the
benchmark runs.

Bit for bit against the oracle in the GPU's summation order (dot_order=1): every operator,
MPDATA, the probed band and its solve, a GCR(4) solve (iterations capped so the oracle
finishes in seconds; the capped trajectory is as deterministic as a converged one) and a
full NFT step. At these sizes the grid-level sums run with 4,096 vector blocks and ~16,000
gather blocks per reduction, far beyond what the smaller meshes reach. Size-independent
properties complete it: MPDATA conserves the mass of a field that is zero on the boundary,
and full steps with uncapped solves converge with finite state.

The production mesh made a doubly periodic box (iper=2; 32,512 coincident pairs and 64
corner quadruples) is held to the oracle the same way: operators, mpdatamper, a capped GCR
solve and a step, with the set kernels running over ~130 blocks.

Reference loops: mpsolver DO ITIME (lippsLFroude_implgwopks3D_unstr3.f:285-463), gcrk
(:1776-1972), laplc (:1288-1353), mpdatam (:3029-3311), derivative (:3943-4022); the
periodic branches: cyclicr* (:4377-4446), mpdatamper (:3617-3941).
"""
import dataclasses
import time

import numpy as np
import pytest

from common import diag_close, interior_blob, production_cfg, seeded_fields, stress_cfg
from oracle import Oracle
from src_amd import mesh as M
from src_amd import nft

pytestmark = pytest.mark.gpu

CAPPED_ITR = 3   # GCR(4) restart cycles of the capped solves: 12 iterations per solve


def _setup(cfg, what):
    t0 = time.perf_counter()
    m = M.generate(cfg.mesh)
    print(f"[{what}] mesh {m.npoin} nodes, {m.nedge} edges ({time.perf_counter() - t0:.1f} s)", flush=True)
    t0 = time.perf_counter()
    s = nft.Solver(m, nft.make_params(cfg))
    print(f"[{what}] GPU context ({time.perf_counter() - t0:.1f} s)", flush=True)
    t0 = time.perf_counter()
    o = Oracle(m, cfg, dot_order=1)
    print(f"[{what}] oracle ({time.perf_counter() - t0:.1f} s)", flush=True)
    return m, s, o


def _mass_conserved(m, s, o, iflag=4):
    """MPDATA of a field that is zero at every boundary node: no boundary flux (the inflow
    value of theta' is 0, slip boundaries carry none), interior and semi-edge fluxes cancel
    in pairs, so sum(vol rho psi) is unchanged up to rounding."""
    psi = interior_blob(m)
    uuu = seeded_fields(m)["uuu"]
    out = s.mpdatam(uuu, psi, iflag)
    assert np.array_equal(out, o.mpdatam(uuu, psi, iflag))
    vr = m.vol * s.field("rh0")
    before, after = (vr * psi).sum(), (vr * out).sum()
    scale = (vr * np.abs(psi)).sum()
    print(f"mass before {before:.15e} after {after:.15e}, relative change {(after - before) / scale:.2e}")
    assert abs(after - before) <= 1e-12 * scale
    assert np.abs(out - psi).max() > 1e-6   # the field did move


# ---- production mesh -------------------------------------------------------------------

@pytest.fixture(scope="module")
def prod():
    cfg = production_cfg(itr=CAPPED_ITR)
    m, s, o = _setup(cfg, "production")
    yield m, cfg, s, o
    s.close()
    o.close()


def test_production_operators_bitwise(prod):
    m, cfg, s, o = prod
    f = seeded_fields(m)
    try:
        for columnar in (1, 0):
            assert s.set_incidence(columnar)
            for idx in (1, 2, 3):
                assert np.array_equal(s.derivative(f["p"], idx), o.derivative(f["p"], idx)), (columnar, idx)
            for impfl in (0, 1):
                assert np.array_equal(s.laplc(f["p"], impfl), o.laplc(f["p"], impfl)), (columnar, impfl)
            a = s.prfr(f["p"], f["u"], f["v"], f["w"], 1)
            b = o.prfr(f["p"], f["u"], f["v"], f["w"], 1)
            for k in range(3):
                assert np.array_equal(a[k], b[k]), (columnar, k)
            assert np.array_equal(s.div(f["u"], f["v"], f["w"], -1), o.div(f["u"], f["v"], f["w"], -1))
    finally:
        s.set_incidence(1)


def test_production_mpdatam_bitwise(prod):
    m, cfg, s, o = prod
    f = seeded_fields(m)
    for iflag, name in ((1, "u"), (2, "v"), (3, "w"), (4, "p")):
        assert np.array_equal(s.mpdatam(f["uuu"], f[name], iflag), o.mpdatam(f["uuu"], f[name], iflag)), iflag


def test_production_mpdatam_conserves_mass(prod):
    m, cfg, s, o = prod
    _mass_conserved(m, s, o)


def test_production_band_bitwise(prod):
    """The in-column band of laplc probed on the GPU (12 column colours at this size) and
    its banded LU solve equal the oracle's."""
    m, cfg, s, o = prod
    for name in ("band_m2", "band_m1", "band_0", "band_p1", "band_p2"):
        assert np.array_equal(s.field(name), o.field(name)), name
    r = seeded_fields(m)["r"]
    assert np.array_equal(s.precon(r, 1), o.precon(r, 1))


def test_production_gcr_bitwise(prod):
    """One GCR(4) solve (12 iterations): pressure, velocities, iteration count and residual
    ratios equal the oracle's; every fused dot and grid-level sum of the solve is in it."""
    m, cfg, s, o = prod
    f = seeded_fields(m)
    p0 = np.zeros(m.npoin)
    pg, ug, st, rc = s.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    po, uo = o.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    so = o.stats()
    print(f"niter {st.niter} / {so.niter}, eer {st.eer:.6e} / {so.eer:.6e}", flush=True)
    assert rc == 1 and st.niter == so.niter == 4 * CAPPED_ITR
    assert np.array_equal(pg, po)
    for k in range(3):
        assert np.array_equal(ug[k], uo[k])
    assert st.eer == so.eer and st.eem == so.eem


def test_production_step_bitwise(prod):
    """The potential-flow initialisation and one full NFT step (MPDATA x4, two capped GCR
    solves, projections, forcing, extrapolation) equal the oracle's bit for bit; so do the
    per-step diagnostics' extrema, and their sums to 1e-10."""
    m, cfg, s, o = prod
    s.state_init()
    its = [s.gcr_stats().niter]
    s.run_steps(1)
    its.append(s.gcr_stats().niter)
    o.init()
    ito = [o.stats().niter]
    o.step(1)
    ito.append(o.stats().niter)
    print("iterations", its, ito, flush=True)
    assert its == ito
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    for flag in (1, 2, 3):
        d = s.diagnostics(flag)
        mine = np.concatenate([np.array(d.div), np.array(d.cour), [d.crnic], d.u, d.v, d.w, d.t])
        diag_close(mine, o.diagnostics(flag))


def test_production_steps_converge():
    """Two steps of the production configuration itself (uncapped GCR(4) to 1e-8 relative):
    every solve converges and the state stays finite."""
    cfg = production_cfg()
    m = M.generate(cfg.mesh)
    s = nft.Solver(m, nft.make_params(cfg))
    s.state_init()
    st = s.run_steps(2)
    print(f"2 steps: {st.gcr_solves} solves, {st.gcr_iters} iterations, unconverged {st.gcr_unconverged}")
    assert st.gcr_solves == 4 and st.gcr_unconverged == 0
    for name in ("unk", "uuu", "prpot"):
        assert np.isfinite(s.state(name)).all(), name
    s.close()


# ---- stress mesh -----------------------------------------------------------------------

@pytest.fixture(scope="module")
def stress():
    cfg = stress_cfg(itr=CAPPED_ITR)
    m, s, o = _setup(cfg, "stress")
    yield m, cfg, s, o
    s.close()
    o.close()


def test_stress_laplc_bitwise(stress):
    m, cfg, s, o = stress
    p = seeded_fields(m)["p"]
    try:
        for columnar in (1, 0):
            assert s.set_incidence(columnar)
            assert np.array_equal(s.laplc(p, 1), o.laplc(p, 1)), columnar
        assert np.array_equal(s.derivative(p, 3), o.derivative(p, 3))
    finally:
        s.set_incidence(1)


def test_stress_mpdatam_bitwise(stress):
    m, cfg, s, o = stress
    f = seeded_fields(m)
    for iflag, name in ((1, "u"), (4, "p")):
        assert np.array_equal(s.mpdatam(f["uuu"], f[name], iflag), o.mpdatam(f["uuu"], f[name], iflag)), iflag


def test_stress_band_bitwise(stress):
    """128 stretched levels: the probed band and its solve (columns staged 32 per block in
    LDS) equal the oracle's."""
    m, cfg, s, o = stress
    for name in ("band_m2", "band_m1", "band_0", "band_p1", "band_p2"):
        assert np.array_equal(s.field(name), o.field(name)), name
    r = seeded_fields(m)["r"]
    assert np.array_equal(s.precon(r, 1), o.precon(r, 1))


def test_stress_mpdatam_conserves_mass(stress):
    m, cfg, s, o = stress
    _mass_conserved(m, s, o)


def test_stress_steps_converge():
    """Two steps of the stress configuration (Fr = 0.1, uncapped GCR(4) to 1e-8): every
    solve converges and the state stays finite."""
    cfg = stress_cfg()
    m = M.generate(cfg.mesh)
    s = nft.Solver(m, nft.make_params(cfg))
    s.state_init()
    st = s.run_steps(2)
    print(f"2 steps: {st.gcr_solves} solves, {st.gcr_iters} iterations, unconverged {st.gcr_unconverged}")
    assert st.gcr_solves == 4 and st.gcr_unconverged == 0
    for name in ("unk", "uuu", "prpot"):
        assert np.isfinite(s.state(name)).all(), name
    s.close()


# ---- production mesh as a periodic box ------------------------------------------------

@pytest.fixture(scope="module")
def prod_periodic():
    base = production_cfg(itr=CAPPED_ITR, precon=0)
    cfg = dataclasses.replace(base, mesh=dataclasses.replace(base.mesh, periodic=True))
    m, s, o = _setup(cfg, "production periodic")
    yield m, cfg, s, o
    s.close()
    o.close()


def test_production_periodic_operators_bitwise(prod_periodic):
    m, cfg, s, o = prod_periodic
    assert s.periodic and m.coin.shape[1] + m.coin2.shape[1] + m.coin4.shape[1] == 32576
    f = seeded_fields(m)
    for idx in (1, 2, 3):
        assert np.array_equal(s.derivative(f["p"], idx), o.derivative(f["p"], idx)), idx
    for impfl in (0, 1):
        assert np.array_equal(s.laplc(f["p"], impfl), o.laplc(f["p"], impfl)), impfl
    assert np.array_equal(s.div(f["u"], f["v"], f["w"], -1), o.div(f["u"], f["v"], f["w"], -1))
    for iflag, name in ((1, "u"), (4, "p")):
        assert np.array_equal(s.mpdatam(f["uuu"], f[name], iflag), o.mpdatam(f["uuu"], f[name], iflag)), iflag


def test_production_periodic_gcr_and_step_bitwise(prod_periodic):
    """A GCR(4) solve without preconditioner (capped at 12 iterations), then the
    initialisation and one full step (mpdatamper x4, two capped solves)."""
    m, cfg, s, o = prod_periodic
    f = seeded_fields(m)
    p0 = np.zeros(m.npoin)
    pg, ug, st, rc = s.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    po, uo = o.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    assert st.niter == o.stats().niter == 4 * CAPPED_ITR
    assert np.array_equal(pg, po)
    s.state_init()
    s.run_steps(1)
    o.init()
    o.step(1)
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
