"""HIP path (through the C ABI) against the CPU oracle and the reference's golden vectors.

Bar: bit-exact where the arithmetic is a fixed-order restatement (gathers in the
reference's scatter order, pointwise loops, Thomas sweeps, MPDATA); GCR inner products
are tree reductions, so solver results are compared with tolerances that are stated in
each test (FP64: 1e-12 relative L2 per field after the time steps).
"""
import os

import numpy as np
import pytest

from common import diag_close as _diag_close, mesh_checksum, mesh_for, rel_l2, seeded_fields, small_cfg, tiny_cfg
from oracle import Oracle
from src_amd import mesh as M
from src_amd import nft

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def solver(m, cfg, **kw):
    return nft.Solver(m, nft.make_params(cfg, **kw))


@pytest.fixture(scope="module", params=[("tiny", 1), ("tiny", 0), ("small", 1), ("small", 0)],
                ids=["tiny-columnar", "tiny-sell64", "small-columnar", "small-sell64"])
def pair(request):
    key, columnar = request.param
    m = mesh_for(key)
    cfg = tiny_cfg() if key == "tiny" else small_cfg()
    s = solver(m, cfg)
    assert s.set_incidence(columnar), "generated meshes are extruded: the columnar form must exist"
    o = Oracle(m, cfg)
    yield m, cfg, s, o
    s.close()


def test_setup_fields_bitwise(pair):
    m, cfg, s, o = pair
    for k in ("alpha", "d1", "coefw", "line_a", "line_inv", "line_c"):
        assert np.array_equal(s.field(k), o.field(k)), k


def test_derivative_bitwise(pair):
    m, cfg, s, o = pair
    f = seeded_fields(m)
    for idx in (1, 2, 3):
        assert np.array_equal(s.derivative(f["p"], idx), o.derivative(f["p"], idx)), idx


def test_laplc_bitwise(pair):
    m, cfg, s, o = pair
    f = seeded_fields(m)
    for impfl in (0, 1):
        assert np.array_equal(s.laplc(f["p"], impfl), o.laplc(f["p"], impfl)), impfl


def test_prfr_bcv_div_bitwise(pair):
    m, cfg, s, o = pair
    f = seeded_fields(m)
    for impfl in (0, 1):
        a = s.prfr(f["p"], f["u"], f["v"], f["w"], impfl)
        b = o.prfr(f["p"], f["u"], f["v"], f["w"], impfl)
        for k in range(3):
            assert np.array_equal(a[k], b[k]), (impfl, k)
    for flag in (0.0, 1.0):
        a = s.bcv(f["u"], f["v"], f["w"], flag)
        b = o.bcv(f["u"], f["v"], f["w"], flag)
        for k in range(3):
            assert np.array_equal(a[k], b[k]), (flag, k)
    for iflg in (-1, 1):
        assert np.array_equal(s.div(f["u"], f["v"], f["w"], iflg), o.div(f["u"], f["v"], f["w"], iflg))


@pytest.mark.parametrize("key", ["tiny", "small"])
@pytest.mark.parametrize("kind", [dict(precon=1, itrp1=2, omega=0.5), dict(precon=1, itrp1=1, omega=0.7),
                                  dict(precon=2, itrp1=1, omega=1.0), dict(precon=2, itrp1=2, omega=0.8)])
def test_precon_bitwise(key, kind):
    m = mesh_for(key)
    cfg = (tiny_cfg if key == "tiny" else small_cfg)(**kind)
    s, o = solver(m, cfg), Oracle(m, cfg)
    f = seeded_fields(m)
    assert np.array_equal(s.precon(f["r"], 1), o.precon(f["r"], 1))
    s.close()


@pytest.mark.parametrize("key", ["tall", "tall3"])
def test_long_column_line_solve_bitwise(key):
    """Columns of 100 and 130 levels: the coarsened line scans equal the oracle's
    restatement bit for bit, alone and inside a GCR solve."""
    m = mesh_for(key)
    cfg = tiny_cfg(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=100)
    s, o = solver(m, cfg), Oracle(m, cfg, dot_order=1)
    f = seeded_fields(m)
    assert np.array_equal(s.precon(f["r"], 1), o.precon(f["r"], 1))
    p0 = np.zeros(m.npoin)
    pg, ug, st, rc = s.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    po, uo = o.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    assert st.niter == o.stats().niter
    assert np.array_equal(pg, po)
    s.close()


def test_mpdatam_bitwise(pair):
    m, cfg, s, o = pair
    f = seeded_fields(m)
    for iflag in (1, 2, 3, 4):
        a = s.mpdatam(f["uuu"], f["p"], iflag)
        b = o.mpdatam(f["uuu"], f["p"], iflag)
        assert np.array_equal(a, b), iflag


@pytest.mark.parametrize("key", ["wave64", "patches"])
def test_mpdatam_column_patches(key):
    """Columns of 64 levels in a closed box: the bounds and limiter sweeps stage column patches
    (up to 16 columns, with every column their slots reach) in LDS and gather neighbour
    values from there. Single-field mpdatam (FP64 and the FP32 form) and two full steps (the
    four fields of a step in one sweep) equal the oracle bit for bit (63 columns in 5
    patches; 1,024 columns in 69, dealt to the XCDs in ranges)."""
    print(f"case {key}", flush=True)
    if key == "wave64":
        m, cfg = mesh_for("wave64"), tiny_cfg(**BAND)
    else:
        spec = M.MeshSpec(nx=32, ny=32, nlev=64, lx=25000.0, ly=25000.0, htop=20000.0, hill_h0=1000.0,
                          hill_a=2000.0, seed=4)
        m = M.generate(spec)
        cfg = M.RunConfig("patches", spec, u0=10.0, bvf=0.01, dt=20.0, steps=2, lord=4, eps0=1e-8, exit_mode=1,
                          itr=200, precon=3, itrp1=1, omega=1.0, sponge_depth=3000.0, sponge_tau=300.0)
    f = seeded_fields(m)
    o = Oracle(m, cfg, dot_order=1)
    s = solver(m, cfg)
    for iflag, name in ((1, "u"), (2, "v"), (3, "w"), (4, "p")):
        assert np.array_equal(s.mpdatam(f["uuu"], f[name], iflag), o.mpdatam(f["uuu"], f[name], iflag)), iflag
    s32 = solver(m, cfg, precision="mixed")
    for iflag, name in ((1, "u"), (4, "p")):
        assert np.array_equal(s32.mpdatam(f["uuu"], f[name], iflag), o.mpdatam32(f["uuu"], f[name], iflag)), iflag
    s32.close()
    s.state_init()
    o.init()
    s.run_steps(2)
    for _ in range(2):
        o.step()
    for name in ("unk", "frc", "uuu", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    s.close()


def test_mpdatam_against_reference_vectors():
    g = np.load(os.path.join(GOLD, "tiny_ops.npz"), allow_pickle=False)
    m = mesh_for("tiny")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    s = solver(m, tiny_cfg())
    v = g["vec_in"]
    uuu = np.ascontiguousarray(v[5:8])
    assert np.array_equal(s.mpdatam(uuu, v[0], 1), g["mpdatam_1"])
    assert np.array_equal(s.mpdatam(uuu, v[0], 4), g["mpdatam_4"])
    for k in (1, 2, 3):
        assert np.array_equal(s.derivative(v[0], k), g[f"derivative_{k}"])
    assert np.array_equal(s.laplc(v[0], 1), g["laplc_1"])
    assert np.array_equal(s.precon(v[4], 1), g["precon"])
    s.close()


# The HIP path and the oracle in dot_order=1 use the same inner-product reduction tree,
# so everything below is compared bit for bit. The oracle in dot_order=0 (the reference's
# sequential sums) equals the reference bit for bit (tests/test_oracle.py).
LINE = dict(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)


@pytest.mark.parametrize("key,kw", [
    ("tiny", dict()),                                                 # reference parity mode
    ("small", LINE),
    ("small", dict(precon=1, itrp1=2, omega=0.5, itr=10)),
])
def test_gcrk_bitwise_device_order(key, kw):
    m = mesh_for(key)
    cfg = (tiny_cfg if key == "tiny" else small_cfg)(**kw)
    s, o = solver(m, cfg), Oracle(m, cfg, dot_order=1)
    f = seeded_fields(m)
    p0 = np.zeros(m.npoin)
    pg, ug, st, rc = s.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    po, uo = o.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    so = o.stats()
    assert st.niter == so.niter
    assert np.array_equal(pg, po)
    for k in range(3):
        assert np.array_equal(ug[k], uo[k])
    assert st.eer == so.eer and st.eem == so.eem
    s.close()


def test_incidence_forms_give_identical_steps():
    """Columnar and SELL-64 gathers hold the same slots in the same order: three
    line-preconditioned steps on the 40,000-node mesh give bit-identical states."""
    m = mesh_for("small")
    cfg = small_cfg(**LINE)
    out = []
    for columnar in (1, 0):
        s = solver(m, cfg)
        s.set_incidence(columnar)
        s.state_init()
        s.run_steps(3)
        out.append({k: s.state(k) for k in ("unk", "uuu", "prpot")})
        s.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def _steps(m, cfg, nsteps, **kw):
    s = solver(m, cfg, **kw)
    s.state_init()
    its = [s.gcr_stats().niter]
    for _ in range(nsteps):
        s.run_steps(1)
        its.append(s.gcr_stats().niter)
    return s, its


def _oracle_steps(m, cfg, nsteps, dot_order=1, precision=0):
    o = Oracle(m, cfg, dot_order=dot_order, precision=precision)
    o.init()
    its = [o.stats().niter]
    for _ in range(nsteps):
        o.step()
        its.append(o.stats().niter)
    return o, its


@pytest.mark.parametrize("key,kw,nsteps", [("tiny", dict(), 10), ("small", dict(), 2), ("small", LINE, 5)])
def test_steps_bitwise_device_order(key, kw, nsteps):
    """Full NFT steps (MPDATA x4, two GCR solves, projections): the GPU state equals the
    oracle's bit for bit, iteration counts included."""
    m = mesh_for(key)
    cfg = (tiny_cfg if key == "tiny" else small_cfg)(**kw)
    s, its = _steps(m, cfg, nsteps)
    o, ito = _oracle_steps(m, cfg, nsteps)
    assert its == ito
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    s.close()


@pytest.mark.parametrize("fixture,precon", [("small_line_oracle.npz", 2), ("small_band_oracle.npz", 3)])
def test_steps_small_100_line_vs_sequential_oracle(fixture, precon):
    """100 steps, line-preconditioned GCR(3) at 1e-8 (two-point-flux line, and the band of
    laplc): the GPU against the oracle in the reference's sequential summation order
    (fixtures small_line_oracle.npz, small_band_oracle.npz). Only the inner-product
    summation order differs; tolerance: prognostic fields u, v, w, theta' and the advector
    within 1e-12 relative L2 (sampled field and full-field norm), GCR iterations within one
    per solve."""
    path = os.path.join(GOLD, fixture)
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    g = np.load(path, allow_pickle=False)
    m = mesh_for("small")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    s = solver(m, small_cfg(**dict(LINE, precon=precon)))
    s.state_init()
    its = [s.gcr_stats().niter]
    done = 0
    for nt in (1, 10, 100):
        while done < nt:
            s.run_steps(1)
            its.append(s.gcr_stats().niter)
            done += 1
        st = int(g[f"s{nt}_stride"][0])
        ref_its = g[f"s{nt}_niter"].tolist()
        assert all(abs(a - b) <= 1 for a, b in zip(its, ref_its)), (its, ref_its)
        for name in ("unk", "uuu"):
            a = s.state(name)
            for k in range(a.shape[0]):
                err = rel_l2(a[k, ::st], g[f"s{nt}_{name}_sample"][k])
                print(f"step {nt} {name}[{k}] sampled rel L2 {err:.3e}")
                assert err <= 1e-12, (nt, name, k, err)
                l2 = np.sqrt((a[k] * a[k]).sum())
                assert abs(l2 - g[f"s{nt}_{name}_l2"][k]) <= 1e-12 * g[f"s{nt}_{name}_l2"][k]
    s.close()


def test_steps_small_100_converged_vs_reference():
    """100 steps on the 40,000-node mesh in a regime where the reference's own GCR converges:
    the reference built with no preconditioner (gcrk's iprc=0, precon's p = rhs branch) and a
    relative exit test (max|r| <= 1e-8 max|b|), GCR(3); fixture small_conv.npz written by
    the reference itself. The GPU FP64 path with the same settings, against it directly (only
    the inner-product summation order differs): u, v, w, theta' and the advector within 1e-12
    relative L2 (sampled field and full-field norm) after 1, 10 and 100 steps, GCR iterations
    within one per solve."""
    path = os.path.join(GOLD, "small_conv.npz")
    g = np.load(path, allow_pickle=False)
    m = mesh_for("small")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    s = solver(m, small_cfg(precon=0, exit_mode=1, eps0=1e-8, itr=400))
    s.state_init()
    its = [s.gcr_stats().niter]
    done = 0
    for nt in (1, 10, 100):
        while done < nt:
            st = s.run_steps(1)
            assert st.gcr_unconverged == 0
            its.append(s.gcr_stats().niter)
            done += 1
        st_ = int(g[f"s{nt}_stride"][0])
        ref_its = g[f"s{nt}_niter"].tolist()
        assert len(ref_its) == len(its)
        assert all(abs(a - b) <= 1 for a, b in zip(its, ref_its)), (its, ref_its)
        for name in ("unk", "uuu"):
            a = s.state(name)
            for k in range(a.shape[0]):
                err = rel_l2(a[k, ::st_], g[f"s{nt}_{name}_sample"][k])
                l2 = np.sqrt((a[k] * a[k]).sum())
                ref = g[f"s{nt}_{name}_l2"][k]
                print(f"step {nt} {name}[{k}] sampled rel L2 {err:.3e}, norm rel {abs(l2 - ref) / ref:.1e}")
                assert err <= 1e-12, (nt, name, k, err)
                assert abs(l2 - ref) <= 1e-12 * ref
    print("iterations per solve (GPU)", its[:12], "...")
    s.close()


def test_steps_small_100_reference_parity_mode():
    """100 steps in the reference's own settings (point Jacobi, 120 GCR iterations per
    solve, never reaching eps0) against the flang reference (small_run.npz). In this
    regime the GCR iterate is far from converged and round-off in the inner-product order
    alone changes the fields by ~1e-3 after one step (oracle vs oracle, measured in
    tests/test_oracle.py), so the check is: same iteration counts, and full-field norms
    within 2%."""
    g = np.load(os.path.join(GOLD, "small_run.npz"), allow_pickle=False)
    m = mesh_for("small")
    s, its = _steps(m, small_cfg(), 100)
    assert its == g["s100_niter"].tolist()
    for name in ("unk", "uuu"):
        a = s.state(name)
        for k in range(a.shape[0]):
            l2 = np.sqrt((a[k] * a[k]).sum())
            ref = g[f"s100_{name}_l2"][k]
            print(f"{name}[{k}] norm {l2:.6e} reference {ref:.6e} rel {abs(l2 - ref) / ref:.2e}")
            assert abs(l2 - ref) <= 0.02 * ref
    s.close()


@pytest.mark.parametrize("precision", ["mixed", "mixed_op"])
@pytest.mark.parametrize("precon", [2, 3])
def test_mixed_precision_close_to_fp64(precon, precision):
    """FP32 MPDATA and line preconditioner (either line), FP64 GCR (mixed_op: with the FP32
    laplc inside the iterations): per-field deviation from the FP64 form after 10 steps
    below 1e-5 relative L2, and the solves still converge."""
    m = mesh_for("small")
    cfg = small_cfg(precon=precon, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
    a = solver(m, cfg)
    b = solver(m, cfg, precision=precision)
    for s in (a, b):
        s.state_init()
        st = s.run_steps(10)
        assert st.gcr_unconverged == 0
    for name in ("unk", "uuu"):
        x, y = a.state(name), b.state(name)
        for k in range(x.shape[0]):
            err = rel_l2(y[k], x[k])
            print(f"mixed vs fp64 after 10 steps: {name}[{k}] rel L2 {err:.3e}")
            assert err <= 1e-5, (name, k, err)
    a.close(); b.close()


@pytest.mark.parametrize("precon,drop", [(2, None), (3, None), (3, 1e-4), (3, 1e-2)])
def test_fp32_operator_solve_meets_fp64_residual(precon, drop):
    """precision 2 (FP32 laplc inside GCR): the solve ends only when the true residual of
    the pressure it returns, recomputed here in FP64 by the oracle, is within the
    tolerance; gcrk's velocity outputs (prfr of the incoming pressure) are those of the FP64
    solve; the velocities prfr makes from the returned pressure agree with the FP64 solve's
    to 1e-6 relative L2; the iteration count stays within 35% of the FP64 solve's."""
    print(f"case precon={precon} drop={drop}", flush=True)
    m = mesh_for("small")
    cfg = small_cfg(precon=precon, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
    f = seeded_fields(m)
    u, v, w = f["u"], f["v"], 0.1 * f["w"]
    out = {}
    for prec in ("fp64", "mixed_op"):
        s = solver(m, cfg, precision=prec)
        if prec == "mixed_op":
            assert s.set_mixed_drop() == 1e-3   # the default
            if drop is not None:
                assert s.set_mixed_drop(drop) == drop and s.set_mixed_drop(0.0) == drop
            with pytest.raises(nft.NftError):
                s.set_mixed_drop(1.5)
        p, vel, st, rc = s.gcrk(np.zeros(m.npoin), u, v, w, 1)
        out[prec] = (p, vel, st.niter, st.tol, rc)
        s.close()
    p32, vel32, nit32, tol, rc = out["mixed_op"]
    _, vel64, nit64, tol64, rc64 = out["fp64"]
    assert rc == 0 and rc64 == 0 and tol == tol64
    o = Oracle(m, cfg)
    us, vs, ws = o.prfr(p32, u, v, w, 1)
    r = o.div(us, vs, ws, -1)
    print(f"FP32 operator: {nit32} iterations (FP64 {nit64}); true max|r| {np.abs(r).max():.3e}, tol {tol:.3e}")
    assert np.abs(r).max() <= tol * (1 + 1e-9)
    for k in range(3):
        assert np.array_equal(vel32[k], vel64[k])
    p64 = out["fp64"][0]
    corr32, corr64 = o.prfr(p32, u, v, w, 1), o.prfr(p64, u, v, w, 1)
    for k in range(3):
        err = rel_l2(corr32[k], corr64[k])
        print(f"corrected velocity {k}: rel L2 vs FP64 {err:.2e}")
        assert err <= 1e-6
    assert nit32 <= 1.35 * nit64


def test_gcr_reports_nonconvergence():
    """A forced iteration limit (itr=1) ends the solve early and is reported (rc=1)."""
    m = mesh_for("small")
    cfg = small_cfg(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-12, itr=1)
    s = solver(m, cfg)
    f = seeded_fields(m)
    _, _, st, rc = s.gcrk(np.zeros(m.npoin), f["u"], f["v"], f["w"], 1)
    assert rc == 1 and st.converged == 0 and st.niter == cfg.lord
    s.close()


def test_init_rejects_corrupt_mesh():
    m = mesh_for("tiny")
    bad = m.iedge.copy()
    bad[1, 3] = m.npoin + 5
    import copy
    m2 = copy.copy(m)
    m2._handle = None   # the copy must not free the generator handle it shares
    m2.iedge = bad
    with pytest.raises(nft.NftError, match="out of range"):
        nft.Solver(m2, nft.make_params(tiny_cfg()))


@pytest.mark.parametrize("key,columnar", [("tiny", 1), ("tiny", 0), ("small", 1)])
def test_diagnostics_match_oracle(key, columnar):
    """div(uuu,rhs,1), courant (iflag 1, 2, 3) and stats computed on the GPU after the
    initialisation and after each step, against the oracle's restatement of the reference's
    loops; the records made inside nft_run_steps equal the direct calls bit for bit."""
    m = mesh_for(key)
    cfg = tiny_cfg() if key == "tiny" else small_cfg(**LINE)
    s, o = solver(m, cfg), Oracle(m, cfg, dot_order=1)
    s.set_incidence(columnar)
    s.state_init()
    o.init()
    direct = []
    for step in range(4):
        for flag in (1, 2, 3):
            d = s.diagnostics(flag)
            r = o.diagnostics(flag)
            mine = np.concatenate([np.array(d.div), np.array(d.cour), [d.crnic], d.u, d.v, d.w, d.t])
            _diag_close(mine, r)
        st = o.stats()
        assert d.niter == st.niter and d.nitav == st.nitsm // max(st.icount, 1) and d.eer == st.eer
        if step == 0:
            s.set_diagnostics(every=1, capacity=8, courant_flag=1)
        else:
            direct.append(s.diagnostics(1).row())
        if step < 3:
            s.run_steps(1)
            o.step()
    recs = s.get_diagnostics()
    assert [r.step for r in recs] == [1, 2, 3]
    for r, d in zip(recs, direct):
        assert np.array_equal(r.row(), d)
    s.close()


def test_nan_seeded_state_is_reported_not_hung():
    """A NaN planted in the state (failure injection from the host): the device exit test
    never passes, every solve stops at its iteration limit and is reported unconverged,
    and the run returns instead of hanging."""
    m = mesh_for("tiny")
    cfg = tiny_cfg(itr=5)
    s = solver(m, cfg)
    s.state_init()
    unk = s.state("unk")
    unk[0, m.npoin // 2] = np.nan
    s.set_state("unk", unk)
    st = s.run_steps(1)
    assert st.gcr_solves == 2 and st.gcr_unconverged == 2
    assert not np.isfinite(s.state("unk")).all()
    s.close()


@pytest.mark.parametrize("key,kw,nsteps", [("tiny", dict(), 10), ("small", LINE, 3)])
def test_velprd_steps_bitwise_device_order(key, kw, nsteps):
    """itraj=1: the advector from velprd (gradient + advection fused in one gather, a GCR
    solve with impfl=0, prfr) instead of the extrapolation; the GPU equals the oracle
    (whose restatement equals the reference built with itraj=1) bit for bit."""
    print(f"case {key} {kw} steps={nsteps}", flush=True)
    m = mesh_for(key)
    cfg = (tiny_cfg if key == "tiny" else small_cfg)(**kw)
    s = solver(m, cfg, itraj=1)
    s.state_init()
    o = Oracle(m, cfg, dot_order=1, itraj=1)
    o.init()
    its, ito = [s.gcr_stats().niter], [o.stats().niter]
    st_all = []
    for _ in range(nsteps):
        st_all.append(s.run_steps(1))
        o.step()
        its.append(s.gcr_stats().niter)
        ito.append(o.stats().niter)
    print("iterations", its)
    assert its == ito
    assert all(st.gcr_solves == 3 for st in st_all)
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    s.close()


def test_velprd_mixed_precision_close_to_fp64():
    m = mesh_for("small")
    cfg = small_cfg(**LINE)
    out = []
    for prec in ("fp64", "mixed"):
        s = solver(m, cfg, itraj=1, precision=prec)
        s.state_init()
        s.run_steps(3)
        out.append(s.state("unk"))
        s.close()
    for k in range(4):
        r = rel_l2(out[1][k], out[0][k])
        print(f"unk[{k}] mixed vs fp64 {r:.2e}")
        assert r <= 1e-5


@pytest.mark.parametrize("key", ["wave64", "wave128"])
def test_wave_columns_bitwise(key):
    """Columns of 64 and 128 levels (each wavefront inside one column): laplc, a GCR solve
    and three steps equal the oracle (device summation order) bit for bit, with both
    incidence forms."""
    print(f"case {key}", flush=True)
    m = mesh_for(key)
    cfg = tiny_cfg(**LINE)
    f = seeded_fields(m)
    o = Oracle(m, cfg, dot_order=1)
    lap = o.laplc(f["p"], 1)
    po, uo = o.gcrk(np.zeros(m.npoin), f["u"], f["v"], 0.1 * f["w"], 1)
    nito = o.stats().niter
    o2 = Oracle(m, cfg, dot_order=1)
    o2.init()
    o2.step(3)
    for columnar in (1, 0):
        s = solver(m, cfg)
        assert s.set_incidence(columnar)
        assert np.array_equal(s.laplc(f["p"], 1), lap)
        assert np.array_equal(s.derivative(f["p"], 2), o.derivative(f["p"], 2))
        pg, ug, st, rc = s.gcrk(np.zeros(m.npoin), f["u"], f["v"], 0.1 * f["w"], 1)
        assert st.niter == nito and np.array_equal(pg, po)
        s.state_init()
        s.run_steps(3)
        for name in ("unk", "uuu", "prpot"):
            assert np.array_equal(s.state(name), o2.state(name)), (columnar, name)
        s.close()


BAND = dict(precon=3, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)


@pytest.mark.parametrize("key", ["tiny2", "tall", "wave64", "small"])
def test_band_preconditioner_bitwise(key):
    """precon 3 (in-column band of laplc, probed on the GPU, LU per column, sequential
    substitution per lane from LDS): band, preconditioner and a GCR solve equal the
    oracle's bit for bit (columns of 6, 100, 64 and 20 levels)."""
    print(f"case {key}", flush=True)
    m = mesh_for(key)
    cfg = (small_cfg if key == "small" else tiny_cfg)(**BAND)
    s, o = solver(m, cfg), Oracle(m, cfg, dot_order=1)
    for name in ("band_m2", "band_m1", "band_0", "band_p1", "band_p2"):
        assert np.array_equal(s.field(name), o.field(name)), name
    f = seeded_fields(m)
    assert np.array_equal(s.precon(f["r"], 1), o.precon(f["r"], 1))
    p0 = np.zeros(m.npoin)
    pg, ug, st, rc = s.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    po, uo = o.gcrk(p0, f["u"], f["v"], 0.1 * f["w"], 1)
    print("niter", st.niter, o.stats().niter)
    assert st.niter == o.stats().niter
    assert np.array_equal(pg, po)
    s.close()


@pytest.mark.parametrize("key", ["tiny2", "tall", "wave64", "small"])
def test_band_preconditioner_fp32_bitwise(key):
    """The mixed forms' band solve (FP32 residual, factors and substitution; k_line_penta
    <float>) equals the oracle's FP32 restatement bit for bit, and the FP64 solve to FP32
    rounding."""
    print(f"case {key}", flush=True)
    m = mesh_for(key)
    cfg = (small_cfg if key == "small" else tiny_cfg)(**BAND)
    s, o = solver(m, cfg, precision="mixed"), Oracle(m, cfg)
    r = seeded_fields(m)["r"]
    got = s.precon(r, 1)                      # one sweep from p = 0 with omega = 1: p = M^-1 r
    want = o.line_solve32(r)
    assert np.array_equal(got, want)
    err = rel_l2(got, o.line_solve(r))
    print(f"FP32 band solve vs FP64: rel L2 {err:.2e}")
    assert err < 1e-4
    s.close()


@pytest.mark.parametrize("nx,cpb", [(128, 64), (96, 32), (64, 16), (48, 8), (24, 2)])
def test_band_solve_columns_per_block(nx, cpb):
    """The band solve runs one wavefront per block with one column per lane; on meshes with
    few columns it takes fewer columns per block (halved from 64 until the grid has at least
    256 blocks: nx*nx columns give cpb per block here), and its factors are laid out for that
    width. The band, the FP64 solve and the FP32 solve equal the oracle's at every width."""
    print(f"case {nx}x{nx} columns, {cpb} per block", flush=True)
    spec = M.MeshSpec(nx=nx, ny=nx, nlev=6, lx=500.0 * nx, ly=500.0 * nx, htop=1500.0, hill_h0=200.0,
                      hill_a=2000.0, seed=nx)
    m = M.generate(spec)
    want = 64
    while want > 1 and -(-m.nh // want) < 256:
        want //= 2
    assert m.nh == nx * nx and want == cpb   # the case table above
    cfg = tiny_cfg(**BAND)
    s, o = solver(m, cfg), Oracle(m, cfg, dot_order=1)
    for name in ("band_m2", "band_m1", "band_0", "band_p1", "band_p2"):
        assert np.array_equal(s.field(name), o.field(name)), name
    r = seeded_fields(m)["r"]
    assert np.array_equal(s.precon(r, 1), o.precon(r, 1))
    s.close()
    s32 = solver(m, cfg, precision="mixed")
    assert np.array_equal(s32.precon(r, 1), o.line_solve32(r))
    s32.close()


def test_band_preconditioner_steps_bitwise():
    m = mesh_for("small")
    cfg = small_cfg(**BAND)
    s, its = _steps(m, cfg, 3)
    o, ito = _oracle_steps(m, cfg, 3)
    print("iterations", its)
    assert its == ito
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    s.close()


@pytest.mark.parametrize("key", ["tiny", "small", "wave64"])
@pytest.mark.parametrize("columnar", [1, 0])
def test_fp32_operator_bitwise(key, columnar):
    """The FP32 operator that precision 2 applies inside GCR equals the oracle's
    restatement (oc_laplc32: FP32 edge metrics, scaled gradient rounded to FP32, FP64 sums)
    bit for bit, with both incidence forms and both impfl; other precisions refuse it."""
    print(f"case {key} columnar={columnar}", flush=True)
    m = mesh_for(key)
    cfg = small_cfg() if key == "small" else tiny_cfg()
    s = solver(m, cfg, precision="mixed_op")
    assert s.set_incidence(columnar)
    o = Oracle(m, cfg)
    p = seeded_fields(m)["p"]
    for impfl in (0, 1):
        assert np.array_equal(s.laplc32(p, impfl), o.laplc32(p, impfl)), impfl
    s.close()
    s64 = solver(m, cfg)
    with pytest.raises(nft.NftError):
        s64.laplc32(p, 1)
    s64.close()


@pytest.mark.parametrize("key", ["tiny", "small"])
@pytest.mark.parametrize("columnar", [1, 0])
def test_fp32_mpdatam_bitwise(key, columnar):
    """The mixed forms' MPDATA (k_mp_*<float>: FP32 advector, metrics, vol, rho and
    arithmetic, FP64 fields) equals the oracle's FP32 restatement bit for bit, every
    iflag, with both incidence forms."""
    print(f"case {key} columnar={columnar}", flush=True)
    m = mesh_for(key)
    cfg = small_cfg() if key == "small" else tiny_cfg()
    s = solver(m, cfg, precision="mixed")
    assert s.set_incidence(columnar)
    o = Oracle(m, cfg)
    f = seeded_fields(m)
    for iflag, name in ((1, "u"), (2, "v"), (3, "w"), (4, "p")):
        assert np.array_equal(s.mpdatam(f["uuu"], f[name], iflag), o.mpdatam32(f["uuu"], f[name], iflag)), iflag
    s.close()


@pytest.mark.parametrize("key,nsteps,precision,drop", [("tiny", 10, 1, 0.0), ("small", 3, 1, 0.0),
                                                       ("small", 3, 2, 0.0), ("small", 3, 2, 1e-4)])
def test_mixed_steps_bitwise(key, nsteps, precision, drop):
    """The mixed forms — precision 1: FP32 MPDATA and FP32 band solve, FP64 GCR and laplc;
    precision 2: also the FP32 operator inside GCR and the solves in passes from the FP64
    residual (default pass threshold and 1e-4): the initial projection and full time steps
    equal the oracle's mixed forms (oc_mpdatam32, oc_line_solve32, oc_laplc32, the same
    passes, the GPU's summation order) bit for bit, with equal GCR iteration counts."""
    print(f"case {key} precision={precision} drop={drop}", flush=True)
    m = mesh_for(key)
    cfg = (small_cfg if key == "small" else tiny_cfg)(**BAND)
    s = solver(m, cfg, precision=precision)
    if drop:
        s.set_mixed_drop(drop)
    s.state_init()
    its = [s.gcr_stats().niter]
    for _ in range(nsteps):
        s.run_steps(1)
        its.append(s.gcr_stats().niter)
    o = Oracle(m, cfg, dot_order=1, precision=precision, mixed_drop=drop)
    o.init()
    ito = [o.stats().niter]
    for _ in range(nsteps):
        o.step()
        ito.append(o.stats().niter)
    print("iterations", its, ito)
    assert its == ito
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    s.close()
