"""The CPU oracle against the reference's own outputs (golden vectors written by the
flang-compiled reference, tests/golden/make_golden.py) and against closed forms."""
import os

import numpy as np
import pytest

from common import mesh_checksum, mesh_for, seeded_fields, small_cfg, tiny_cfg, rel_l2
from oracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="module")
def tiny_ops():
    g = load("tiny_ops.npz")
    m = mesh_for("tiny")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m)), "regenerated mesh differs from the fixture's"
    return g, m


def test_setup_fields_bitwise(tiny_ops):
    g, m = tiny_ops
    o = Oracle(m, tiny_cfg())
    for k in ("alpha", "d1", "coefw"):
        assert np.array_equal(o.field(k), g[k]), k
    assert o.beta == g["beta"][0]


def test_operators_bitwise(tiny_ops):
    g, m = tiny_ops
    o = Oracle(m, tiny_cfg())
    v = g["vec_in"]
    for k in (1, 2, 3):
        assert np.array_equal(o.derivative(v[0], k), g[f"derivative_{k}"])
    for k in (0, 1):
        assert np.array_equal(o.laplc(v[0], k), g[f"laplc_{k}"])
        assert np.array_equal(np.concatenate(o.bcv(v[1], v[2], v[3], k)), g[f"bcv_{k}"])
        assert np.array_equal(np.concatenate(o.prfr(v[0], v[1], v[2], v[3], k)), g[f"prfr_{k}"])
    assert np.array_equal(o.div(v[1], v[2], v[3], -1), g["div"])
    assert np.array_equal(o.precon(v[4], 1), g["precon"])
    uuu = np.ascontiguousarray(v[5:8])
    assert np.array_equal(o.mpdatam(uuu, v[0], 1), g["mpdatam_1"])
    assert np.array_equal(o.mpdatam(uuu, v[0], 4), g["mpdatam_4"])
    p, outs = o.gcrk(np.zeros(m.npoin), v[1], v[2], v[3], 1)
    assert np.array_equal(p, g["gcrk_p"])
    assert np.array_equal(np.concatenate(outs), g["gcrk_u"])
    st = o.stats()
    assert st.niter == int(g["gcrk_stats"][0])
    assert st.eer == g["gcrk_stats"][1] and st.eem == g["gcrk_stats"][2]


def test_planted_sign_error_is_caught(tiny_ops):
    """A plausible transcription slip (the i2 update of derivative with + instead of -)
    changes the result, so the golden comparison would catch it."""
    g, m = tiny_ops
    v = g["vec_in"]
    wrong = np.zeros(m.npoin)
    i1, i2 = m.iedge[0] - 1, m.iedge[1] - 1
    u = (v[0][i1] + v[0][i2]) * 0.5
    np.add.at(wrong, i1, m.sn[0] * u)
    np.add.at(wrong, i2, m.sn[0] * u)
    assert not np.allclose(wrong, g["derivative_1"])


def test_time_steps_bitwise():
    g = load("tiny_run.npz")
    m = mesh_for("tiny")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    for nt in (1, 10):
        o = Oracle(m, tiny_cfg())
        o.init()
        its = [o.stats().niter]
        for _ in range(nt):
            o.step()
            its.append(o.stats().niter)
        for k in ("unk", "frc", "uuu", "unkdash", "prpot"):
            assert np.array_equal(o.state(k), g[f"s{nt}_{k}"]), (nt, k)
        assert its == g[f"s{nt}_niter"].tolist()


def test_velprd_steps_bitwise():
    """itraj=1: the advector from velprd (F:4448-4531, a third GCR solve per step) instead
    of the extrapolation; against the reference built with that switch."""
    g = load("tiny_velprd.npz")
    m = mesh_for("tiny")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    for nt in (1, 10):
        o = Oracle(m, tiny_cfg(), itraj=1)
        o.init()
        its = [o.stats().niter]
        for _ in range(nt):
            o.step()
            its.append(o.stats().niter)
        for k in ("unk", "frc", "uuu", "unkdash", "prpot"):
            assert np.array_equal(o.state(k), g[f"s{nt}_{k}"]), (nt, k)
        assert its == g[f"s{nt}_niter"].tolist()
    # the extrapolated advector gives a different state: the fixture pins the switch
    o = Oracle(m, tiny_cfg(), itraj=0)
    o.init()
    o.step()
    assert not np.array_equal(o.state("uuu"), g["s1_uuu"])


def _printed(x, ref):
    """x agrees with the reference's e11.4 print `ref` to within 0.6 of its last digit."""
    if ref == 0.0:
        return abs(x) < 1e-300
    e = np.floor(np.log10(abs(ref))) + 1
    return abs(x - ref) <= 0.6e-4 * 10.0 ** e


def test_diagnostics_against_reference_prints():
    """div(uuu,rhs,1), courant and stats after the initialisation and after each of 10
    steps (tiny mesh): the oracle's values against the reference's printed blocks."""
    g = load("tiny_run.npz")
    rows = g["s10_diag"]
    m = mesh_for("tiny")
    o = Oracle(m, tiny_cfg())
    o.init()
    for k in range(rows.shape[0]):
        if k:
            o.step()
        d = o.diagnostics(1)
        st = o.stats()
        mine = np.concatenate([d[:4], [st.eer, st.eem, st.niter, st.nitsm // max(st.icount, 1), d[8]], d[4:8], d[9:]])
        ref = rows[k]
        for j in range(29):
            if j in (6, 7, 8):
                assert mine[j] == ref[j], (k, j, mine[j], ref[j])
            else:
                assert _printed(mine[j], ref[j]), (k, j, mine[j], ref[j])


def test_small_mesh_steps_against_reference():
    """40,000-node mesh, point-Jacobi GCR(3): 1 and 10 steps against the reference."""
    g = load("small_run.npz")
    m = mesh_for("small")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    o = Oracle(m, small_cfg())
    o.init()
    its = [o.stats().niter]
    done = 0
    for nt in (1, 10):
        while done < nt:
            o.step()
            its.append(o.stats().niter)
            done += 1
        s = int(g[f"s{nt}_stride"][0])
        for k in ("unk", "uuu", "frc", "unkdash"):
            a = o.state(k)
            assert np.array_equal(a[:, ::s], g[f"s{nt}_{k}_sample"]), (nt, k)
            # full-field norms (summation order of the fixture's numpy sum may differ)
            l2 = np.sqrt((a * a).sum(axis=1))
            assert np.allclose(l2, g[f"s{nt}_{k}_l2"], rtol=1e-14, atol=0), (nt, k)
        assert its == g[f"s{nt}_niter"].tolist()


def test_small_mesh_converged_steps_against_reference():
    """40,000-node mesh, GCR(3) that converges: the reference built with no preconditioner
    (gcrk's iprc=0) and the relative exit test max|r| <= 1e-8 max|b| (small_conv.npz). The
    oracle with the same settings, in the reference's sequential summation order, equals it
    bit for bit after 1 and 10 steps, iteration counts included; every solve converged
    (counts below the 1200-iteration cap)."""
    g = load("small_conv.npz")
    m = mesh_for("small")
    assert np.array_equal(g["mesh_checksum"], mesh_checksum(m))
    o = Oracle(m, small_cfg(precon=0, exit_mode=1, eps0=1e-8, itr=400))
    o.init()
    its = [o.stats().niter]
    done = 0
    for nt in (1, 10):
        while done < nt:
            o.step()
            its.append(o.stats().niter)
            done += 1
        s = int(g[f"s{nt}_stride"][0])
        for k in ("unk", "uuu", "frc", "unkdash"):
            a = o.state(k)
            assert np.array_equal(a[:, ::s], g[f"s{nt}_{k}_sample"]), (nt, k)
        assert its == g[f"s{nt}_niter"].tolist()
    assert max(g["s100_niter"]) < 1200


# ---- closed forms ------------------------------------------------------------------
@pytest.mark.parametrize("key", ["tiny2", "tall", "tall3"])
def test_line_solve_matches_dense(key):
    """Wavefront scans (6 levels), coarsened scans (100 and 130 levels: 2 and 3 rows per
    lane) against a dense solve of each column's tridiagonal system."""
    m = mesh_for(key)
    o = Oracle(m, tiny_cfg(precon=2, itrp1=1, omega=1.0))
    rhs = np.random.default_rng(3).normal(size=m.npoin)
    z = o.line_solve(rhs)
    a, inv, c = o.field("line_a"), o.field("line_inv"), o.field("line_c")
    # rebuild each column's tridiagonal from its factors and solve densely
    L = m.nlev
    for col in range(m.nh):
        s = slice(col * L, (col + 1) * L)
        aa, ii, cc = a[s], inv[s], c[s]
        # factors: den_k = 1/inv_k = b_k - a_k c_{k-1}; c_k = up_k * inv_k
        den = 1.0 / ii
        up = cc * den
        b = den + aa * np.concatenate([[0.0], cc[:-1]])
        T = np.diag(b) + np.diag(aa[1:], -1) + np.diag(up[:-1], 1)
        assert np.allclose(np.linalg.solve(T, rhs[s]), z[s], rtol=1e-12, atol=1e-12 * np.abs(z[s]).max())


def test_line_operator_symmetric_negative_definite():
    """Per column, vol*rho*M is symmetric (c_ij vol_i rho_i = c_ji vol_j rho_j) and negative
    definite; a swapped index (vol_j for vol_i) or a sign slip breaks one of the two."""
    m = mesh_for("tiny2")
    o = Oracle(m, tiny_cfg(precon=2))
    a, inv, c = o.field("line_a"), o.field("line_inv"), o.field("line_c")
    L = m.nlev
    for col in range(m.nh):
        s = slice(col * L, (col + 1) * L)
        aa, ii, cc = a[s], inv[s], c[s]
        den = 1.0 / ii
        up = cc * den
        b = den + aa * np.concatenate([[0.0], cc[:-1]])
        T = np.diag(b) + np.diag(aa[1:], -1) + np.diag(up[:-1], 1)
        W = np.diag(m.vol[s]) @ T
        assert np.allclose(W, W.T, rtol=1e-12, atol=1e-14 * np.abs(W).max())
        assert np.linalg.eigvalsh(0.5 * (W + W.T)).max() < 0


def test_laplc_of_linear_vanishes_inside():
    m = mesh_for("small")
    o = Oracle(m, tiny_cfg())
    # interior nodes two rings away from the boundary
    bnd = np.zeros(m.npoin, bool)
    bnd[m.ibpoin[0] - 1] = True
    near = bnd.copy()
    for _ in range(2):
        nxt = near.copy()
        for ie in (m.iedge, m.kbface):
            nxt[ie[0][near[ie[1] - 1]] - 1] = True
            nxt[ie[1][near[ie[0] - 1]] - 1] = True
        near = nxt
    p = 3.0 * m.coord[0] - 2.0 * m.coord[1] + 0.5 * m.coord[2]
    lp = o.laplc(p, 0)
    scale = np.abs(o.laplc(np.sin(m.coord[0] / 800.0) * 1000.0, 0)).max()
    assert np.abs(lp[~near]).max() <= 1e-9 * scale


@pytest.mark.parametrize("precon,extra", [(2, dict(itrp1=1, omega=1.0)), (1, dict(itrp1=2, omega=0.5))])
def test_gcrk_residual_is_the_true_residual(precon, extra):
    """The recursively updated GCR residual equals b + L p recomputed from the returned p
    with a dense operator; GCR never raises ||r||_2 and stays above the least-squares
    minimum of ||L p + b|| (numpy lstsq). A wrong sign or a dropped term in the p/r
    updates breaks the first check."""
    m = mesh_for("tiny2")
    cfg = tiny_cfg(precon=precon, exit_mode=1, eps0=1e-14, itr=100, lord=4, **extra)
    o = Oracle(m, cfg)
    rng = np.random.default_rng(5)
    u, v, w = 10 + rng.normal(size=m.npoin), rng.normal(size=m.npoin), rng.normal(size=m.npoin)
    p, _ = o.gcrk(np.zeros(m.npoin), u, v, w, 1)
    u2, v2, w2 = o.prfr(p, u, v, w, 1)   # gcrk's own velocity output is the initial one (:1829)
    r = o.div(u2, v2, w2, -1)
    b = o.div(*o.bcv(u, v, w, 1.0), -1)
    n = m.npoin
    A = np.stack([o.laplc(np.eye(n)[k], 1) for k in range(n)], axis=1)
    assert np.abs(b + A @ p - r).max() <= 1e-11 * np.abs(b).max()
    p_ls = np.linalg.lstsq(A, -b, rcond=None)[0]
    r_ls = np.linalg.norm(b + A @ p_ls)
    assert r_ls * (1 - 1e-9) <= np.linalg.norm(r) < np.linalg.norm(b)


def test_gcrk_reaches_relative_tolerance_small_mesh():
    """40,000-node mesh, line preconditioner, relative exit 1e-8: the residual recomputed
    from the returned pressure meets the tolerance."""
    m = mesh_for("small")
    cfg = small_cfg(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200, lord=3)
    o = Oracle(m, cfg)
    rng = np.random.default_rng(5)
    u, v, w = 10 + rng.normal(size=m.npoin), rng.normal(size=m.npoin), 0.1 * rng.normal(size=m.npoin)
    p, _ = o.gcrk(np.zeros(m.npoin), u, v, w, 1)
    r = o.div(*o.prfr(p, u, v, w, 1), -1)
    b = o.div(*o.bcv(u, v, w, 1.0), -1)
    st = o.stats()
    assert 0 < st.niter < 600
    assert np.abs(r).max() <= 1e-8 * np.abs(b).max() * 1.0001


def test_mpdatam_conserves_mass_and_keeps_constant_without_flow():
    m = mesh_for("tiny2")
    o = Oracle(m, tiny_cfg())
    psi = np.random.default_rng(1).normal(size=m.npoin)
    zero = np.zeros((3, m.npoin))
    assert np.array_equal(o.mpdatam(zero, psi, 4), psi)
    # theta' (iflag 4) with inflow value 0 and a flow tangent to every boundary except x-faces:
    # total mass changes only through the x-faces
    uuu = np.zeros((3, m.npoin)); uuu[0] = 5.0
    out = o.mpdatam(uuu, psi, 4)
    assert np.isfinite(out).all()
    # monotone: donor-cell + FCT keeps values within the initial range (constant inflow 0)
    lo, hi = min(psi.min(), 0.0), max(psi.max(), 0.0)
    assert out.min() >= lo - 1e-12 and out.max() <= hi + 1e-12


def test_summation_order_effect():
    """Same algorithm, inner products summed sequentially (reference) or with the HIP
    path's tree: with a converging solve (line preconditioner, 1e-8) the prognostic fields
    agree to ~1e-13 after two steps; in the reference's non-converging point-Jacobi
    setting round-off alone moves them by far more, which bounds what any reordered
    implementation can show against the reference there."""
    m = mesh_for("small")
    conv = small_cfg(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
    a, b = Oracle(m, conv, dot_order=0), Oracle(m, conv, dot_order=1)
    for o in (a, b):
        o.init()
        o.step(2)
    for k in range(4):
        assert rel_l2(b.state("unk")[k], a.state("unk")[k]) <= 1e-12
    a, b = Oracle(m, small_cfg(), dot_order=0), Oracle(m, small_cfg(), dot_order=1)
    for o in (a, b):
        o.init()
        o.step(1)
    assert max(rel_l2(b.state("unk")[k], a.state("unk")[k]) for k in range(4)) > 1e-8


@pytest.mark.parametrize("key", ["tiny2", "tall"])
def test_probed_band_is_laplc_and_its_solve_is_exact(key):
    """precon 3: the probed in-column band equals the entries of laplc's dense matrix (built
    from unit vectors) bit for bit, and the banded LU solve equals a dense solve of it."""
    m = mesh_for(key)
    o = Oracle(m, tiny_cfg(precon=3, itrp1=1, omega=1.0))
    n, L = m.npoin, m.nlev
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = o.laplc(e, 1)
    B = np.zeros((n, n))
    for off, name in ((-2, "m2"), (-1, "m1"), (0, "0"), (1, "p1"), (2, "p2")):
        b = o.field("band_" + name)
        for i in range(n):
            if 0 <= i % L + off < L:
                assert b[i] == A[i, i + off], (name, i)
                B[i, i + off] = b[i]
    r = np.random.default_rng(1).normal(size=n)
    z = o.line_solve(r)
    zd = np.linalg.solve(B, r)
    assert np.abs(z - zd).max() <= 1e-12 * np.abs(zd).max()
    # the mixed forms' FP32 substitution: a solve of the same band to FP32 accuracy, and
    # not the FP64 one
    z32 = o.line_solve32(r)
    assert np.linalg.norm(B @ z32 - r) <= 1e-5 * np.linalg.norm(r)
    assert 0 < np.abs(z32 - zd).max() <= 1e-4 * np.abs(zd).max()
    assert np.array_equal(z32, z32.astype(np.float32).astype(np.float64))   # FP32 values


@pytest.mark.parametrize("key,impfl", [("tiny2", 1), ("tiny2", 0), ("small", 1)])
def test_fp32_operator_restatement(key, impfl):
    """The mixed form's FP32 operator in the oracle (oc_laplc32) against its definition
    composed here from independent pieces: the FP64 derivative of a mesh copy whose edge
    metrics sn, sb are rounded to FP32 (sbb kept), laplc's scaling, bcv and rho written out
    in numpy, the scaled gradient rounded to FP32, and the divergence of that mesh copy.
    Equal bit for bit; and within FP32 rounding of laplc, but not equal to it."""
    import dataclasses
    m = mesh_for(key)
    cfg = tiny_cfg() if key.startswith("tiny") else small_cfg()
    o = Oracle(m, cfg)
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)
    mr = dataclasses.replace(m, sn=f32(m.sn), sb=f32(m.sb), _handle=None)   # the copy owns no native mesh
    orr = Oracle(mr, cfg)
    p = np.random.default_rng(5).standard_normal(m.npoin)
    g = [orr.derivative(p, k) for k in (1, 2, 3)]
    dt, v = cfg.dt, m.vol
    g = [dt * gk / v for gk in g]
    fz = o.field("coefw") if impfl else np.ones(m.npoin)
    g = [0.5 * g[0], 0.5 * g[1], 0.5 * g[2] * fz]
    g = list(o.bcv(g[0], g[1], g[2], 0.0))
    rho = o.field("rh0")
    g = [f32(gk * rho) for gk in g]
    a = [orr.derivative(g[k], k + 1) for k in range(3)]
    want = (a[0] + a[1] + a[2]) / (v * rho)
    got = o.laplc32(p, impfl)
    assert np.array_equal(got, want)
    ref = o.laplc(p, impfl)
    err = rel_l2(got, ref)
    print(f"FP32 operator vs laplc: rel L2 {err:.2e}")
    assert 0 < err < 1e-5


@pytest.mark.parametrize("key", ["tiny2", "small"])
def test_fp32_mpdatam_restatement_tracks_fp64(key):
    """The mixed forms' MPDATA restatement (oc_mpdatam32, FP32 arithmetic) stays within FP32
    rounding of the FP64 mpdatam, which is the reference's bit for bit: a dropped term, a
    wrong sign or a transposed operand moves the result by orders of magnitude more. Every
    inflow treatment (iflag 1-4)."""
    m = mesh_for(key)
    o = Oracle(m, small_cfg() if key == "small" else tiny_cfg())
    f = seeded_fields(m)
    for iflag, name in ((1, "u"), (2, "v"), (3, "w"), (4, "p")):
        a = o.mpdatam(f["uuu"], f[name], iflag)
        b = o.mpdatam32(f["uuu"], f[name], iflag)
        moved = rel_l2(a, f[name])
        err = rel_l2(b, a)
        print(f"iflag {iflag}: FP32 vs FP64 {err:.2e} (the step moves the field by {moved:.2e})")
        assert 0 < err <= 2e-7 and moved > 1e3 * err
