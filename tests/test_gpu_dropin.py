"""The Fortran drop-in subroutines (src_amd/fortran/nft_hip.f90) called the way the
reference's mpsolver calls them.

tests/fortran/nftdrop.f90 fills the reference's COMMON blocks from its own input files,
calls nft_attach, then derivative, laplc, prfr, bcv, div, precon, mpdatam(uuu,unk(1,k),k)
and gcrk(p,unk1(1,1),...,uuu(1,1),...,1) with npoinx leading dimensions, and courant/stats;
every output and /itero/ (lippsLFroude_implgwopks3D_unstr3.f :1953-1963) must equal the
oracle bit for bit. courant and stats work on the arrays they are given and leave the
device-resident time-step state alone. In the periodic box (iper = 2 in inputinc.d) the
program reads the coincident sets after the mesh into /coin/, /per2/, /per4/, nft_attach
builds the periodic context from them, and mpdatamper replaces mpdatam.
"""
import os
import subprocess

import numpy as np
import pytest

from common import BAND_KW, diag_close, mesh_for, periodic_cfg, seeded_fields, small_cfg, tiny_cfg
from oracle import Oracle
from src_amd import nft
from src_amd.refio import write_reference_inputs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROP = os.path.join(ROOT, "tests", "fortran", "nftdrop")


@pytest.mark.parametrize("key,kw", [("tiny", dict()), ("small", dict(BAND_KW, exit_mode=0, itr=5)),
                                    ("periodic2", dict())],
                         ids=["tiny-point-relaxation", "small-band", "periodic-box"])
def test_dropin_subroutines_match_oracle(tmp_path, key, kw):
    if not os.path.exists(DROP):
        pytest.skip("nftdrop not built (flang missing)")
    print(f"case {key} {kw}", flush=True)
    m = mesh_for(key)
    cfg = {"tiny": tiny_cfg, "small": small_cfg, "periodic2": periodic_cfg}[key](**kw)
    d = str(tmp_path)
    write_reference_inputs(d, m, cfg, ntime=1, sm_flag=1, itrp1=cfg.itrp1, omega=cfg.omega)
    f = seeded_fields(m)
    n = m.npoin
    vin = np.stack([f["p"], f["u"], f["v"], f["w"], f["r"], *f["uuu"]])
    vin.tofile(os.path.join(d, "vec_in.bin"))   # (8, n) C order = vin(npoin, 8) in Fortran order
    r = subprocess.run([DROP, str(m.nlev), str(cfg.precon), "0", str(cfg.lord), repr(cfg.strat)], cwd=d,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    out = np.fromfile(os.path.join(d, "dropin_out.bin"), dtype=np.float64)
    assert out.size == 25 * n + 5 + n
    arr = out[:25 * n].reshape(25, n)
    itero = out[25 * n:25 * n + 5]
    cour = out[25 * n + 5:]

    # nft_attach runs the reference's absolute exit test (exit_mode 0)
    o = Oracle(m, cfg, exit_mode=0, dot_order=1)
    p, u, v, w, rr = f["p"], f["u"], f["v"], f["w"], f["r"]
    uuu = f["uuu"]
    for k in range(3):
        assert np.array_equal(arr[k], o.derivative(p, k + 1)), ("derivative", k + 1)
    for k in range(2):
        assert np.array_equal(arr[3 + k], o.laplc(p, k)), ("laplc", k)
    for k, x in enumerate(o.prfr(p, u, v, w, 1)):
        assert np.array_equal(arr[5 + k], x), ("prfr", k)
    for j, flag in enumerate((0.0, 1.0)):
        for k, x in enumerate(o.bcv(u, v, w, flag)):
            assert np.array_equal(arr[8 + 3 * j + k], x), ("bcv", flag, k)
    assert np.array_equal(arr[14], o.div(u, v, w, -1))
    assert np.array_equal(arr[15], o.div(u, v, w, 1))
    assert np.array_equal(arr[16], o.precon(rr, 1))
    for k, fld in enumerate((u, v, w, p)):
        assert np.array_equal(arr[17 + k], o.mpdatam(uuu, fld, k + 1)), ("mpdatam", k + 1)
    po, uo = o.gcrk(np.zeros(n), u, v, 0.1 * w, 1)
    st = o.stats()
    print(f"gcrk: niter {itero[0]:.0f} (oracle {st.niter}), eer {itero[3]:.6e}")
    assert np.array_equal(arr[21], po)
    for k in range(3):
        assert np.array_equal(arr[22 + k], uo[k]), ("gcrk velocity", k)
    assert itero.tolist() == [st.niter, st.nitsm, st.icount, st.eer, st.eem]
    # courant(uuu,cour,1): the numbers equal the C entry point's; the statistics printed
    # (crnmx, ...) are checked against the oracle in test_courant_and_stats_leave_state_alone
    s = nft.Solver(m, nft.make_params(cfg))
    c2, _ = s.courant(uuu, 1)
    s.close()
    assert np.array_equal(cour, c2)
    assert "crnmx,crnmn,crnav,crnsd:" in r.stdout and "umx,umn,uav,usd:" in r.stdout


def test_courant_and_stats_leave_state_alone():
    """nft_courant / nft_stats on host arrays: statistics equal the oracle's diagnostics of
    the same arrays, and the run continues bit for bit as if they had not been called."""
    m = mesh_for("small")
    cfg = small_cfg(**BAND_KW)
    f = seeded_fields(m)
    a, b = nft.Solver(m, nft.make_params(cfg)), nft.Solver(m, nft.make_params(cfg))
    for s in (a, b):
        s.state_init()
        s.run_steps(1)
    unk_any = np.stack([f["u"], f["v"], f["w"], f["p"]])
    o = Oracle(m, cfg, dot_order=1)
    o.state("uuu")[:] = f["uuu"]
    o.state("unk")[:] = unk_any
    for flag in (1, 2, 3):
        cour, d = b.courant(f["uuu"], flag)
        ref = o.diagnostics(flag)
        assert d.crnic == ref[8]
        mine = np.concatenate([ref[:4], np.array(d.cour), [d.crnic], ref[9:]])
        diag_close(mine, ref)
    ds = b.stats(unk_any)
    ref = o.diagnostics(1)
    diag_close(np.concatenate([ref[:9], ds.u, ds.v, ds.w, ds.t]), ref)
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(a.state(name), b.state(name)), name
    a.run_steps(1)
    b.run_steps(1)
    for name in ("unk", "uuu", "prpot"):
        assert np.array_equal(a.state(name), b.state(name)), name
    a.close(); b.close()


def test_wrappers_reject_wrong_shapes():
    """The ctypes wrappers check array shapes before calling the library (which reads npoin
    values, or k x ld for state blocks, from the pointers it gets)."""
    m = mesh_for("tiny")
    s = nft.Solver(m, nft.make_params(tiny_cfg()))
    n = m.npoin
    with pytest.raises(nft.NftError):
        s.laplc(np.zeros(n - 1), 1)
    with pytest.raises(nft.NftError):
        s.derivative(np.zeros((2, n)), 1)
    with pytest.raises(nft.NftError):
        s.set_state("unk", np.zeros((4, n + 3)))
    with pytest.raises(nft.NftError):
        s.set_state("uuu", np.zeros((4, n)))
    with pytest.raises(nft.NftError):
        s.mpdatam(np.zeros((3, n - 1)), np.zeros(n), 1)
    with pytest.raises(nft.NftError):
        s.state("nope")
    s.close()
