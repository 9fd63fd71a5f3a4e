"""The Fortran host side: nftdrv reads the reference's own input files and drives the GPU
time loop through iso_c_binding; its state must equal the Python/ctypes path bit for bit
(same library, same inputs) and the oracle within the parity tolerance."""
import os
import subprocess

import numpy as np
import pytest

from common import mesh_for, parse_diagnostics, periodic_cfg, tiny_cfg, rel_l2
from oracle import Oracle
from src_amd import nft
from src_amd.refio import write_reference_inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "src_amd", "lib", "nftdrv")


def test_driver_built():
    assert os.path.exists(DRV), "nftdrv not built (flang missing?)"


@pytest.mark.gpu
def test_fortran_driver_matches_ctypes_path(tmp_path):
    if not os.path.exists(DRV):
        pytest.skip("nftdrv not built")
    m = mesh_for("tiny")
    cfg = tiny_cfg()
    nsteps = 3
    d = str(tmp_path)
    write_reference_inputs(d, m, cfg, ntime=nsteps, sm_flag=1, itrp1=cfg.itrp1, omega=cfg.omega)
    r = subprocess.run([DRV, str(m.nlev), "1", "0", str(cfg.lord), repr(cfg.strat), "0"], cwd=d,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    st = np.fromfile(os.path.join(d, "nftdrv_state.bin"), dtype=np.float64).reshape(8, m.npoin)
    s = nft.Solver(m, nft.make_params(cfg))
    s.state_init()
    s.run_steps(nsteps)
    unk, uuu, prpot = s.state("unk"), s.state("uuu"), s.state("prpot")
    s.close()
    assert np.array_equal(st[0:4], unk)
    assert np.array_equal(st[4:7], uuu)
    assert np.array_equal(st[7], prpot)
    o = Oracle(m, cfg, dot_order=1)
    o.init()
    o.step(nsteps)
    assert np.array_equal(st[0:4], o.state("unk"))
    # The diagnostics it prints (computed on the GPU): the block after the initialisation
    # against the reference's own print (same trajectory); the blocks after each step
    # against the oracle in the GPU's summation order (in this absolute-tolerance mode the
    # GCR iterates, and so the later blocks, depend on the summation order).
    rows = parse_diagnostics(r.stdout)
    assert rows.shape == (nsteps + 1, 29)
    ref0 = np.load(os.path.join(ROOT, "tests", "golden", "tiny_run.npz"))["s10_diag"][0]
    o2 = Oracle(m, cfg, dot_order=1)
    o2.init()
    for k in range(nsteps + 1):
        if k:
            o2.step()
        d = o2.diagnostics(1)
        st2 = o2.stats()
        ora = np.concatenate([d[:4], [st2.eer, st2.eem, st2.niter, st2.nitsm // max(st2.icount, 1), d[8]], d[4:8], d[9:]])
        for ref in ([ref0, ora] if k == 0 else [ora]):
            for j in range(29):
                x, y = rows[k, j], ref[j]
                if j in (6, 7, 8):
                    assert x == y, (k, j, x, y)
                else:
                    assert (x == 0 if y == 0 else abs(x - y) <= 1.2e-4 * 10.0 ** (np.floor(np.log10(abs(y))) + 1)), (k, j, x, y)


@pytest.mark.gpu
def test_fortran_driver_multigrid(tmp_path):
    """mg_flag=1 in input_mg_v3.d: nftdrv reads mesh3d_c.d, mesh3d_cc.d, match1.d and
    match2.d as the reference's `input` does (level-2 metrics quirk included) and runs
    precon's V-cycle on the GPU; its state equals the oracle's (multigrid on, GPU summation
    order) bit for bit."""
    if not os.path.exists(DRV):
        pytest.skip("nftdrv not built")
    from common import MG_SWEEPS, mg_levels_for
    m = mesh_for("tiny")
    cfg = tiny_cfg()
    nsteps = 3
    d = str(tmp_path)
    write_reference_inputs(d, m, cfg, ntime=nsteps, sm_flag=1, itrp1=cfg.itrp1, omega=cfg.omega,
                           mg=dict(levels=mg_levels_for("tiny"), **MG_SWEEPS))
    r = subprocess.run([DRV, str(m.nlev), "1", "0", str(cfg.lord), repr(cfg.strat), "0"], cwd=d,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    st = np.fromfile(os.path.join(d, "nftdrv_state.bin"), dtype=np.float64).reshape(8, m.npoin)
    o = Oracle(m, cfg, dot_order=1)
    o.set_multigrid(mg_levels_for("tiny", as_read=True), transfer=0, **MG_SWEEPS)
    o.init()
    o.step(nsteps)
    assert np.array_equal(st[0:4], o.state("unk"))
    assert np.array_equal(st[4:7], o.state("uuu"))
    assert np.array_equal(st[7], o.state("prpot"))


@pytest.mark.gpu
def test_fortran_driver_velprd(tmp_path):
    """itraj=1 through nftdrv: equals the oracle with velprd bit for bit."""
    if not os.path.exists(DRV):
        pytest.skip("nftdrv not built")
    m = mesh_for("tiny")
    cfg = tiny_cfg()
    nsteps = 3
    d = str(tmp_path)
    write_reference_inputs(d, m, cfg, ntime=nsteps, sm_flag=1, itrp1=cfg.itrp1, omega=cfg.omega)
    r = subprocess.run([DRV, str(m.nlev), "1", "0", str(cfg.lord), repr(cfg.strat), "0", "1"], cwd=d,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    st = np.fromfile(os.path.join(d, "nftdrv_state.bin"), dtype=np.float64).reshape(8, m.npoin)
    o = Oracle(m, cfg, dot_order=1, itraj=1)
    o.init()
    o.step(nsteps)
    assert np.array_equal(st[0:4], o.state("unk"))
    assert np.array_equal(st[4:7], o.state("uuu"))
    assert np.array_equal(st[7], o.state("prpot"))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_fortran_driver_band_line_and_precision(tmp_path, precision):
    """precon=3 (the band line) with each precision through nftdrv, relative exit: the
    state equals the ctypes path with the same settings bit for bit (same library, same
    inputs)."""
    if not os.path.exists(DRV):
        pytest.skip("nftdrv not built")
    print(f"case precision={precision}", flush=True)
    m = mesh_for("tiny")
    cfg = tiny_cfg(precon=3, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
    nsteps = 3
    d = str(tmp_path)
    write_reference_inputs(d, m, cfg, ntime=nsteps, sm_flag=1, itrp1=cfg.itrp1, omega=cfg.omega)
    r = subprocess.run([DRV, str(m.nlev), "3", str(precision), str(cfg.lord), repr(cfg.strat), "1"], cwd=d,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    st = np.fromfile(os.path.join(d, "nftdrv_state.bin"), dtype=np.float64).reshape(8, m.npoin)
    s = nft.Solver(m, nft.make_params(cfg, precision=precision))
    s.state_init()
    s.run_steps(nsteps)
    unk, uuu, prpot = s.state("unk"), s.state("uuu"), s.state("prpot")
    s.close()
    assert np.array_equal(st[0:4], unk)
    assert np.array_equal(st[4:7], uuu)
    assert np.array_equal(st[7], prpot)


@pytest.mark.gpu
def test_fortran_driver_periodic_box(tmp_path):
    """iper=2: nftdrv reads the coincident sets after the mesh and runs the periodic scheme
    through nft_init_periodic; 3 steps equal the ctypes path and the oracle (tree order)
    bit for bit."""
    if not os.path.exists(DRV):
        pytest.skip("nftdrv not built")
    m = mesh_for("periodic2")
    cfg = periodic_cfg(converged=True)
    nsteps = 3
    d = str(tmp_path)
    write_reference_inputs(d, m, cfg, ntime=nsteps, sm_flag=1, itrp1=cfg.itrp1, omega=cfg.omega)
    r = subprocess.run([DRV, str(m.nlev), str(cfg.precon), "0", str(cfg.lord), repr(cfg.strat), str(cfg.exit_mode)],
                       cwd=d, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    st = np.fromfile(os.path.join(d, "nftdrv_state.bin"), dtype=np.float64).reshape(8, m.npoin)
    s = nft.Solver(m, nft.make_params(cfg))
    s.state_init()
    s.run_steps(nsteps)
    assert np.array_equal(st[0:4], s.state("unk"))
    assert np.array_equal(st[4:7], s.state("uuu"))
    assert np.array_equal(st[7], s.state("prpot"))
    s.close()
    o = Oracle(m, cfg, dot_order=1)
    o.init()
    o.step(nsteps)
    assert np.array_equal(st[0:4], o.state("unk"))
    assert np.array_equal(st[7], o.state("prpot"))
