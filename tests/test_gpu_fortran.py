"""The Fortran host side: nftdrv reads the reference's own input files and drives the GPU
time loop through iso_c_binding; its state must equal the Python/ctypes path bit for bit
(same library, same inputs) and the oracle within the parity tolerance."""
import os
import subprocess

import numpy as np
import pytest

from common import mesh_for, tiny_cfg, rel_l2
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
