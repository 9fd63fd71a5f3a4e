"""Binary mesh file (host only) and the full-precision state file (GPU): round trips
that must be exact, and malformed files rejected with an error, not a crash."""
import os

import numpy as np
import pytest

from common import mesh_for, small_cfg
from src_amd import nft

LINE = dict(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
FIELDS = ("iedge", "sn", "kbface", "sb", "ibpoin", "sbb", "anp", "coord", "vol")


@pytest.mark.parametrize("key", ["tiny", "tiny2"])
def test_mesh_file_round_trip(tmp_path, key):
    m = mesh_for(key)
    path = str(tmp_path / "mesh.bin")
    nft.mesh_save(path, m)
    out = nft.mesh_load(path)
    for k in FIELDS:
        assert out[k].dtype == getattr(m, k).dtype and np.array_equal(out[k], getattr(m, k)), k
    assert np.array_equal(out["col_start"], m.col_start())


def test_mesh_file_rejects_bad_input(tmp_path):
    m = mesh_for("tiny")
    good = tmp_path / "mesh.bin"
    nft.mesh_save(str(good), m)
    data = good.read_bytes()
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTAMESH" + data[8:])
    with pytest.raises(nft.NftError, match="not a mesh file"):
        nft.mesh_load(str(bad))
    cut = tmp_path / "cut.bin"
    cut.write_bytes(data[: len(data) // 2])
    with pytest.raises(nft.NftError, match="truncated"):
        nft.mesh_load(str(cut))
    with pytest.raises(nft.NftError, match="cannot open"):
        nft.mesh_load(str(tmp_path / "missing.bin"))


@pytest.mark.gpu
def test_state_file_resume_is_bit_exact(tmp_path):
    """3 steps, save, a fresh context loads the file and runs 2 more: the state equals 5
    uninterrupted steps bit for bit, and the step and GCR counters carry over."""
    m = mesh_for("small")
    cfg = small_cfg(**LINE)
    a = nft.Solver(m, nft.make_params(cfg))
    a.state_init()
    a.run_steps(5)
    ref = {k: a.state(k) for k in ("unk", "frc", "uuu", "unkdash", "prpot")}
    st_ref = a.gcr_stats()
    b = nft.Solver(m, nft.make_params(cfg))
    b.state_init()
    b.run_steps(3)
    path = str(tmp_path / "state.bin")
    b.save_state(path)
    b.close()
    c = nft.Solver(m, nft.make_params(cfg))
    c.load_state(path)
    assert c.steps_done() == 3
    c.run_steps(2)
    assert c.steps_done() == 5
    for k, v in ref.items():
        assert np.array_equal(c.state(k), v), k
    assert c.gcr_stats().nitsm == st_ref.nitsm and c.gcr_stats().icount == st_ref.icount
    with open(path, "r+b") as f:
        f.seek(8)
        f.write(np.int64(m.npoin + 1).tobytes())
    with pytest.raises(nft.NftError, match="nodes"):
        c.load_state(path)
    a.close(); c.close()
