"""The GCR grid-level sums at the small-grid limit.

On grids of at most 256 blocks the finalize of the divergence's dots runs inside the
direction kernel, the direction's inside the next update and, with the band line, the exit
test inside the band solve; each replays the finalize's summation exactly. The 65,536-node
mesh below (32 x 32 columns x 64 levels) launches exactly 256 gather and 256 vector blocks,
so every virtual wavefront that can hold partials is full; one more row of columns (33 x 32,
264 blocks) takes the separate finalize kernels. On both, the initial projection and two
time steps equal the oracle's (the GPU's summation order) bit for bit, iteration counts
included, for GCR(4) and GCR(3), FP64 and the FP32-operator form.
"""
import numpy as np
import pytest

from oracle import Oracle
from src_amd import mesh as M
from src_amd import nft

pytestmark = pytest.mark.gpu

SPECS = {
    "fold256": M.MeshSpec(nx=32, ny=32, nlev=64, lx=25000.0, ly=25000.0, htop=20000.0, hill_h0=1000.0,
                          hill_a=2000.0, seed=4),
    "nofold264": M.MeshSpec(nx=33, ny=32, nlev=64, lx=25800.0, ly=25000.0, htop=20000.0, hill_h0=1000.0,
                            hill_a=2000.0, seed=4),
}


def _cfg(spec, lord):
    return M.RunConfig("fold", spec, u0=10.0, bvf=0.01, dt=20.0, steps=2, lord=lord, eps0=1e-8, exit_mode=1,
                       itr=200, precon=3, itrp1=1, omega=1.0, sponge_depth=3000.0, sponge_tau=300.0)


@pytest.mark.parametrize("key,lord,precision", [("fold256", 4, 0), ("fold256", 3, 0), ("fold256", 4, 2),
                                                ("nofold264", 4, 0)])
def test_small_grid_sums_bitwise(key, lord, precision):
    print(f"case {key} lord={lord} precision={precision}", flush=True)
    spec = SPECS[key]
    m = M.generate(spec)
    assert m.npoin == spec.nx * spec.ny * spec.nlev
    cfg = _cfg(spec, lord)
    s = nft.Solver(m, nft.make_params(cfg, precision=precision))
    s.state_init()
    its = [s.gcr_stats().niter]
    for _ in range(2):
        s.run_steps(1)
        its.append(s.gcr_stats().niter)
    o = Oracle(m, cfg, dot_order=1, precision=precision)
    o.init()
    ito = [o.stats().niter]
    for _ in range(2):
        o.step()
        ito.append(o.stats().niter)
    print("iterations", its, ito, flush=True)
    assert its == ito
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name
    s.close()
