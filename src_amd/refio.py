"""Writers for the reference's input files (lippsLFroude_implgwopks3D_unstr3.f `input`,
:2172-2679): mesh3d.d (plus the two coarse-mesh units, here copies of the fine mesh),
inputinc.d, input_mg_v3.d, inputre.d and the identity match files match1.d/match2.d.

These files let the reference program run on the same inputs as this package; the
package itself takes the mesh through its C ABI (nft_init) and never parses them.
"""
from __future__ import annotations

import os

from .mesh import Mesh, RunConfig


def fmt(x: float) -> str:
    """Shortest text that reads back as the same double."""
    return repr(float(x))


def write_reference_inputs(d: str, mesh: Mesh, cfg: RunConfig, ntime: int,
                           sm_flag: int = 1, itrp1: int = 2, omega: float = 0.5,
                           eps0: float | None = None, itr: int | None = None,
                           lord: int | None = None) -> None:
    os.makedirs(d, exist_ok=True)
    mesh.write_text(os.path.join(d, "mesh3d.d"))
    for name in ("mesh3d_c.d", "mesh3d_cc.d"):
        p = os.path.join(d, name)
        if os.path.lexists(p):
            os.remove(p)
        os.symlink("mesh3d.d", p)
    N = mesh.npoin
    with open(os.path.join(d, "match1.d"), "w") as f:
        f.write("\n".join(str(i + 1) for i in range(N)) + "\n")
    with open(os.path.join(d, "match2.d"), "w") as f:
        f.write("\n".join(str(i + 1) for i in range(N)) + "\n")
    eps0 = cfg.eps0 if eps0 is None else eps0
    itr = cfg.itr if itr is None else itr
    lord = cfg.lord if lord is None else lord
    sp = cfg.mesh
    with open(os.path.join(d, "inputinc.d"), "w") as f:
        f.write("ntime nrkstage istst mrestart isfera iper\n")
        f.write(f"{ntime} 2 2 0 0 0\n")
        f.write("cfl delt adapt relax\n")
        f.write(f"0.5 {fmt(cfg.dt)} 0. 0.\n")
        f.write("p0inf u0inf v0inf w0inf Re\n")
        f.write(f"0. {fmt(cfg.u0)} 0. 0. 1000.\n")
        f.write("elliptic solver\n")
        f.write("ielast eps0 itr lord\n")
        f.write(f"0 {fmt(eps0)} {itr} {lord}\n")
        f.write("itrp betap\n")
        f.write("1 1.\n")
        f.write("irl\n")
        f.write("1\n")
        f.write("dxabl dxabr dyabd dyabu dzabd dzabu\n")
        f.write(f"0. 0. 0. 0. 0. {fmt(cfg.sponge_depth)}\n")
        f.write("towxl towxr towyd towyu towzd towzu\n")
        f.write(f"0. 0. 0. 0. 0. {fmt(cfg.sponge_tau)}\n")
    with open(os.path.join(d, "input_mg_v3.d"), "w") as f:
        f.write("mg_flag\n0\n")
        f.write("smoothers\nsm_flag1 sm_flag2\n")
        f.write(f"{sm_flag} {sm_flag}\n")
        f.write("omega\n")
        f.write(f"{fmt(omega)}\n")
        f.write("sweeps\nitrp1 itrp2 itrp_c1 itrp_c2 itrp_cc\n")
        f.write(f"{itrp1} 0 0 0 0\n")
        f.write("match_flag\n1\n")
    with open(os.path.join(d, "inputre.d"), "w") as f:
        f.write("nrstart\n0\nntimet\n0\nirestart nre\n0 0\n")
