"""Writers for the reference's input files (lippsLFroude_implgwopks3D_unstr3.f `input`,
:2172-2679): mesh3d.d, the two coarse-mesh units mesh3d_c.d / mesh3d_cc.d with their node
maps match1.d / match2.d (copies of the fine mesh and identity maps unless multigrid
levels are given), inputinc.d, input_mg_v3.d and inputre.d.

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
                           lord: int | None = None, mg: dict | None = None) -> None:
    """mg = None: the coarse-mesh units are the fine mesh and mg_flag=0. Otherwise
    mg = dict(levels=[(mesh_c, match1), (mesh_cc, match2)], itrp2=, itrp_c1=, itrp_c2=,
    itrp_cc=) writes the coarse meshes and matches and sets mg_flag=1."""
    os.makedirs(d, exist_ok=True)
    mesh.write_text(os.path.join(d, "mesh3d.d"))
    for name in ("mesh3d_c.d", "mesh3d_cc.d", "match1.d", "match2.d"):
        p = os.path.join(d, name)
        if os.path.lexists(p):
            os.remove(p)
    if mg is None:
        for name in ("mesh3d_c.d", "mesh3d_cc.d"):
            os.symlink("mesh3d.d", os.path.join(d, name))
        N = mesh.npoin
        matches = [range(1, N + 1), range(1, N + 1)]
    else:
        (m1, t1), (m2, t2) = mg["levels"]
        m1.write_text(os.path.join(d, "mesh3d_c.d"))
        m2.write_text(os.path.join(d, "mesh3d_cc.d"))
        matches = [t1, t2]
    for name, t in zip(("match1.d", "match2.d"), matches):
        with open(os.path.join(d, name), "w") as f:
            f.write("\n".join(str(int(i)) for i in t) + "\n")
    eps0 = cfg.eps0 if eps0 is None else eps0
    itr = cfg.itr if itr is None else itr
    lord = cfg.lord if lord is None else lord
    sp = cfg.mesh
    with open(os.path.join(d, "inputinc.d"), "w") as f:
        f.write("ntime nrkstage istst mrestart isfera iper\n")
        # iper = 2: the doubly periodic box, whose coincident node sets follow the mesh
        f.write(f"{ntime} 2 2 0 0 {2 if mesh.coin is not None else 0}\n")
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
        f.write(f"mg_flag\n{0 if mg is None else 1}\n")
        f.write("smoothers\nsm_flag1 sm_flag2\n")
        f.write(f"{sm_flag} {sm_flag}\n")
        f.write("omega\n")
        f.write(f"{fmt(omega)}\n")
        f.write("sweeps\nitrp1 itrp2 itrp_c1 itrp_c2 itrp_cc\n")
        if mg is None:
            f.write(f"{itrp1} 0 0 0 0\n")
        else:
            f.write(f"{itrp1} {mg['itrp2']} {mg['itrp_c1']} {mg['itrp_c2']} {mg['itrp_cc']}\n")
        f.write("match_flag\n1\n")
    with open(os.path.join(d, "inputre.d"), "w") as f:
        f.write("nrstart\n0\nntimet\n0\nirestart nre\n0 0\n")
