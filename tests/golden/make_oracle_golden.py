"""Writes tests/golden/small_line_oracle.npz and small_band_oracle.npz: 100 NFT steps on
the 40,000-node mesh with the vertical line preconditioner (two-point-flux tridiagonal,
and the in-column band of laplc) and a 1e-8 relative GCR(3) exit, computed by the CPU
oracle in the reference's own (sequential) summation order. The reference itself has no
line preconditioner and does not converge with its point relaxation in reasonable time,
so this converging run uses the oracle, which equals the reference bit for bit on every
operator (tiny_ops.npz, tiny_run.npz, small_run.npz).

    python tests/golden/make_oracle_golden.py [line|band]
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

from common import mesh_checksum, mesh_for, small_cfg  # noqa: E402
from oracle import Oracle  # noqa: E402

LINE = dict(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
BAND = dict(LINE, precon=3)


def main():
    kind = sys.argv[1] if len(sys.argv) > 1 else "line"
    m = mesh_for("small")
    o = Oracle(m, small_cfg(**(BAND if kind == "band" else LINE)))
    o.init()
    its = [o.stats().niter]
    out = {"mesh_checksum": mesh_checksum(m)}
    for step in range(1, 101):
        o.step()
        its.append(o.stats().niter)
        if step in (1, 10, 100):
            s = 10 if step == 100 else 40
            for k in ("unk", "uuu", "frc", "unkdash"):
                a = o.state(k)
                if step < 100 or k in ("unk", "uuu"):
                    out[f"s{step}_{k}_sample"] = a[:, ::s].copy()
                out[f"s{step}_{k}_l2"] = np.sqrt((a * a).sum(axis=1))
            out[f"s{step}_stride"] = np.array([s])
            out[f"s{step}_niter"] = np.array(its, dtype=np.int32)
            print(step, its[-5:], flush=True)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", f"small_{kind}_oracle.npz"), **out)


if __name__ == "__main__":
    main()
