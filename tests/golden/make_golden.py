"""Writes the golden fixtures of tests/golden from the flang-compiled reference.

Run in the build container (it needs the reference source to build oracle/_ref):
    python tests/golden/make_golden.py [--tiny-only]
It builds the reference with oracle/build_ref.sh, writes the reference's input files for
generated meshes, runs
  * the harness in `ops` mode: the reference's own derivative, laplc, bcv, prfr, div,
    precon, mpdatam and gcrk on seeded vectors over the tiny mesh (tiny_ops.npz), and
  * the reference program itself (`run` mode) for 1 and 10 steps on the tiny mesh
    (tiny_run.npz, with the diagnostics it prints after the initialisation and every
    step) and for 1, 10 and 100 steps on the 40,000-node mesh (small_run.npz),
  * the multigrid branch of precon (mg_flag=1) on the tiny mesh's generated nested coarse
    meshes: its pieces in `mgops` mode and 1 and 10 steps of the program (tiny_mg.npz),
  * 1 and 10 steps with the velprd advector (itraj=1) on the tiny mesh (tiny_velprd.npz),
  * with --converged only: 1, 10 and 100 steps on the 40,000-node mesh of a reference build
    whose GCR converges (no preconditioner, relative exit test; small_conv.npz),
  * with --periodic only: the doubly periodic box (iper=2) on a generated 280-node periodic
    mesh: the `ops` outputs (mpdatamper in place of mpdatam) and 1 and 10 steps of the
    program, both with the reference build of tiny_ops.npz and with the converging build of
    small_conv.npz (tiny_per.npz),
and stores inputs and outputs. Only numbers are stored.
"""
from __future__ import annotations

import glob
import os
import re
import subprocess
import sys
import resource
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from src_amd import mesh as M  # noqa: E402
from src_amd.refio import write_reference_inputs  # noqa: E402
from common import parse_diagnostics  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# reference-parity settings: point Jacobi (sm_flag1=1), 2 sweeps, omega 0.5, absolute exit
PARITY = dict(sm_flag=1, itrp1=2, omega=0.5)


def tiny_config():
    return M.RunConfig("tiny", M.tiny_spec(), u0=10.0, bvf=0.01, dt=10.0, steps=10, lord=3,
                       eps0=1e-7, exit_mode=0, itr=30, precon=1, itrp1=2, omega=0.5,
                       sponge_depth=500.0, sponge_tau=300.0)


def small_config():
    c = M.CONFIGS["small"]
    return M.RunConfig("small", c.mesh, u0=c.u0, bvf=c.bvf, dt=c.dt, steps=100, lord=3,
                       eps0=1e-8, exit_mode=0, itr=40, precon=1, itrp1=2, omega=0.5,
                       sponge_depth=3000.0, sponge_tau=300.0)


def strat_literal(cfg):
    return repr(cfg.strat)


def build_ref(cfg, itraj=0, iprc=1, relexit=0):
    exe = subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh"), str(cfg.lord),
                          strat_literal(cfg), str(itraj), str(iprc), str(relexit)], check=True,
                         capture_output=True, text=True)
    return exe.stdout.strip().splitlines()[-1]


def converged_config():
    """The 40,000-node mesh and flow with a GCR(3) that converges: no preconditioner (gcrk's
    iprc=0, precon's p = rhs branch) and the relative exit max|r| <= 1e-8 max|b|."""
    c = M.CONFIGS["small"]
    return M.RunConfig("small", c.mesh, u0=c.u0, bvf=c.bvf, dt=c.dt, steps=100, lord=3,
                       eps0=1e-8, exit_mode=1, itr=400, precon=0, itrp1=2, omega=0.5,
                       sponge_depth=3000.0, sponge_tau=300.0)


def write_converged():
    """1, 10 and 100 steps of the reference itself in a converging regime (small_conv.npz):
    built with iprc=0 and the relative exit test, the other settings as small_run.npz."""
    ccfg = converged_config()
    exe = build_ref(ccfg, iprc=0, relexit=1)
    m = M.generate(ccfg.mesh)
    stride = {1: 40, 10: 40, 100: 10}
    with tempfile.TemporaryDirectory() as d:
        runs = {}
        for nt in (1, 10, 100):
            st = run_steps(exe, m, ccfg, d, nt)
            s = stride[nt]
            for k in ("unk", "uuu", "frc", "unkdash"):
                a = st[k]
                runs[f"s{nt}_{k}_sample"] = a[:, ::s].copy()
                runs[f"s{nt}_{k}_l2"] = np.sqrt((a * a).sum(axis=1))
            runs[f"s{nt}_prpot_sample"] = st["prpot"][::s].copy()
            runs[f"s{nt}_prpot_l2"] = np.array([np.sqrt((st["prpot"] ** 2).sum())])
            runs[f"s{nt}_stride"] = np.array([s])
            runs[f"s{nt}_niter"] = st["niter"]
            print(f"converged run: {nt} steps done, niter {st['niter'].tolist()[:12]}...")
        np.savez_compressed(os.path.join(OUT, "small_conv.npz"), mesh_checksum=mesh_checksum(m), **runs)
    print("converged fixtures written")


def periodic_config(converged=False):
    """The periodic box (src_amd.mesh.periodic_spec, 8 x 7 columns x 5 levels) with the
    tiny mesh's flow; converged: no preconditioner and the relative exit test to 1e-8."""
    if converged:
        return M.RunConfig("periodic", M.periodic_spec(), u0=10.0, bvf=0.01, dt=10.0, steps=10, lord=3,
                           eps0=1e-8, exit_mode=1, itr=400, precon=0, itrp1=2, omega=0.5,
                           sponge_depth=500.0, sponge_tau=300.0)
    return M.RunConfig("periodic", M.periodic_spec(), u0=10.0, bvf=0.01, dt=10.0, steps=10, lord=3,
                       eps0=1e-7, exit_mode=0, itr=30, precon=1, itrp1=2, omega=0.5,
                       sponge_depth=500.0, sponge_tau=300.0)


def write_periodic():
    """tiny_per.npz: the reference with iper=2 on the periodic box."""
    cfg, ccfg = periodic_config(), periodic_config(converged=True)
    m = M.generate(cfg.mesh)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        exe = build_ref(cfg)
        out.update(run_ops(exe, m, cfg, d))
        exe_c = build_ref(ccfg, iprc=0, relexit=1)
        for tag, e, c in (("", exe, cfg), ("conv_", exe_c, ccfg)):
            for nt in (1, 10):
                st = run_steps(e, m, c, d, nt)
                for k in ("unk", "uuu", "frc", "unkdash", "prpot", "niter"):
                    out[f"{tag}s{nt}_{k}"] = st[k]
                print(f"periodic {tag or 'parity'} run: {nt} steps, niter {st['niter'].tolist()}")
    np.savez_compressed(os.path.join(OUT, "tiny_per.npz"), mesh_checksum=mesh_checksum(m),
                        coin_checksum=np.array([a.sum() for a in (m.coin, m.coin2, m.coin4)], dtype=np.int64),
                        strat=np.array([cfg.strat]), **out)
    print("periodic fixtures written")


def mesh_checksum(m):
    return np.array([m.npoin, m.nedge, m.nbface, m.nboun, m.iedge.sum(), m.kbface.sum(),
                     m.ibpoin.sum(), m.sn.sum(), np.abs(m.sn).sum(), m.sb.sum(), m.vol.sum(),
                     m.sbb.sum(), m.coord.sum()], dtype=np.float64)


def _unlimited_stack():
    # the reference keeps large automatic arrays (npoinx/nedgex-sized locals) on the stack
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def seeded_vectors(m, seed=7):
    rng = np.random.default_rng(seed)
    n = m.npoin
    x, y, z = m.coord
    v = np.zeros((8, n))
    v[0] = 50.0 * np.sin(x / 700.0) * np.cos(y / 900.0) + 10.0 * z / 1000.0 + rng.normal(0, 1.0, n)
    v[1] = 10.0 + rng.normal(0, 1.0, n)
    v[2] = rng.normal(0, 1.0, n)
    v[3] = rng.normal(0, 0.5, n)
    v[4] = rng.normal(0, 1e-3, n)
    v[5] = 10.0 + rng.normal(0, 0.5, n)
    v[6] = rng.normal(0, 0.5, n)
    v[7] = rng.normal(0, 0.2, n)
    return v


def run_ops(exe, m, cfg, d):
    write_reference_inputs(d, m, cfg, ntime=1, **PARITY)
    with open(os.path.join(d, "harness.d"), "w") as f:
        f.write(strat_literal(cfg) + "\n")
    v = seeded_vectors(m)
    with open(os.path.join(d, "vec_in.bin"), "wb") as f:
        for k in range(8):
            f.write(v[k].tobytes())
    r = subprocess.run([exe, "ops"], cwd=d, capture_output=True, text=True,
                       preexec_fn=_unlimited_stack)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    raw = np.fromfile(os.path.join(d, "vec_out.bin"), dtype=np.float64)
    n = m.npoin
    out = {}
    pos = 0

    def take(name, k):
        nonlocal pos
        out[name] = raw[pos:pos + k].copy()
        pos += k

    take("alpha", n); take("d1", n); take("coefw", n); take("beta", 1)
    for k in (1, 2, 3):
        take(f"derivative_{k}", n)
    take("laplc_0", n); take("laplc_1", n)
    take("bcv_0", 3 * n); take("bcv_1", 3 * n)
    take("prfr_0", 3 * n); take("prfr_1", 3 * n)
    take("div", n)
    take("precon", n)
    take("mpdatam_1", n); take("mpdatam_4", n)
    take("gcrk_p", n); take("gcrk_u", 3 * n); take("gcrk_stats", 3)
    assert pos == raw.size, (pos, raw.size)
    out["vec_in"] = v
    return out


# multigrid branch (mg_flag=1): sweeps of the V-cycle on the tiny mesh's nested levels
MG_SWEEPS = dict(itrp2=1, itrp_c1=2, itrp_c2=2, itrp_cc=3)


def run_mgops(exe, m, levels, cfg, d):
    write_reference_inputs(d, m, cfg, ntime=1, mg=dict(levels=levels, **MG_SWEEPS), **PARITY)
    with open(os.path.join(d, "harness.d"), "w") as f:
        f.write(strat_literal(cfg) + "\n")
    v = seeded_vectors(m)
    with open(os.path.join(d, "vec_in.bin"), "wb") as f:
        for k in range(8):
            f.write(v[k].tobytes())
    r = subprocess.run([exe, "mgops"], cwd=d, capture_output=True, text=True,
                       preexec_fn=_unlimited_stack)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    raw = np.fromfile(os.path.join(d, "vec_out.bin"), dtype=np.float64)
    n, n1, n2 = m.npoin, levels[0][0].npoin, levels[1][0].npoin
    out, pos = {}, 0

    def take(name, k):
        nonlocal pos
        out[name] = raw[pos:pos + k].copy()
        pos += k

    take("d2", n1); take("d3", n2); take("beta23", 2)
    take("restrict_1", n1); take("restrict_2", n2)
    take("prolong_1", n); take("prolong_2", n1)
    take("laplc_c_1", n1); take("laplc_c_0", n1); take("laplc_cc_1", n2)
    take("precon", n)
    take("gcrk_p", n); take("gcrk_u", 3 * n); take("gcrk_stats", 3)
    assert pos == raw.size, (pos, raw.size)
    return out


def read_restart(d, n):
    files = glob.glob(os.path.join(d, "Rstart_Out*.d"))
    assert len(files) == 1, files
    with open(files[0]) as f:
        lines = f.read().splitlines()
    os.remove(files[0])

    def block(start, k):
        return np.array([[float(t) for t in ln.split()] for ln in lines[start:start + n]]).T

    st = {}
    st["uuu"] = block(0, 3)
    st["uuu1"] = block(n, 3)
    st["unkdash"] = block(2 * n, 3)
    st["unk"] = block(3 * n, 4)
    st["frc"] = block(4 * n, 4)
    last = block(5 * n, 4)
    st["prpot"], st["rhs"], st["cour"], st["coefw"] = last
    return st


NITER_RE = re.compile(r"niter,nitav:\s*(\d+)\s+(\d+)")


def run_steps(exe, m, cfg, d, ntime, mg=None):
    write_reference_inputs(d, m, cfg, ntime=ntime, mg=mg, **PARITY)
    r = subprocess.run([exe, "run"], cwd=d, capture_output=True, text=True,
                       preexec_fn=_unlimited_stack)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-3000:] + r.stderr[-3000:])
    st = read_restart(d, m.npoin)
    niter = [int(a) for a, _ in NITER_RE.findall(r.stdout)]
    st["niter"] = np.array(niter, dtype=np.int32)  # init solve first, then one per step
    st["diag"] = parse_diagnostics(r.stdout)       # init block first, then one per step
    return st


def main():
    if "--converged" in sys.argv:
        write_converged()
        return
    if "--periodic" in sys.argv:
        write_periodic()
        return
    only_tiny = "--tiny-only" in sys.argv
    tcfg = tiny_config()
    exe = build_ref(tcfg)
    m = M.generate(tcfg.mesh)
    with tempfile.TemporaryDirectory() as d:
        ops = run_ops(exe, m, tcfg, d)
        np.savez_compressed(os.path.join(OUT, "tiny_ops.npz"), mesh_checksum=mesh_checksum(m),
                            strat=np.array([tcfg.strat]), **ops)
        runs = {}
        for nt in (1, 10):
            st = run_steps(exe, m, tcfg, d, nt)
            for k, val in st.items():
                runs[f"s{nt}_{k}"] = val
        np.savez_compressed(os.path.join(OUT, "tiny_run.npz"), mesh_checksum=mesh_checksum(m), **runs)
        # multigrid branch on generated nested coarse meshes
        levels = M.multigrid_meshes(tcfg.mesh)
        mg = run_mgops(exe, m, levels, tcfg, d)
        for nt in (1, 10):
            st = run_steps(exe, m, tcfg, d, nt, mg=dict(levels=levels, **MG_SWEEPS))
            for k in ("unk", "uuu", "frc", "unkdash", "prpot", "niter"):
                mg[f"s{nt}_{k}"] = st[k]
        np.savez_compressed(os.path.join(OUT, "tiny_mg.npz"), mesh_checksum=mesh_checksum(m),
                            level_checksums=np.stack([mesh_checksum(lm) for lm, _ in levels]),
                            match1=levels[0][1], match2=levels[1][1],
                            sweeps=np.array([MG_SWEEPS[k] for k in ("itrp2", "itrp_c1", "itrp_c2", "itrp_cc")]),
                            **mg)
        # the velprd advector (itraj=1), from a reference build with that switch
        exe1 = build_ref(tcfg, itraj=1)
        vp = {}
        for nt in (1, 10):
            st = run_steps(exe1, m, tcfg, d, nt)
            for k in ("unk", "uuu", "frc", "unkdash", "prpot", "niter"):
                vp[f"s{nt}_{k}"] = st[k]
        np.savez_compressed(os.path.join(OUT, "tiny_velprd.npz"), mesh_checksum=mesh_checksum(m), **vp)
    print("tiny fixtures written")
    if only_tiny:
        return

    scfg = small_config()
    exe = build_ref(scfg)
    m = M.generate(scfg.mesh)
    stride = {1: 40, 10: 40, 100: 10}
    with tempfile.TemporaryDirectory() as d:
        runs = {}
        for nt in (1, 10, 100):
            st = run_steps(exe, m, scfg, d, nt)
            s = stride[nt]
            for k in ("unk", "uuu", "frc", "unkdash"):
                a = st[k]
                runs[f"s{nt}_{k}_sample"] = a[:, ::s].copy()
                runs[f"s{nt}_{k}_l2"] = np.sqrt((a * a).sum(axis=1))
                runs[f"s{nt}_{k}_sum"] = a.sum(axis=1)
                runs[f"s{nt}_{k}_max"] = np.abs(a).max(axis=1)
            runs[f"s{nt}_prpot_sample"] = st["prpot"][::s].copy()
            runs[f"s{nt}_prpot_l2"] = np.array([np.sqrt((st["prpot"] ** 2).sum())])
            runs[f"s{nt}_stride"] = np.array([s])
            runs[f"s{nt}_niter"] = st["niter"]
            print(f"small: {nt} steps done, niter per step {st['niter'].tolist()}")
        np.savez_compressed(os.path.join(OUT, "small_run.npz"), mesh_checksum=mesh_checksum(m), **runs)
    print("small fixtures written")


if __name__ == "__main__":
    main()
