"""Build recipes: the gfx950 HIP library, the mesh generator and (for tests) the oracle.

    python -m src_amd.build          # everything buildable here
Outputs stay in-tree (src_amd/lib/*.so, oracle/libnft_oracle.so, oracle/_ref/) so they
travel with the repository snapshot to the GPU machine.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "src_amd", "csrc")
LIB = os.path.join(ROOT, "src_amd", "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             # no FP contraction: results must round like the reference's scalar code
             "-ffp-contract=off",
             # packed FP32 instructions are kept out of generated code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout[-4000:] + r.stderr[-4000:])
    return r


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libnft_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "nft_capi.hip")])
    os.replace(tmp, out)
    return out


def build_mesh():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libnft_mesh.so")
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-o", out, os.path.join(CSRC, "meshgen.cpp")])
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "libnft_oracle.so")


def build_fortran():
    """Fortran drop-in module and driver (flang), linked against libnft_hip.so."""
    fc = shutil.which("flang") or shutil.which("amdflang")
    if fc is None:
        return None
    fdir = os.path.join(ROOT, "src_amd", "fortran")
    if not os.path.isdir(fdir):
        return None
    bdir = os.path.join(LIB, "fmod")
    os.makedirs(bdir, exist_ok=True)
    srcs = [os.path.join(fdir, f) for f in ("nft_hip.f90", "nftdrv.f90") if os.path.exists(os.path.join(fdir, f))]
    if not srcs:
        return None
    exe = os.path.join(LIB, "nftdrv")
    _run([fc, "-O2", "-J", bdir, *srcs, "-o", exe, "-L" + LIB, "-lnft_hip", "-Wl,-rpath,$ORIGIN"])
    # test program of the drop-in subroutines (tests/test_gpu_dropin.py runs it)
    tsrc = os.path.join(ROOT, "tests", "fortran", "nftdrop.f90")
    if os.path.exists(tsrc):
        tdir = os.path.join(bdir, "tests")
        os.makedirs(tdir, exist_ok=True)
        _run([fc, "-O2", "-J", tdir, os.path.join(fdir, "nft_hip.f90"), tsrc,
              "-o", os.path.join(ROOT, "tests", "fortran", "nftdrop"), "-L" + LIB, "-lnft_hip",
              "-Wl,-rpath,$ORIGIN/../../src_amd/lib"])
    return exe


def build_all(verbose=True):
    outs = [build_mesh(), build_hip(), build_oracle()]
    f = build_fortran()
    if f:
        outs.append(f)
    if verbose:
        for o in outs:
            print("built", os.path.relpath(o, ROOT))
    return outs


if __name__ == "__main__":
    build_all()
    sys.exit(0)
