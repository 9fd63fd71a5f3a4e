"""Times the reference program (oracle/_ref, flang -O3 FP64 build) against the C oracle
on identical work: the 40,000-node mesh in the reference's own point-relaxation mode,
where the oracle equals the reference bit for bit (same GCR iteration counts). Each side
runs 2 and 12 time steps; the difference is the cost of 10 steps without input parsing
and set-up. The ratio calibrates the oracle-based CPU baseline that bench.py reports
(the reference binary and its input files are not available on the GPU host).

    python oracle/calibrate_cpu.py         (needs /root/reference to build oracle/_ref)
"""
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    sys.path.insert(0, p)

import subprocess  # noqa: E402

from make_golden import PARITY, _unlimited_stack, build_ref, small_config  # noqa: E402
from oracle import Oracle  # noqa: E402
from src_amd import mesh as M  # noqa: E402
from src_amd.refio import write_reference_inputs  # noqa: E402


def ref_time(exe, m, cfg, d, ntime):
    write_reference_inputs(d, m, cfg, ntime=ntime, **PARITY)
    t = time.perf_counter()
    r = subprocess.run([exe, "run"], cwd=d, capture_output=True, text=True, preexec_fn=_unlimited_stack)
    dt = time.perf_counter() - t
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
    return dt


def oracle_time(m, cfg, ntime):
    t = time.perf_counter()
    o = Oracle(m, cfg, sm_flag=PARITY["sm_flag"], itrp1=PARITY["itrp1"], omega=PARITY["omega"])
    o.init()
    o.step(ntime)
    dt = time.perf_counter() - t
    it = o.stats().nitsm
    o.close()
    return dt, it


def main():
    cfg = small_config()
    exe = build_ref(cfg)
    m = M.generate(cfg.mesh)
    with tempfile.TemporaryDirectory() as d:
        r2, r12 = ref_time(exe, m, cfg, d, 2), ref_time(exe, m, cfg, d, 12)
    (o2, i2), (o12, i12) = oracle_time(m, cfg, 2), oracle_time(m, cfg, 12)
    ref10, orc10 = r12 - r2, o12 - o2
    print(f"reference (flang -O3, 1 thread): {ref10:.2f} s per 10 steps")
    print(f"C oracle   (gcc -O2, 1 thread): {orc10:.2f} s per 10 steps ({i12 - i2} GCR iterations)")
    print(f"oracle / reference time ratio: {orc10 / ref10:.3f}")


if __name__ == "__main__":
    main()
