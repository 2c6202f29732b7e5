"""Times the reference's own CPU heat application (apps/heat_cpu, built from
its sources into oracle/_ref/heat_cpu by oracle/Makefile.ref) on the 64^3
heat workload: dt = eps_constant = 1e-3 (rho = 0.512 / 2.024 for 64 cells),
4 z-slabs (MPI ranks), 100 timed steps after 10 warm-up steps, for each
backend; then the 256^3 step (pascal, 8 z-slabs, 10 timed steps). heat_cpu's option = strong records per-step phase times (rhs, the
z, y and x solves, the ghost exchange) to the CSV named by TIMING_CSV
(apps/heat_cpu/solve_theta.cpp:356-377); this prints their medians and the
mean step time as JSON lines. Test infrastructure: the CPU baseline of the
heat step, run where /root/reference and MPICH are present.

    make -C oracle -f Makefile.ref && python oracle/time_heat_cpu.py
"""
from __future__ import annotations

import csv
import json
import os
import statistics
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
HEAT = os.path.join(HERE, "_ref", "heat_cpu")
MPIEXEC = "/opt/conda/bin/mpiexec"


def run(kind, n=64, parts=(1, 1, 4), nt=110, warmup=10, rho=0.512 / 2.024):
    px, py, pz = parts
    text = (f"nx = {n}\nny = {n}\nnz = {n}\nnpx = {px}\nnpy = {py}\nnpz = {pz}\n"
            f"rho = {rho!r}\neps = 0.001\noption = strong\nNt = {nt}\nwarmup = {warmup}\n"
            f"tdma_backend = {kind}\nperiodic = 0\n")
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, "in.txt")
        out = os.path.join(tmp, "timing.csv")
        with open(inp, "w") as f:
            f.write(text)
        env = dict(os.environ, TIMING_CSV=out, MPICH_CXX="g++")
        subprocess.run([MPIEXEC, "-n", str(px * py * pz), HEAT, inp], cwd=tmp, env=env,
                       capture_output=True, text=True, timeout=1200, check=True)
        with open(out) as f:
            rows = [r for r in csv.reader(f) if r and not r[0].startswith("#")]
    # long form: rank, t_step, event, time_sec; a phase's time in a step is
    # its slowest rank's
    per = {}
    for r in rows[1:]:
        key = (int(r[1]), r[2].strip())
        per[key] = max(per.get(key, 0.0), float(r[3]))
    steps = sorted({k[0] for k in per})
    events = sorted({k[1] for k in per})
    med = {e: statistics.median(per[(t, e)] for t in steps) * 1e3 for e in events}
    step = statistics.mean(sum(per[(t, e)] for e in events) for t in steps) * 1e3
    return {"backend": kind, "grid": f"{n}^3", "ranks": px * py * pz, "steps": len(steps),
            "median_ms": med, "mean_step_ms": step}


def main():
    if not os.path.exists(HEAT):
        sys.exit("reference heat build missing: make -C oracle -f Makefile.ref")
    for kind in ("pascal", "filtered", "filtered_v2"):
        print(json.dumps(run(kind)), flush=True)
    # the 256^3 step (same dt rule), 8 z-slabs, 10 timed steps
    print(json.dumps(run("pascal", n=256, parts=(1, 1, 8), nt=12, warmup=2)), flush=True)


if __name__ == "__main__":
    main()
