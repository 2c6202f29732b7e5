"""Writes tests/golden/heat_l2.json: the L2 errors printed by the reference's
own CPU heat application (apps/heat_cpu built unmodified into
oracle/_ref/heat_cpu by oracle/Makefile.ref) for small option=order runs.

    make -C oracle -f Makefile.ref && python tests/golden/make_heat_golden.py

Each case is an input file in heat_cpu's key = value format; the stored
value is the text heat_cpu prints after "Global L2 error =" (6 significant
digits, std::cout default).
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HEAT = os.path.join(ROOT, "oracle", "_ref", "heat_cpu")
MPIEXEC = "/opt/conda/bin/mpiexec"
OUT = os.path.join(HERE, "heat_l2.json")


def cases():
    out = []
    for n in (32,):
        for rho in (0.25, 0.40):
            for kind in ("pascal", "filtered", "filtered_v2"):
                for parts in ((1, 1, 1), (1, 1, 2), (1, 1, 4), (2, 2, 2), (1, 2, 1), (4, 1, 1)):
                    out.append(dict(n=n, rho=rho, kind=kind, parts=parts, periodic=0))
    for kind in ("pascal", "filtered", "filtered_v2"):
        for parts in ((1, 1, 1), (1, 1, 4), (2, 2, 2)):
            out.append(dict(n=64, rho=0.25, kind=kind, parts=parts, periodic=0))
            out.append(dict(n=32, rho=0.25, kind=kind, parts=parts, periodic=1))
    return out


def run(case):
    n, (px, py, pz) = case["n"], case["parts"]
    text = (f"nx = {n}\nny = {n}\nnz = {n}\nnpx = {px}\nnpy = {py}\nnpz = {pz}\n"
            f"rho = {case['rho']}\neps = 0.005\noption = order\n"
            f"tdma_backend = {case['kind']}\nperiodic = {case['periodic']}\n")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "in.txt")
        with open(path, "w") as f:
            f.write(text)
        res = subprocess.run([MPIEXEC, "-n", str(px * py * pz), HEAT, path], cwd=tmp,
                             capture_output=True, text=True, timeout=600, check=True)
    for line in res.stdout.splitlines():
        if "Global L2 error" in line:
            return line.split("=")[-1].strip()
    raise RuntimeError("no L2 line in heat_cpu output:\n" + res.stdout)


def main():
    if not os.path.exists(HEAT):
        sys.exit("reference heat build missing: make -C oracle -f Makefile.ref")
    table = []
    for c in cases():
        c = dict(c, l2=run(c))
        print(c, flush=True)
        table.append(c)
    with open(OUT, "w") as f:
        json.dump(table, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
