"""The heat ADI harness on the GPU (ftd_heat_* through the C ABI) against the
oracle's restatement of apps/heat_cpu, and the hipGraph replay against the
eager steps.

Bit-for-bit where every sweep runs the same partitioning in both: the
directions the caller splits, and undivided ones that fit the single-lane
Thomas kernel (n <= 64; longer undivided lines run in internal partitions on
the GPU). For PaScaL the L2 text must equal what the reference heat_cpu
prints for the same case (tests/golden/heat_l2.json)."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from filtered_tdma_amd.heat import SolveTheta, dt_policy

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "heat_l2.json")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def golden():
    with open(GOLDEN) as f:
        return [c for c in json.load(f) if c["n"] <= 64]


@pytest.mark.parametrize("case", golden(),
                         ids=lambda c: f"{c['kind']}-n{c['n']}-r{c['rho']}-"
                                       f"{'x'.join(map(str, c['parts']))}-per{c['periodic']}")
def test_heat_gpu_matches_oracle_and_reference_l2(case):
    """Field bit-for-bit equal to the oracle's GPU-semantics restatement
    (every sweep runs the reference's partitioning). The golden L2 was
    printed by the reference's CPU heat_cpu: PaScaL is exact, so its L2 text
    must match too; the filtered kinds choose J differently on the CPU and
    the GPU (filtered_tdma.cpp:63-115 vs filtered_tdma_kernels.cuh:220-261),
    so for them the CPU-semantics oracle is what matches heat_cpu (checked
    on the CPU) and the GPU is held to the GPU-semantics oracle."""
    n, parts, kind = case["n"], tuple(case["parts"]), case["kind"]
    per = (1, 1, 1) if case["periodic"] else (0, 0, 0)
    dt, nt = dt_policy("order", case["rho"], n)
    note(f"heat {kind} n={n} parts={parts} periodic={per} dt={dt} Nt={nt}")
    sim = SolveTheta(n, n, n, parts=parts, backend=kind, dt=dt, periodic=per)
    try:
        sim.run(nt)
        l2 = sim.l2()
        field = sim.field()
    finally:
        sim.close()
    l2o, ref = O.heat_run(n, n, n, parts, kind, dt, nt, periodic=per, sem=O.SEM_GPU,
                          return_theta=True)
    inner = (slice(1, -1),) * 3
    np.testing.assert_array_equal(field[inner], ref[inner])
    assert l2 == l2o
    if kind == "pascal":
        assert "%g" % l2 == case["l2"]


@pytest.mark.parametrize("n", [100, 128])
def test_heat_gpu_matches_oracle_at_larger_sizes(n):
    """Every direction split by the caller (P = 2, so the GPU runs the
    oracle's partitioning): the field equals the oracle's bit for bit at
    sizes whose right-hand-side runs are 16 (100^3) and 32 (128^3) planes
    long and whose x tiles are halved to cover the CUs (100^3)."""
    parts = (2, 2, 2)
    dt = dt_policy("order", 0.3, n)[0]
    note(f"heat pascal n={n} parts={parts} against the oracle")
    sim = SolveTheta(n, n, n, parts=parts, backend="pascal", dt=dt)
    try:
        sim.run(3, use_graph=True)
        l2 = sim.l2()
        field = sim.field()
    finally:
        sim.close()
    l2o, ref = O.heat_run(n, n, n, parts, "pascal", dt, 3, sem=O.SEM_GPU, return_theta=True)
    inner = (slice(1, -1),) * 3
    np.testing.assert_array_equal(field[inner], ref[inner])
    assert l2 == l2o


@pytest.mark.parametrize("kind,parts,periodic", [
    ("filtered", (1, 1, 4), (0, 0, 0)),
    ("filtered_v2", (2, 2, 2), (0, 0, 0)),
    ("pascal", (1, 1, 1), (1, 1, 1)),
])
def test_heat_graph_replay_equals_eager(kind, parts, periodic):
    """A step captured into a hipGraph and replayed gives the eager steps'
    field bit for bit, replay after replay, with eager work in between."""
    n, steps = 48, 6
    dt = dt_policy("order", 0.3, n)[0]
    note(f"heat graph {kind} parts={parts} periodic={periodic}")
    eager = SolveTheta(n, n, n, parts=parts, backend=kind, dt=dt, periodic=periodic)
    graph = SolveTheta(n, n, n, parts=parts, backend=kind, dt=dt, periodic=periodic)
    try:
        for s in range(steps):
            eager.run(1, use_graph=False)
            graph.run(1, use_graph=True)
            torch.cuda.synchronize()
            _ = torch.ones(1024, device="cuda").sum().item()  # unrelated eager work
            np.testing.assert_array_equal(graph.field(), eager.field())
    finally:
        eager.close()
        graph.close()


def test_heat_config1_filtered_vs_pascal_vs_thomas():
    """The 64^3, 100-step, dt = eps_constant = 1e-3, 4 z-slab case:
    filtered and filtered_v2 against pascal and single-partition Thomas,
    and each against the oracle's restatement."""
    n, dt, steps = 64, 1e-3, 100
    res = {}
    for kind, parts in (("pascal", (1, 1, 4)), ("filtered", (1, 1, 4)),
                        ("filtered_v2", (1, 1, 4)), ("pascal", (1, 1, 1))):
        note(f"heat config1 {kind} parts={parts}")
        sim = SolveTheta(n, n, n, parts=parts, backend=kind, dt=dt)
        try:
            sim.run(steps, use_graph=True)
            res[(kind, parts)] = (sim.l2(), sim.field())
        finally:
            sim.close()
        l2o = O.heat_run(n, n, n, parts, kind, dt, steps, sem=O.SEM_GPU)
        assert res[(kind, parts)][0] == l2o
    thomas = res[("pascal", (1, 1, 1))][1]
    inner = (slice(1, -1),) * 3
    for key, (l2, f) in res.items():
        assert np.max(np.abs(f[inner] - thomas[inner])) <= 1e-6, key


@pytest.mark.parametrize("kind,n,parts,periodic", [
    ("pascal", 48, (1, 1, 1), (0, 0, 0)),     # Thomas sweeps (N <= 64)
    ("pascal", 100, (1, 1, 1), (0, 0, 0)),    # internal PaScaL partitions (N > 64)
    ("pascal", 64, (2, 4, 8), (0, 0, 0)),     # PaScaL with the caller's partitions
    ("pascal", 40, (1, 2, 1), (1, 0, 0)),     # periodic x: materialized there, fused in y, z
    ("pascal", 37, (1, 1, 3), (0, 0, 0)),     # odd sizes, uneven partitions
    # filtered kinds: x and y (one partition) fused as the exact solve; z over
    # the caller's partitions by the fused filtered sweep (k_adi_filt)
    ("filtered", 100, (1, 1, 4), (0, 0, 0)),      # 25 rows per partition (M = 32)
    ("filtered_v2", 48, (1, 1, 4), (0, 0, 0)),    # 12 rows (M = 16)
    ("filtered_v2", 200, (1, 1, 4), (0, 0, 0)),   # 50 rows (M = 64)
    ("filtered", 37, (1, 1, 3), (0, 0, 0)),       # uneven 13/12/12
    ("filtered_v2", 64, (2, 2, 8), (0, 0, 0)),    # 8 z partitions; x, y materialized
    ("filtered", 150, (1, 1, 16), (0, 0, 0)),     # 16 z partitions of 9-10 rows
    ("filtered", 40, (1, 2, 1), (0, 0, 0)),       # y over 2 partitions, x and z exact
    ("filtered_v2", 48, (1, 4, 1), (0, 0, 0)),    # y over 4 partitions of 12 rows
    ("filtered", 70, (1, 3, 2), (0, 0, 0)),       # y 24/23/23 rows and z 35 rows (M = 64)
    ("filtered_v2", 48, (3, 2, 2), (0, 0, 0)),    # x split: materialized there
    ("filtered_v2", 48, (1, 1, 4), (0, 0, 1)),    # periodic z: materialized
    ("filtered", 48, (1, 4, 1), (0, 1, 0)),       # periodic y: materialized
])
def test_heat_fused_sweeps_equal_materialized(kind, n, parts, periodic):
    """Fused sweeps (coefficients generated in the kernel, rhs and x in the
    field) give the materialized build + solve + scatter field bit for bit,
    eagerly and from a graph. The filtered kinds fuse the directions of one
    caller partition, where their solve is the exact one, and their z sweep
    over the caller's z partitions (after the first step, which sets the
    solver's column-0 bounds)."""
    dt = dt_policy("order", 0.3, n)[0]
    note(f"heat fused vs materialized {kind} n={n} parts={parts} periodic={periodic}")
    sims = [SolveTheta(n, n, n, parts=parts, backend=kind, dt=dt, periodic=periodic)
            for _ in range(2)]
    try:
        sims[1].set_fused(False)
        for use_graph in (False, True):
            for sim in sims:
                sim.run(3, use_graph=use_graph)
            np.testing.assert_array_equal(sims[0].field(), sims[1].field())
        plan = sims[0].last_plan()
        for d, name in ((1, "y"), (2, "z")):
            fused_filtered = kind != "pascal" and parts[d] > 1 and not periodic[d]
            assert (f"{name}:filtered_fused" in plan) == fused_filtered, plan
        if kind != "pascal":  # the filtered kinds' split x stays materialized
            assert ("x:materialized" in plan) == (parts[0] > 1 or bool(periodic[0])), plan
        assert "fused" not in sims[1].last_plan(), sims[1].last_plan()
    finally:
        for sim in sims:
            sim.close()


@pytest.mark.parametrize("n,plans", [
    (100, {0: (32, 16), 1: (16, 0), 2: (64, 0)}),
    (128, {0: (64, 0), 1: (64, 64), 2: (16, 16)}),
    (256, {}),                                  # the automatic plans at the ADI workload
])
def test_heat_plans_keep_fused_equal_materialized(n, plans):
    """Each sweep shape (rows per lane, systems per workgroup) set through
    ftd_heat_set_plan: the fused field equals the materialized one, which
    runs the same partitioning, bit for bit."""
    dt = 1e-4
    note(f"heat plans n={n} plans={plans}")
    sims = [SolveTheta(n, n, n, parts=(1, 1, 1), backend="pascal", dt=dt) for _ in range(2)]
    try:
        for sim in sims:
            for d, (m, s) in plans.items():
                sim.set_plan(d, m, s)
        sims[1].set_fused(False)
        for sim in sims:
            sim.run(2, use_graph=True)
        np.testing.assert_array_equal(sims[0].field(), sims[1].field())
    finally:
        for sim in sims:
            sim.close()


def test_heat_256_adi_step_matches_oracle():
    """The 256^3 ADI step of the uneven-batch/ADI workload with the automatic
    plans (one caller partition per direction; lines of 256 rows solved in
    8 internal partitions of 32 rows, x tiles and right-hand-side runs sized
    for the device), three steps from a hipGraph, against the oracle with
    the same partitioning (parts = (8, 8, 8)): field and L2 bit for bit."""
    n, dt, steps = 256, 1e-4, 3
    note(f"heat 256^3 automatic plans vs oracle parts=(8, 8, 8), {steps} graph steps")
    sim = SolveTheta(n, n, n, parts=(1, 1, 1), backend="pascal", dt=dt)
    try:
        sim.run(steps, use_graph=True)
        l2 = sim.l2()
        field = sim.field()
    finally:
        sim.close()
    l2o, ref = O.heat_run(n, n, n, (8, 8, 8), "pascal", dt, steps, sem=O.SEM_GPU,
                          return_theta=True, nthreads=16)
    inner = (slice(1, -1),) * 3
    np.testing.assert_array_equal(field[inner], ref[inner])
    assert l2 == l2o
