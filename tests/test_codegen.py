"""The device code the library is built from, read as gfx950 assembly (CPU
only: hipcc cross-compiles).

The register-resident kernels hold whole partitions in registers (up to 512
VGPRs and AGPRs per lane), and under that pressure the compiler moves values
between VGPRs and AGPRs. A move that must reach every lane has been placed
after the target label of an `s_cbranch_execz` but before the
`s_or_b64 exec, exec, s[..]` that restores the lanes of the enclosing
lane-divergent branch: it then runs only in the lanes that took the branch,
and the others keep a stale register (a learned-J guarded solve gave values
near 1e292 in the partition whose lanes skip the branch). The kernels are
written so that no such branch sits where it happens; this test compiles
every source with the library's flags and fails if any kernel has a
register move between such a label and its exec restore."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from filtered_tdma_amd import build as B

MOVES = re.compile(r"v_(accvgpr_write|accvgpr_read|mov|cndmask)\w*\s")


def moves_before_exec_restore(asm: str):
    """(kernel, line) of register moves between an s_cbranch_execz target
    and the s_or_b64 exec restore that follows it."""
    found = []
    for m in re.finditer(r"^(_Z\S+):.*?^\.Lfunc_end", asm, re.M | re.S):
        body = m.group(0).splitlines()
        targets = set(re.findall(r"s_cbranch_execz (\.LBB\d+_\d+)", m.group(0)))
        for k, line in enumerate(body):
            lab = re.match(r"^(\.LBB\d+_\d+):", line)
            if not lab or lab.group(1) not in targets:
                continue
            seg = []
            for nxt in body[k + 1:k + 80]:
                s = nxt.strip()
                if s.startswith("s_or_b64 exec, exec"):
                    found.extend((m.group(1), t) for t in seg)
                    break
                if s.startswith((".LBB", "s_cbranch", "s_branch", "s_endpgm")):
                    break
                if MOVES.match(s):
                    seg.append(s)
    return found


def test_scanner_flags_a_move_before_the_exec_restore():
    asm = """_Zk:
\ts_and_saveexec_b64 s[4:5], vcc
\ts_cbranch_execz .LBB1_2
\tv_add_f64 v[0:1], v[0:1], v[2:3]
.LBB1_2:
\tv_accvgpr_write_b32 a4, v5
\ts_or_b64 exec, exec, s[4:5]
\tv_accvgpr_write_b32 a6, v7
.Lfunc_end0:
"""
    assert moves_before_exec_restore(asm) == [("_Zk", "v_accvgpr_write_b32 a4, v5")]
    assert moves_before_exec_restore(asm.replace("\tv_accvgpr_write_b32 a4, v5\n", "")) == []


@pytest.mark.skipif(not shutil.which(B.HIPCC) and not os.path.exists(B.HIPCC),
                    reason="hipcc is not installed")
def test_no_register_move_runs_under_a_branch_mask_at_a_join():
    flags = [f for f in B.FLAGS if f != "-shared"]
    sources = [s for s in B.SOURCES if s.endswith(".hip")]
    with tempfile.TemporaryDirectory() as tmp:
        procs = []
        for src in sources:
            out = os.path.join(tmp, os.path.basename(src) + ".s")
            cmd = [B.HIPCC, *flags, "--cuda-device-only", "-S", "-x", "hip", src, "-o", out]
            procs.append((out, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT, text=True)))
        bad = []
        for out, pr in procs:
            log, _ = pr.communicate(timeout=900)
            assert pr.returncode == 0, log[-2000:]
            with open(out) as f:
                bad += moves_before_exec_restore(f.read())
    assert not bad, bad[:10]
