"""Code-generation check of the persistent batch kernels (CPU only: hipcc
cross-compiles engine.hip to gfx950 assembly, nothing runs on a GPU).

Each persistent loop fetches its next work item with an atomic and broadcasts
the index with v_readfirstlane; the loop exit compares that SGPR value with
the item count, so the whole wavefront leaves together.  If the broadcast of
the fetched index ever lands in a nested loop, the compiler has split the
persistent loop by lanes (lane 0 fetching in an outer loop while lanes 1-63
iterate an inner one with a stale index), which never terminates on the
device.  This test fails on that shape.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "rtdl_amd", "csrc", "engine.hip")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def _functions(asm: str) -> dict[str, list[str]]:
    funcs, cur = {}, None
    for line in asm.split("\n"):
        m = re.match(r"^(_Z\S+):", line)
        if m:
            cur = m.group(1)
            funcs[cur] = []
            continue
        if cur:
            funcs[cur].append(line)
            if "s_endpgm" in line:
                cur = None
    return funcs


def _fetch_broadcast_depths(body: list[str]) -> set[int]:
    """Loop depths of the v_readfirstlane broadcasts of a returning global
    atomic's result (the work-item fetch and the row reservation).  Other
    broadcasts (of values every lane already agrees on) may sit at any depth."""
    depths, depth, pending = set(), 0, {}
    for line in body:
        if re.match(r"^(\.LBB\d+_\d+|; %bb\.\d+):", line):
            d = re.search(r"Depth=(\d+)", line)
            depth = int(d.group(1)) if d else 0
        h = re.search(r"This (?:Inner )?Loop Header: Depth=(\d+)", line)
        if h:
            depth = int(h.group(1))
        a = re.match(r"\s*global_atomic_\w+\s+(v\d+),.*\b(sc0|glc)\b", line)
        if a:
            pending[a.group(1)] = True
            continue
        r = re.match(r"\s*v_readfirstlane_b32\s+s\d+,\s*(v\d+)", line)
        if r and pending.pop(r.group(1), False):
            depths.add(depth)
    return depths


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not available")
def test_persistent_loops_fetch_in_the_outermost_loop(tmp_path):
    out = tmp_path / "engine.s"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "--cuda-device-only", "-S", SRC, "-o", str(out)]
    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stderr[-2000:]
    funcs = _functions(out.read_text())
    checked = 0
    for name, body in funcs.items():
        if not re.search(r"batch_(points|rays|pip)_kernel|nbr_(count|emit|sort)_kernel", name):
            continue
        depths = _fetch_broadcast_depths(body)
        assert depths, f"{name}: no readfirstlane broadcast of the fetched item"
        assert max(depths) <= 1, f"{name}: item broadcast inside a nested loop (depths {sorted(depths)})"
        checked += 1
    assert checked >= 20
