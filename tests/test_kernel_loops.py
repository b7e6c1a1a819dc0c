"""Shape of the compiled persistent row loops (CPU: gfx950 assembly from
hipcc, no GPU needed).

A persistent kernel claims rows with one thread's global atomic, publishes
the index through LDS between two barriers and loops.  If the compiler
splits that loop in two (an outer loop holding the claim, an inner loop
holding the barriers that the other threads never leave), a workgroup spins
on its first row for ever: the windowed symbolic kernel compiled that way
while it had an early `continue`.  This test compiles every kernel source
and checks, in each function, that the first barrier after a claim lies in
the claim's own loop."""
import os
import re
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "research_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-munsafe-fp-atomics", "--cuda-device-only", "-S"]
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def loops(asm_text):
    """Per function: ordered (op, innermost loop) of barriers and global
    atomic adds, and each loop header's parent loop, from the loop comments
    the compiler writes beside the block labels."""
    funcs, cur = {}, None
    for line in asm_text.splitlines():
        m = re.match(r"^(_Z\w+):", line)
        if m:
            cur = {"items": [], "parent": {}, "loop": {}, "block": "entry"}
            funcs[m.group(1)] = cur
            continue
        if cur is None:
            continue
        if line.startswith(".Lfunc_end"):
            cur = None
            continue
        m = re.match(r"^\.L(BB\d+_\d+):(.*)$", line)
        if m:
            blk, com = m.group(1), m.group(2)
            cur["block"], cur["header_of"] = blk, None
            h = re.search(r"Header=(BB\d+_\d+)", com)
            cur["loop"][blk] = blk if "Loop Header" in com else (h.group(1) if h else None)
            p = re.search(r"Parent Loop (BB\d+_\d+)", com)
            cur["last_parent"] = p.group(1) if p else None
            if "Loop Header" in com:
                cur["parent"][blk] = cur["last_parent"]
            continue
        s = line.strip()
        if s.startswith(";"):
            p = re.search(r"Parent Loop (BB\d+_\d+)", s)
            if p:
                cur["last_parent"] = p.group(1)
            if "Loop Header" in s:
                cur["loop"][cur["block"]] = cur["block"]
                cur["parent"][cur["block"]] = cur["last_parent"]
            continue
        op = s.split()[0] if s else ""
        if (op == "s_barrier" or op.startswith("global_atomic_add") or op.startswith("ds_write")
                or op.startswith("global_store")):
            cur["items"].append((op, cur["loop"].get(cur["block"])))
    return funcs


def is_claim(its, i):
    """A row claim publishes its result through LDS: the first store after
    the atomic is an LDS write (an atomic that appends to a global list is
    followed by a global store instead)."""
    if not its[i][0].startswith("global_atomic_add"):
        return False
    for op, _ in its[i + 1:]:
        if op.startswith("ds_write"):
            return True
        if op.startswith("global_store") or op == "s_barrier":
            return False
    return False


def split_claims(funcs):
    bad = []
    for name, f in funcs.items():
        its = f["items"]
        for i, (op, lp) in enumerate(its):
            if not is_claim(its, i):
                continue
            nxt = [x for x in its[i + 1:] if x[0] == "s_barrier"]
            if not nxt:
                continue
            chain, x = [], nxt[0][1]
            while x:
                chain.append(x)
                x = f["parent"].get(x)
            if nxt[0][1] != lp and lp in chain:
                bad.append(f"{name}: claim in loop {lp}, its barriers in inner loop {nxt[0][1]}")
    return bad


def test_parser_flags_a_split_loop():
    """The checker itself, on a hand-written split loop and a sound one."""
    split = """_Zk:
.LBB0_1:                               ; =>This Loop Header: Depth=1
\tglobal_atomic_add v1, v2, v3, s[0:1] sc0
\tds_write_b32 v4, v1 offset:512
.LBB0_2:                               ;   Parent Loop BB0_1 Depth=1
                                        ; =>  This Loop Header: Depth=2
\ts_barrier
\ts_barrier
.LBB0_3:                               ;   in Loop: Header=BB0_1 Depth=1
.Lfunc_end0:
"""
    sound = split.replace("""                                        ; =>  This Loop Header: Depth=2
""", "").replace(";   Parent Loop BB0_1 Depth=1", ";   in Loop: Header=BB0_1 Depth=1")
    assert len(split_claims(loops(split))) == 1
    assert split_claims(loops(sound)) == []
    # an atomic append to a global list (followed by a global store) is not a claim
    append = split.replace("\tds_write_b32 v4, v1 offset:512", "\tglobal_store_dword v4, v1, s[2:3]")
    assert split_claims(loops(append)) == []


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_persistent_loops_not_split():
    with tempfile.TemporaryDirectory() as tmp:
        def build(src):
            out = os.path.join(tmp, src + ".s")
            subprocess.run([HIPCC] + FLAGS + ["-o", out, os.path.join(CSRC, src)], check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            return src, open(out).read()

        with ThreadPoolExecutor(4) as ex:
            asm = dict(ex.map(build, SOURCES))
    bad, claims = [], 0
    for src, text in asm.items():
        funcs = loops(text)
        claims += sum(1 for f in funcs.values() for i in range(len(f["items"]))
                      if is_claim(f["items"], i))
        bad += [f"{src}: {b}" for b in split_claims(funcs)]
    assert claims > 5, "no row claims found: the assembly format changed"
    assert not bad, "\n".join(bad)
