"""Wait states around the hand-issued DPP updates of the tile LU, and the order of
the device-scope releases that publish tiles and rows.

The diagonal-block elimination issues its row-broadcast FMAs as asm statements
(v_fmac_*_dpp row_newbcast) and hands values between the wave's halves with
v_permlane32_swap, in asm statements the compiler's hazard recognizer does not see
through.  A DPP source VGPR, or either operand of a v_permlane32_swap, must not be
written by a VALU instruction in the two wait states before, and EXEC must not be
written by a VALU instruction in the five before a DPP instruction.  This test compiles the factor kernels to gfx950
assembly and checks every DPP instruction in it against those two rules, counting
one wait state per instruction and N + 1 per s_nop N, and treating a label inside
the window as a violation (the branch into it could come from anywhere).

A release writes the L2 back (buffer_wbl2) and must wait for that write-back
(s_waitcnt vmcnt(0)) before the flag store that follows it; the compiler's own
release omitted that wait where no store of the wave was pending, so the flag
could be seen before the tiles.  The second check compiles the solve kernels and
requires such a wait after every buffer_wbl2, before the next vector store or
atomic and before any label.

The persistent factor kernels publish without a write-back and consume without a
per-task invalidate: every tile store is write-through (global_store ... sc1), each
wave waits for its own stores (s_waitcnt vmcnt(0), marked "publish: stores done")
before the barrier that comes ahead of the flag stores, and the only cache invalidate
is the one each workgroup makes when it starts (no tile of the run can be cached
before it is made, since every slot is written once and read only after its flag; a
factor that reuses pool slots runs the kernels built to read tiles with coherent
loads instead).
The third check reads those rules off the compiled persistent kernels (no
buffer_wbl2, one buffer_inv before the first barrier, stores write-through), and that
no kernel is built for TgSplit mode (where a workgroup's waves may sit on different CUs
and the barrier would not order their stores through one CU's memory pipeline).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sparse-operator-graph-lu_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
REG = re.compile(r"\bv(\d+)\b|\bv\[(\d+):(\d+)\]")
PUBLISH = "publish: stores done"  # the comment on the factor's per-wave publish wait


def vregs(text):
    out = set()
    for m in REG.finditer(text):
        if m.group(1) is not None:
            out.add(int(m.group(1)))
        else:
            out.update(range(int(m.group(2)), int(m.group(3)) + 1))
    return out


def instructions(asm, keep_comments=False):
    out = []
    for line in asm.splitlines():
        s = line.strip() if keep_comments and PUBLISH in line else line.split(";")[0].strip()
        if not s or s.startswith(";"):
            continue
        if s.endswith(":"):  # block labels (.LBB...) start with a dot too
            out.append(("label", s))
        elif not s.startswith("."):
            out.append(("inst", s))
    return out


def dpp_violations(asm, vgpr_states=2, exec_states=5):
    insts = instructions(asm)
    bad, ndpp = [], 0
    for idx, (kind, s) in enumerate(insts):
        if kind != "inst":
            continue
        op = s.split()[0]
        swap = op.startswith("v_permlane")  # reads (and writes) both operands
        if "_dpp" not in op and "row_newbcast" not in s and not swap:
            continue
        ndpp += 1
        opnds = s.split(None, 1)[1].split(",")
        src0 = vregs(opnds[1]) | (vregs(opnds[0]) if swap else set())
        waits, j = 0, idx - 1
        while j >= 0 and waits < exec_states:
            k2, t = insts[j]
            if k2 == "label":
                if waits < vgpr_states:
                    bad.append((s, "label %d wait states before" % waits))
                break
            top = t.split()[0]
            if top == "s_nop":
                waits += int(t.split()[1], 0) + 1
                j -= 1
                continue
            if top.startswith("v_"):
                dst = t.split(None, 1)[1].split(",")[0] if " " in t else ""
                if top.startswith("v_cmpx") or dst.strip() == "exec":
                    bad.append((s, "EXEC written %d wait states before by %s" % (waits, t)))
                elif (waits < vgpr_states and not top.startswith(("v_cmp", "v_readlane", "v_readfirstlane"))
                      and vregs(dst) & src0):
                    bad.append((s, "source written %d wait states before by %s" % (waits, t)))
            waits += 1
            j -= 1
    return ndpp, bad


def test_checker_finds_planted_hazards():
    asm = "\n".join([
        "f:", "\ts_nop 4",
        "\tv_fmac_f32_dpp v27, v27, v53 row_newbcast:4 row_mask:0xf bank_mask:0xf",
        "\tv_mov_b32_dpp v51, v27 row_newbcast:5 row_mask:0xf bank_mask:0xf",   # 0 states after v27's write
        "\tv_add_f64 v[4:5], v[0:1], v[2:3]",
        "\tv_add_f32_e32 v9, v1, v2",
        "\tv_fmac_f64_dpp v[4:5], v[4:5], v[0:1] row_newbcast:4 row_mask:0xf bank_mask:0xf",  # 1 state
        "\tv_mov_b32_e32 v8, v1", "\ts_nop 1",
        "\tv_fmac_f32_dpp v8, v8, v53 row_newbcast:4 row_mask:0xf bank_mask:0xf",  # 2 states: fine
        "\tv_cmpx_lt_f32_e32 vcc, v1, v2", "\ts_nop 1",
        "\tv_mov_b32_dpp v50, v7 row_newbcast:1 row_mask:0xf bank_mask:0xf",  # EXEC 2 states before
        "\tv_cndmask_b32_e64 v60, v1, v2, s[4:5]", "\ts_nop 0",
        "\tv_permlane32_swap_b32_e32 v61, v62",  # v60 written 1 state before: not an operand
        "\tv_cndmask_b32_e64 v63, v1, v2, s[4:5]",
        "\tv_permlane32_swap_b32_e32 v63, v64",  # its first operand written just before
    ])
    n, bad = dpp_violations(asm)
    assert n == 7
    assert len(bad) == 4
    assert "v27" in bad[0][1] and "v[4:5]" in bad[1][1] and "EXEC" in bad[2][1] and "v63" in bad[3][1]


def release_violations(asm):
    """Every buffer_wbl2 must be followed by s_waitcnt vmcnt(0) (or a wait covering it)
    before the next vector store / atomic and before any label."""
    insts = instructions(asm)
    bad, n = [], 0
    for idx, (kind, s) in enumerate(insts):
        if kind != "inst" or not s.startswith("buffer_wbl2"):
            continue
        n += 1
        for k2, t in insts[idx + 1:]:
            if k2 == "label":
                bad.append((s, "label before the wait: " + t))
                break
            if t.startswith("s_waitcnt") and re.search(r"vmcnt\(0\)", t):
                break
            if re.match(r"(global|buffer|flat)_(store|atomic)", t):
                bad.append((s, "store before the wait: " + t))
                break
    return n, bad


def test_release_checker_finds_planted_mistakes():
    asm = "\n".join([
        "f:", "\tbuffer_wbl2 sc1", "\ts_waitcnt vmcnt(0)", "\tglobal_store_dwordx2 v[6:7], v[4:5], off sc1",
        "\tbuffer_wbl2 sc1", "\tglobal_store_dwordx2 v[6:7], v[4:5], off sc1",          # flag first
        "\tbuffer_wbl2 sc1", "\ts_waitcnt lgkmcnt(0)", ".L1:", "\ts_waitcnt vmcnt(0)",  # label first
        "\tbuffer_wbl2 sc1", "\ts_waitcnt vmcnt(0) lgkmcnt(0)", "\tglobal_atomic_add v1, v2, s[0:1]",
    ])
    n, bad = release_violations(asm)
    assert n == 4 and len(bad) == 2
    assert "store" in bad[0][1] and "label" in bad[1][1]


def _compile(tmp_path, name):
    out = tmp_path / (name + ".s")
    cmd = [HIPCC, "-std=c++17", "-O3", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
           "--cuda-device-only", "-S", "-o", str(out), os.path.join(PKG, "csrc", "device", name + ".hip")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return out.read_text()


@pytest.fixture(scope="module")
def factor_asm(tmp_path_factory):
    return _compile(tmp_path_factory.mktemp("asm"), "factor")


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not installed")
def test_releases_wait_for_the_write_back(tmp_path):
    asm = _compile(tmp_path, "solve")
    n, bad = release_violations(asm)
    assert n > 0
    assert not bad, bad[:5]


def kernel_bodies(asm, pattern):
    """{symbol: instruction list} of the functions whose symbol contains pattern."""
    lines = asm.split("\n")
    starts = [i for i, ln in enumerate(lines) if re.match(r"^_Z\S+:(\s|$)", ln)]
    out = {}
    for k, st in enumerate(starts):
        name = lines[st].split(":")[0]
        if pattern in name:
            end = starts[k + 1] if k + 1 < len(starts) else len(lines)
            out[name] = instructions("\n".join(lines[st + 1:end]), keep_comments=True)
    return out


def publish_violations(insts):
    """Rules of the write-through publish on one kernel's instructions; returns
    (number of publish waits, list of violations)."""
    bad, n = [], 0
    for idx, (kind, s) in enumerate(insts):
        if kind != "inst":
            continue
        if s.startswith("buffer_wbl2"):
            bad.append((s, "an L2 write-back in a write-through kernel"))
        if re.match(r"(global|buffer|flat)_store", s) and not re.search(r"\bsc1\b", s):
            bad.append((s, "a store that is not write-through"))
        if s.startswith("buffer_inv"):
            before = [t for k2, t in insts[:idx] if k2 == "inst"]
            if any(t.startswith("s_barrier") for t in before) or any(
                    re.match(r"(global|buffer|flat)_load", t) for t in before):
                bad.append((s, "a cache invalidate after the kernel's start"))
        if PUBLISH in s:
            n += 1
            if not s.startswith("s_waitcnt") or "vmcnt(0)" not in s:
                bad.append((s, "publish wait is not a full vmcnt wait"))
            for k2, t in insts[idx + 1:]:
                if k2 == "inst" and t.startswith("s_barrier"):
                    break
                if k2 == "inst" and re.match(r"(global|buffer|flat)_(store|atomic)", t):
                    bad.append((s, "store between the publish wait and its barrier: " + t))
                    break
            else:
                bad.append((s, "no barrier after the publish wait"))
    return n, bad


def test_publish_checker_finds_planted_mistakes():
    insts = instructions("\n".join([
        "\tglobal_store_dword v1, v2, s[0:1] sc1",
        "\ts_waitcnt vmcnt(0) ; publish: stores done", "\ts_barrier",             # fine
        "\tglobal_store_dword v1, v2, s[0:1]",                                    # not write-through
        "\ts_waitcnt vmcnt(0) ; publish: stores done",
        "\tglobal_store_dword v1, v2, s[0:1] sc1", "\ts_barrier",               # store after the wait
        "\tglobal_load_dword v3, v1, s[0:1] sc1", "\tbuffer_inv sc1",           # late invalidate
        "\tbuffer_wbl2 sc1",                                                     # write-back
    ]), keep_comments=True)
    n, bad = publish_violations(insts)
    assert n == 2
    assert [b[1][:12] for b in bad] == ["a store that", "store betwee", "a cache inva", "an L2 write-"]
    n, bad = publish_violations(instructions("\tbuffer_inv sc1\n\ts_barrier\n\tglobal_load_dword v3, v1, s[0:1]"))
    assert not bad  # the start-of-kernel invalidate


def test_factor_publish_is_write_through(factor_asm):
    bodies = kernel_bodies(factor_asm, "k_factor_persistent")
    assert len(bodies) >= 12
    for name, insts in bodies.items():
        n, bad = publish_violations(insts)
        assert n >= 1, name
        assert not bad, (name, bad[:5])
        assert sum(1 for k, s in insts if k == "inst" and s.startswith("buffer_inv")) == 1, name
    # no kernel of the library is built for TgSplit mode
    modes = re.findall(r"\.amdhsa_tg_split\s+(\d+)", factor_asm)
    assert modes and set(modes) == {"0"}


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not installed")
def test_factor_kernels_keep_dpp_wait_states(factor_asm):
    asm = factor_asm
    n, bad = dpp_violations(asm)
    assert asm.count("v_fmac_f32_dpp") > 0 and asm.count("v_fmac_f64_dpp") > 0
    assert n > 1000
    assert not bad, bad[:5]
