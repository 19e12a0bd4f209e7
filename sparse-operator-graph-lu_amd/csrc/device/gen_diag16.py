#!/usr/bin/env python3
"""Writes diag16_asm.h: the 16 elimination steps of a 16x16 diagonal block (FP32)
as one scheduled gfx950 asm statement.

Why asm: each step's 16 updates are v_fmac_f32_dpp (the row broadcast of lane k
folded into the FMA), which the compiler does not form by itself, and an asm
statement counts as zero wait states to the compiler's hazard recognizer, so
compiler code interleaved with such statements is padded with s_nop (an s_nop 1
costs about as much as two VALU instructions on one wave).  Here every step is
scheduled by hand and software-pipelined: while step k's 16 updates issue, the
pivot chain of step k + 1 (broadcast, clamp, reciprocal, multiplier and selects)
runs between them.

Per lane, as in the C++ form of the same routine: x[c] is row i (lanes 0-15,
32-63) or column i (lanes 16-31) of the block, y[c] the matching row of L^-1 or
(U^T)^-1, li = lane & 15.  Step k, with p = clamp(x[k] of lane k):
    f = x[k] / p (as x[k] * rcp(p));  nf = (li > k) ? -f : -0
    x[k] = (li > k) ? f : x[k];  pd = (li == k) ? p : pd
    x[c] += bcast_k(x[c]) * nf  (c > k);  y[c] += bcast_k(y[c]) * nf  (c <= k)
The caller stores the clamped pivot pd over the diagonal and takes 1 / pd (the
same rcp of the same value) as the scale of the (U^T)^-1 rows.  Lanes 16-31 take
x[k] = f as well: those values are dead after step k.

Wait states kept (gfx950, one wave): a DPP source register is not written by a
VALU instruction in the two preceding slots; an SGPR lane mask written by v_cmp is
read by v_cndmask no earlier than two slots later; a v_rcp result is consumed no
earlier than one slot later.  tests/test_dpp_hazards.py checks the first rule on
the compiled kernels.

Run:  python3 gen_diag16.py  (rewrites diag16_asm.h next to this file)
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def dpp(k):
    return "row_newbcast:%d row_mask:0xf bank_mask:0xf" % k


def fmac(reg, nf, k):
    return "v_fmac_f32_dpp %%[%s], %%[%s], %%[%s] %s" % (reg, reg, nf, dpp(k))


def chain(j, chol):
    """Pivot chain of step j as (instruction, min slots after previous chain item
    it depends on) groups; returned as a list of instructions in issue order with
    'GAP' markers where at least one other instruction must come between."""
    nf = "na" if j % 2 == 0 else "nb"
    c = {}
    c["G"] = "v_cmp_lt_u32_e64 %%[sg], %d, %%[li]" % j
    c["E"] = "v_cmp_eq_u32_e64 %%[se], %d, %%[li]" % j
    c["P"] = "v_mov_b32_dpp %%[P], %%[x%d] %s" % (j, dpp(j))
    if chol:
        c["CS"] = "v_cmp_gt_f32_e64 %[ss], %[T], %[P]"
        c["PC"] = "v_cndmask_b32_e64 %[Pc], %[P], %[T], %[ss]"
    else:
        c["CN"] = "v_cmp_gt_f32_e64 %[sn], 0, %[P]"
        c["CS"] = "v_cmp_lt_f32_e64 %[ss], |%[P]|, %[T]"
        c["TT"] = "v_cndmask_b32_e64 %[t], %[T], -%[T], %[sn]"
        c["PC"] = "v_cndmask_b32_e64 %[Pc], %[P], %[t], %[ss]"
    c["R"] = "v_rcp_f32_e32 %[rp], %[Pc]"
    c["F"] = "v_mul_f32_e32 %%[f], %%[x%d], %%[rp]" % j
    c["NF"] = "v_cndmask_b32_e64 %%[%s], %%[nz], -%%[f], %%[sg]" % nf
    c["X"] = "v_cndmask_b32_e64 %%[x%d], %%[x%d], %%[f], %%[sg]" % (j, j)
    c["PD"] = "v_cndmask_b32_e64 %[pd], %[pd], %[Pc], %[se]"
    return c


def schedule(fm, ch, chol):
    """Interleave the 16 updates fm (fm[0] writes the next pivot column) with the
    chain ch.  Slots are filled from fm where a wait state is needed; s_nop only
    when fm has run out."""
    out = []
    pending = list(fm)

    def filler(n):
        for _ in range(n):
            out.append(pending.pop(0) if pending else "s_nop 0")

    if fm:
        out.append(pending.pop(0))       # the next pivot column, updated first
    if ch is None:
        out.extend(pending)
        return out
    out.append(ch["G"])
    out.append(ch["E"])                  # two slots after the pivot column's write
    out.append(ch["P"])
    if chol:
        out.append(ch["CS"])
        filler(2)                        # v_cmp -> v_cndmask: two slots
        out.append(ch["PC"])
    else:
        out.append(ch["CN"])
        out.append(ch["CS"])
        filler(1)                        # CN -> TT: CS and one more
        out.append(ch["TT"])
        filler(1)                        # CS -> PC: filler, TT, filler
        out.append(ch["PC"])
    out.append(ch["R"])
    filler(1)                            # v_rcp -> use: one slot
    out.append(ch["F"])
    out.append(ch["NF"])                 # sg written long before
    out.append(ch["X"])
    out.append(ch["PD"])
    out.extend(pending)
    return out


def steps(chol):
    lines = ["s_nop 1"]                  # x may have been written by VALU just before
    lines += schedule([], chain(0, chol), chol)
    for k in range(16):
        nf = "na" if k % 2 == 0 else "nb"
        fm = [fmac("x%d" % c, nf, k) for c in range(k + 1, 16)] + [fmac("y%d" % c, nf, k) for c in range(0, k + 1)]
        lines += schedule(fm, chain(k + 1, chol) if k < 15 else None, chol)
    return lines


def check(lines, chol):
    """Re-verify the schedule's wait states on the generated text (independent of
    how schedule() placed them)."""
    import re
    def regs_written(l):
        m = re.match(r"(v_\w+)\s+%\[(\w+)\]", l)
        return {m.group(2)} if m else set()
    def dpp_src(l):
        m = re.match(r"v_(?:fmac|mov)_\w+_dpp\s+%\[\w+\],\s+%\[(\w+)\]", l)
        return m.group(1) if m else None
    def sgpr_read(l):
        return set(re.findall(r"%\[(s[ngse])\]", l.split(",", 1)[1])) if l.startswith("v_cndmask") else set()
    def slots(l):
        if l.startswith("s_nop"):
            return int(l.split()[1]) + 1
        return 1
    for i, l in enumerate(lines):
        src = dpp_src(l)
        if src:
            w = 0
            for j in range(i - 1, -1, -1):
                if w >= 2:
                    break
                if src in regs_written(lines[j]) and not lines[j].startswith("s_nop"):
                    raise AssertionError("DPP source %s written %d slots before: %s / %s" % (src, w, lines[j], l))
                w += slots(lines[j])
        for s in sgpr_read(l):
            w = 0
            for j in range(i - 1, -1, -1):
                m = re.match(r"v_cmp\w*\s+%\[(\w+)\]", lines[j])
                if m and m.group(1) == s:
                    if w < 2:
                        raise AssertionError("mask %s read %d slots after v_cmp: %s" % (s, w, l))
                    break
                w += slots(lines[j])
        if l.startswith("v_mul_f32") and "%[rp]" in l:
            if lines[i - 1].startswith("v_rcp"):
                raise AssertionError("rcp result used in the next slot")


def emit():
    out = ["// Generated by gen_diag16.py; do not edit.  The 16 elimination steps of a 16x16",
           "// diagonal block in FP32, scheduled for gfx950 (the generator's docstring gives",
           "// the per-lane algorithm and the wait-state rules the schedule keeps).",
           "#pragma once", ""]
    for chol in (False, True):
        lines = steps(chol)
        check(lines, chol)
        body = "\n".join('        "%s\\n\\t"' % l for l in lines)
        xs = ", ".join('[x%d] "+v"(x[%d])' % (c, c) for c in range(16))
        ys = ", ".join('[y%d] "+v"(y[%d])' % (c, c) for c in range(16))
        out += ["template <>",
                "__device__ __forceinline__ void diag16_elim_f32<%s>(float (&x)[16], float (&y)[16], float& pd, int li) {" % ("true" if chol else "false"),
                "    float P, Pc, t, rp, f, na, nb;",
                "    unsigned long long sn, ss, sg, se;",
                "    asm volatile(",
                body,
                "        : " + xs + ",",
                "          " + ys + ",",
                '          [pd] "+v"(pd), [P] "=&v"(P), [Pc] "=&v"(Pc), [t] "=&v"(t), [rp] "=&v"(rp), [f] "=&v"(f),',
                '          [na] "=&v"(na), [nb] "=&v"(nb), [sn] "=&s"(sn), [ss] "=&s"(ss), [sg] "=&s"(sg), [se] "=&s"(se)',
                '        : [li] "v"(li), [T] "v"(%s), [nz] "v"(-0.0f));' % ("1e-20f" if chol else "1e-9f"),
                "}", ""]
    with open(os.path.join(HERE, "diag16_asm.h"), "w") as fh:
        fh.write("\n".join(out))
    return out


if __name__ == "__main__":
    emit()
