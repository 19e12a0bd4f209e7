#!/usr/bin/env python3
"""Writes diag16_asm.h: the 16 elimination steps of a 16x16 diagonal block (FP32)
as one scheduled gfx950 asm statement.

Why asm: each step's updates are v_fmac_f32_dpp (the row broadcast of lane k folded
into the FMA), which the compiler does not form by itself, and an asm statement
counts as zero wait states to the compiler's hazard recognizer, so compiler code
interleaved with such statements is padded with s_nop (an s_nop 1 costs about as
much as two VALU instructions on one wave).  Here every step is scheduled by hand
and software-pipelined: while step k's updates issue, the pivot chain of step k + 1
(broadcast, clamp, reciprocal, multiplier, selects and the hand-over between the
wave's halves) runs between them.

Layout: the two halves of the wave split the block's 16 columns by parity, so that
one instruction updates two columns.  Lane L (li = L & 15, half h = L >> 5) holds in
x[j] (j = 0..7) entry 2j + h of row li (lanes 0-15, 32-47) or of column li (lanes
16-31, 48-63) of the block, and in y[j] entry 2j + h of the matching row of L^-1 or
(U^T)^-1.  Step k's pivot column k is x[k >> 1] of half o = k & 1, the owner.  Step
k, with p = clamp(x[k >> 1] of lane k, owner half):
    owner half:  f = x[k>>1] / p (as x[k>>1] * rcp(p))
                 n = (li > k) ? -f : -0;  x[k>>1] = (li > k) ? f : x[k>>1]
                 pd = (li == k) ? p : pd
    both halves: n of the owner half (v_permlane32_swap of two copies of n: lanes
                 32-63 of the first take lanes 0-31 of the second and the other way
                 round, so the first copy ends with half 0's n in both halves and the
                 second with half 1's)
    x[j] += bcast_k(x[j]) * n  where the column 2j + h > k,
    y[j] += bcast_k(y[j]) * n  where 2j + h <= k;
a register whose two columns are on different sides of k is updated under a DPP
row_mask that leaves the other half's rows unwritten (0xc: lanes 32-63, 0x3: lanes
0-31).  The owner-only selects take their lane masks from lg0 / lg1 (li in the half
h = 0 / 1, 0 in the other, so li > k never holds there) and le0 / le1 (li, or 99).
Per element the arithmetic (FMA operands, order, the rcp) is that of one row per
lane; pd is set in the owner half of column li only, the caller hands it over.
Lanes of the column layout take x[k >> 1] = f as well: those values are dead after
step k.

Wait states kept (gfx950, one wave): a DPP source register, or an operand of
v_permlane32_swap, is not written by a VALU instruction in the two preceding
slots; an SGPR lane mask written by v_cmp is read by v_cndmask no earlier than two
slots later; a v_rcp result is consumed no earlier than one slot later.
tests/test_dpp_hazards.py checks the first rule on the compiled kernels.

Run:  python3 gen_diag16.py  (rewrites diag16_asm.h next to this file)
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HALF0, HALF1, BOTH = 0x3, 0xc, 0xf  # DPP row masks: lanes 0-31, 32-63, all


def dpp(k, mask=BOTH):
    return "row_newbcast:%d row_mask:0x%x bank_mask:0xf" % (k, mask)


def fmac(reg, nf, k, mask):
    return "v_fmac_f32_dpp %%[%s], %%[%s], %%[%s] %s" % (reg, reg, nf, dpp(k, mask))


def nregs(k):
    """The two copies of step k's multiplier and the one that holds it in both halves
    after the swap (the first for owner half 0, the second for owner half 1)."""
    p = k % 2
    a, b = "n%da" % p, "n%db" % p
    return a, b, (a if k % 2 == 0 else b)


def chain(j, chol):
    """Pivot chain of step j, as named instructions (schedule() orders them)."""
    o = j & 1
    xr = "x%d" % (j >> 1)
    na, nb, _ = nregs(j)
    c = {}
    c["G"] = "v_cmp_lt_u32_e64 %%[sg], %d, %%[lg%d]" % (j, o)
    c["E"] = "v_cmp_eq_u32_e64 %%[se], %d, %%[le%d]" % (j, o)
    c["P"] = "v_mov_b32_dpp %%[P], %%[%s] %s" % (xr, dpp(j))
    if chol:
        c["CS"] = "v_cmp_gt_f32_e64 %[ss], %[T], %[P]"
        c["PC"] = "v_cndmask_b32_e64 %[Pc], %[P], %[T], %[ss]"
    else:
        c["CN"] = "v_cmp_gt_f32_e64 %[sn], 0, %[P]"
        c["CS"] = "v_cmp_lt_f32_e64 %[ss], |%[P]|, %[T]"
        c["TT"] = "v_cndmask_b32_e64 %[t], %[T], -%[T], %[sn]"
        c["PC"] = "v_cndmask_b32_e64 %[Pc], %[P], %[t], %[ss]"
    c["R"] = "v_rcp_f32_e32 %[rp], %[Pc]"
    c["F"] = "v_mul_f32_e32 %%[f], %%[%s], %%[rp]" % xr
    c["NA"] = "v_cndmask_b32_e64 %%[%s], %%[nz], -%%[f], %%[sg]" % na
    c["NB"] = "v_cndmask_b32_e64 %%[%s], %%[nz], -%%[f], %%[sg]" % nb
    c["X"] = "v_cndmask_b32_e64 %%[%s], %%[%s], %%[f], %%[sg]" % (xr, xr)
    c["PD"] = "v_cndmask_b32_e64 %[pd], %[pd], %[Pc], %[se]"
    c["SW"] = "v_permlane32_swap_b32_e32 %%[%s], %%[%s]" % (na, nb)
    return c


def updates(k):
    """Step k's updates, the next pivot column's register first."""
    _, _, nf = nregs(k)
    fm = []
    for m in range(8):  # x: columns 2m + h > k
        if 2 * m > k:
            fm.append(fmac("x%d" % m, nf, k, BOTH))
        elif 2 * m + 1 > k:
            fm.append(fmac("x%d" % m, nf, k, HALF1))
    for m in range(8):  # y: columns 2m + h <= k
        if 2 * m + 1 <= k:
            fm.append(fmac("y%d" % m, nf, k, BOTH))
        elif 2 * m <= k:
            fm.append(fmac("y%d" % m, nf, k, HALF0))
    return fm


def schedule(fm, ch, chol):
    """Interleave step k's updates fm (fm[0] writes the next pivot column) with the
    chain ch of step k + 1.  Slots are filled from fm where a wait state is needed;
    s_nop only when fm has run out."""
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
    out.append(ch["NA"])                 # sg written long before
    out.append(ch["NB"])
    out.append(ch["X"])
    out.append(ch["PD"])                 # X, PD: two slots between NB and the swap
    out.append(ch["SW"])
    out.extend(pending)
    return out


def steps(chol):
    lines = ["s_nop 1"]                  # x may have been written by VALU just before
    lines += schedule([], chain(0, chol), chol)
    for k in range(16):
        lines += schedule(updates(k), chain(k + 1, chol) if k < 15 else None, chol)
    lines.append("s_nop 1")              # the caller may read pd or x through DPP next
    return lines


def check(lines, chol):
    """Re-verify the schedule's wait states on the generated text (independent of
    how schedule() placed them)."""
    import re
    def regs_written(l):
        m = re.match(r"(v_\w+)\s+%\[(\w+)\]", l)
        if not m:
            return set()
        if m.group(1).startswith("v_permlane"):  # a swap writes both operands
            return set(re.findall(r"%\[(\w+)\]", l))
        return {m.group(2)}
    def hazard_src(l):
        m = re.match(r"v_(?:fmac|mov)_\w+_dpp\s+%\[\w+\],\s+%\[(\w+)\]", l)
        if m:
            return {m.group(1)}
        if l.startswith("v_permlane"):
            return set(re.findall(r"%\[(\w+)\]", l))
        return set()
    def sgpr_read(l):
        return set(re.findall(r"%\[(s[ngse])\]", l.split(",", 1)[1])) if l.startswith("v_cndmask") else set()
    def slots(l):
        if l.startswith("s_nop"):
            return int(l.split()[1]) + 1
        return 1
    for i, l in enumerate(lines):
        for src in hazard_src(l):
            w = 0
            for j in range(i - 1, -1, -1):
                if w >= 2:
                    break
                if src in regs_written(lines[j]) and not lines[j].startswith("s_nop"):
                    raise AssertionError("source %s written %d slots before: %s / %s" % (src, w, lines[j], l))
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
        xs = ", ".join('[x%d] "+v"(x[%d])' % (c, c) for c in range(8))
        ys = ", ".join('[y%d] "+v"(y[%d])' % (c, c) for c in range(8))
        out += ["template <>",
                "__device__ __forceinline__ void diag16_elim_f32<%s>(float (&x)[8], float (&y)[8], float& pd, int lg0, int lg1,"
                % ("true" if chol else "false"),
                "                                                   int le0, int le1) {",
                "    float P, Pc, t, rp, f, n0a, n0b, n1a, n1b;",
                "    unsigned long long sn, ss, sg, se;",
                "    asm volatile(",
                body,
                "        : " + xs + ",",
                "          " + ys + ",",
                '          [pd] "+v"(pd), [P] "=&v"(P), [Pc] "=&v"(Pc), [t] "=&v"(t), [rp] "=&v"(rp), [f] "=&v"(f),',
                '          [n0a] "=&v"(n0a), [n0b] "=&v"(n0b), [n1a] "=&v"(n1a), [n1b] "=&v"(n1b),',
                '          [sn] "=&s"(sn), [ss] "=&s"(ss), [sg] "=&s"(sg), [se] "=&s"(se)',
                '        : [lg0] "v"(lg0), [lg1] "v"(lg1), [le0] "v"(le0), [le1] "v"(le1), [T] "v"(%s), [nz] "v"(-0.0f));'
                % ("1e-20f" if chol else "1e-9f"),
                "}", ""]
    with open(os.path.join(HERE, "diag16_asm.h"), "w") as fh:
        fh.write("\n".join(out))
    return out


if __name__ == "__main__":
    emit()
