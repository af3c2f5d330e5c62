"""Checks on gfx950 assembly built here (hipcc -S output and disassembly of our
own code objects): where the round-toward-+infinity regions are and which FP64
instructions round inside and outside them."""
from __future__ import annotations

import re

# FP64 instructions whose result depends on the rounding mode.
ROUNDING = re.compile(r"^\s*(v_(add|mul|fma|sub|subrev|div_fmas|div_scale|rcp|sqrt|ldexp|cvt)[a-z0-9_]*f64)\b")
ENTER = "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 1"
LEAVE = "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0"
PACKED_F32 = re.compile(r"^\s*v_pk_(add|mul|fma)_f32\b")


def regions(lines):
    """Bodies of the round-up regions, in text order."""
    inside = False
    cur = []
    for ln in lines:
        if ENTER in ln:
            inside = True
            cur = []
            continue
        if LEAVE in ln and inside:
            inside = False
            yield cur
            continue
        if inside:
            cur.append(ln)


_LABEL = re.compile(r"^([A-Za-z_.$][\w.$]*):")
_TYPE_FN = re.compile(r"^\s*\.type\s+([\w.$]+),@function")
_BRANCH = re.compile(r"^\s*(s_branch|s_cbranch_\w+)\s+([\w.$]+)")


def _functions(lines):
    """(name, body lines) of every function in hipcc -S output."""
    names = {m.group(1) for ln in lines for m in [_TYPE_FN.match(ln)] if m}
    cur = None
    body = []
    for ln in lines:
        m = _LABEL.match(ln)
        if m and m.group(1) in names:
            if cur is not None:
                yield cur, body
            cur, body = m.group(1), []
            continue
        if cur is not None:
            if ln.startswith(".Lfunc_end"):
                yield cur, body
                cur, body = None, []
                continue
            body.append(ln)
    if cur is not None:
        yield cur, body


def _blocks(body):
    """Basic blocks: (label, [(line, in_asm)], successors, falls_through)."""
    blocks = []
    label, insns = "<entry>", []
    in_asm = False

    def close(fall):
        blocks.append([label, insns, [], fall])

    for ln in body:
        m = _LABEL.match(ln)
        if m:
            close(True)
            label, insns = m.group(1), []
            continue
        st = ln.strip()
        if st.startswith("; %bb."):
            close(True)
            label, insns = st, []
            continue
        if ";;#ASMSTART" in ln:
            in_asm = True
            continue
        if ";;#ASMEND" in ln:
            in_asm = False
            continue
        if not st or st.startswith(";") or st.startswith("."):
            continue
        insns.append((st, in_asm))
        b = _BRANCH.match(st)
        if b:
            blocks.append([label, insns, [b.group(2)], b.group(1) != "s_branch"])
            label, insns = f"<after {len(blocks)}>", []
        elif st.startswith("s_endpgm") or st.startswith("s_setpc_b64"):
            blocks.append([label, insns, [], False])
            label, insns = f"<after {len(blocks)}>", []
    close(False)
    return blocks


def check_asm_regions(lines) -> int:
    """hipcc -S output: follows the MODE register's FP64 rounding field through
    each function's control flow (every function starts in round-to-nearest;
    the region's entry and exit are the only writes). Every rounding FP64
    instruction reached in round-up mode comes from inline asm, every rounding
    inline asm instruction is reached only in round-up mode, and no block is
    reached in both modes at a rounding instruction or holds a call or another
    mode write in round-up mode. Returns the number of region entries."""
    found = sum(1 for ln in lines if ENTER in ln)
    for name, body in _functions(lines):
        blocks = _blocks(body)
        index = {b[0]: i for i, b in enumerate(blocks)}
        succ = []
        for i, (label, insns, targets, fall) in enumerate(blocks):
            s = [index[t] for t in targets if t in index]
            if fall and i + 1 < len(blocks):
                s.append(i + 1)
            succ.append(s)
        state_in = [set() for _ in blocks]
        state_in[0] = {"rn"}
        work = [0]
        while work:
            i = work.pop()
            state = set(state_in[i])
            for st, in_asm in blocks[i][1]:
                if ENTER in st:
                    state = {"ru"}
                    continue
                if LEAVE in st:
                    state = {"rn"}
                    continue
                if ROUNDING.match(st):
                    if in_asm:
                        assert state == {"ru"}, f"{name}: interval asm outside every round-up region: {st}"
                    else:
                        assert "ru" not in state, f"{name}: compiler-generated rounding op inside the round-up region: {st}"
                if "ru" in state:
                    assert not (st.startswith("s_swappc") or st.startswith("s_setpc")), f"{name}: call in the region"
                    assert not (st.startswith("s_setreg") and not in_asm), f"{name}: stray mode write in the region"
            for j in succ[i]:
                if not state <= state_in[j]:
                    state_in[j] |= state
                    work.append(j)
    assert not any(PACKED_F32.match(ln) for ln in lines), "packed FP32 arithmetic in the code"
    return found
