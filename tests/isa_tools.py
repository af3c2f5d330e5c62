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


def check_asm_regions(lines) -> int:
    """hipcc -S output: inside every region each rounding FP64 instruction comes
    from inline asm, and there are no calls or stray mode writes; every inline
    asm block that rounds lies inside a region. Returns the region count."""
    found = 0
    for body in regions(lines):
        found += 1
        in_asm = False
        for ln in body:
            if ";;#ASMSTART" in ln:
                in_asm = True
                continue
            if ";;#ASMEND" in ln:
                in_asm = False
                continue
            m = ROUNDING.match(ln)
            assert not (m and not in_asm), f"compiler-generated rounding op inside the round-up region: {ln.strip()}"
            assert "s_swappc" not in ln and "s_setpc" not in ln, "call inside the round-up region"
            assert "s_setreg" not in ln or in_asm, "stray mode write inside the round-up region"
    inside = False
    in_asm = False
    for ln in lines:
        if ENTER in ln:
            inside = True
        elif LEAVE in ln:
            inside = False
        if ";;#ASMSTART" in ln:
            in_asm = True
            continue
        if ";;#ASMEND" in ln:
            in_asm = False
            continue
        if in_asm and ROUNDING.match(ln):
            assert inside, f"interval asm outside every round-up region: {ln.strip()}"
    assert not any(PACKED_F32.match(ln) for ln in lines), "packed FP32 arithmetic in the code"
    return found
