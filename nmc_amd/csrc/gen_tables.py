"""Writes nmc_tables.h: the tables of the transport loops' natural logarithm
and azimuth sine and cosine, included by the GPU code and the CPU oracle.

NMC_LOG_TABLE: 128 rows of nmc::log_tab (nmco_logf_tab in the oracle).

Entry i covers the reduced mantissa m' of x = 2^k m' whose top seven
mantissa bits are i: m' in [1 + i/128, 1 + (i+1)/128) for i < 64 and, halved
so that m' stays in [3/4, 3/2), m' in [(1 + i/128)/2, (1 + (i+1)/128)/2) for
i >= 64.  Fields (rc, L_hi, L_lo, 0):
  rc    1/c for the interval's midpoint c, rounded to 8 significant bits, so
        that m' rc - 1 is exact in float (|m' rc - 1| < 2^-7);
  L     -log(rc) in double, split into L_hi, a multiple of 2^-15 (so that
        k ln2_hi + L_hi is exact for the loop's k), and L_lo = float(L - L_hi).
The six intervals next to m' = 1 (i = 0, 1 and 124..127) take rc = 1, L = 0:
there log(m') = log1p(m' - 1) with m' - 1 exact, which keeps full relative
precision as log(x) goes to 0.

NMC_SINCOS_TABLE: 64 rows {sin(2 pi j/64), cos(2 pi j/64)} of
nmc::sincos2pi_bits (nmco_sincos2pi_bits), each the double value rounded to
float, with the quarter turns j = 0, 16, 32, 48 exact (0 and +-1).

Run: python3 nmc_amd/csrc/gen_tables.py > nmc_amd/csrc/nmc_tables.h
"""
from __future__ import annotations

import math
import struct


def unit_entry(i: int) -> bool:
    return i < 2 or i >= 124


def entry(i: int) -> tuple[float, float, float]:
    if unit_entry(i):
        return 1.0, 0.0, 0.0
    if i < 64:
        lo, hi = 1.0 + i / 128.0, 1.0 + (i + 1) / 128.0
    else:
        lo, hi = (1.0 + i / 128.0) / 2.0, (1.0 + (i + 1) / 128.0) / 2.0
    c = (lo + hi) / 2.0
    m, e = math.frexp(1.0 / c)                 # 1/c = m 2^e, m in [0.5, 1)
    rc = math.ldexp(round(math.ldexp(m, 8)), e - 8)   # round half to even
    big_l = -math.log(rc)
    l_hi = round(big_l * 32768.0) / 32768.0
    l_lo = struct.unpack("<f", struct.pack("<f", big_l - l_hi))[0]
    return rc, l_hi, l_lo


def table() -> list[tuple[float, float, float]]:
    return [entry(i) for i in range(128)]


def f32(x: float) -> float:
    return struct.unpack("<f", struct.pack("<f", x))[0]


def sincos_table() -> list[tuple[float, float]]:
    exact = {0: (0.0, 1.0), 16: (1.0, 0.0), 32: (0.0, -1.0), 48: (-1.0, 0.0)}
    return [exact.get(j, (f32(math.sin(2 * math.pi * j / 64)), f32(math.cos(2 * math.pi * j / 64))))
            for j in range(64)]


def main() -> None:
    print("// Generated by gen_tables.py (the tables' definitions are there); do")
    print("// not edit.  NMC_LOG_TABLE: 128 rows {rc, L_hi, L_lo, 0} of the transport")
    print("// loops' natural logarithm; NMC_SINCOS_TABLE: 64 rows {sin, cos} of")
    print("// 2 pi j/64 for their azimuth.  Included by the GPU code and the CPU oracle.")
    print("#ifndef NMC_TABLES_H")
    print("#define NMC_TABLES_H")
    print("#define NMC_LOG_TABLE \\")
    print("  { \\")
    for i, (rc, hi, lo) in enumerate(table()):
        sep = "," if i < 127 else ""
        print(f"    {{{rc.hex()}f, {hi.hex()}f, {lo.hex()}f, 0.0f}}{sep} \\")
    print("  }")
    print("#define NMC_SINCOS_TABLE \\")
    print("  { \\")
    for j, (sv, cv) in enumerate(sincos_table()):
        sep = "," if j < 63 else ""
        print(f"    {{{sv.hex()}f, {cv.hex()}f}}{sep} \\")
    print("  }")
    print("#endif")


if __name__ == "__main__":
    main()
