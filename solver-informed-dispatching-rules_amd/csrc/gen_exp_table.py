"""Generate sigp_exp_table.h: the 2^(k/128) table of the table-driven exp
(ARM optimized-routines method, as used by glibc >= 2.28).

For k in [0, 128): H = RN(2^(k/128)) and T = RN((2^(k/128) - H) / H), stored
as (bits(T), bits(H) - (k << 45)) so that adding the scaled exponent k << 45
to the second word rebuilds 2^(k/128) with the integer part folded in.
Computed from first principles with 80-digit decimal arithmetic.

Licence of the method: optimized-routines, Copyright (c) 2018 Arm Limited,
SPDX-License-Identifier: MIT OR Apache-2.0 WITH LLVM-exception.
"""
import struct
from decimal import Decimal, getcontext

getcontext().prec = 80
N = 128


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def main() -> None:
    rows = []
    for k in range(N):
        e = Decimal(2) ** (Decimal(k) / Decimal(N))
        h = float(e)                      # float(Decimal) rounds to nearest
        t = float((e - Decimal(h)) / Decimal(h))
        rows.append((bits(t), (bits(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF))
    out = ["// Generated by gen_exp_table.py -- do not edit.",
           "// 2^(k/128) = H(1+T): pairs (bits(T), bits(H) - (k << 45)), k = 0..127.",
           "#pragma once",
           "#define SIGP_EXP_TAB_INIT {                         \\"]
    for t, h in rows:
        out.append("    0x%016xull, 0x%016xull,                \\" % (t, h))
    out.append("}")
    with open(__file__.replace("gen_exp_table.py", "sigp_exp_table.h"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
