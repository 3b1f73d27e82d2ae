"""Generate sigp_pow_log_table.h: the 128-entry logarithm table of the
table-driven pow (the method of ARM's optimized-routines math/pow.c and
pow_log_data.c, which glibc >= 2.28 uses for pow).

The method writes x = 2^k z with z in [OFF, 2 OFF), OFF = 0x1.69555p-1, and
splits that range by the bits of z into 128 subintervals: subinterval i
holds the z whose bit pattern minus OFF's has i in bits 45..51.  For each
one a point c near its centre is chosen so that 1/c has few bits,

    1/c = round(N / centre) / N          if centre < 1,
    1/c = round(2 N / centre) / (2 N)    otherwise           (N = 128),

which makes r = z/c - 1 exact in one fused multiply-add, and

    logc     = round(2^43 log c) / 2^43   (so k ln2hi + logc is exact),
    logctail = log c - logc, rounded to double.

The two subintervals next to 1 get 1/c = 1 exactly, so log(x) near 1 is the
polynomial alone.  Everything is computed here from that construction with
90-digit decimal arithmetic; no table is copied.

Licence of the method: optimized-routines, Copyright (c) 2018 Arm Limited,
SPDX-License-Identifier: MIT OR Apache-2.0 WITH LLVM-exception.
"""
import struct
from decimal import Decimal, getcontext

getcontext().prec = 90
N = 128
OFF = 0x3FE6955500000000


def _asd(u: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", u))[0]


def table():
    rows = []
    for i in range(N):
        lo = Decimal(_asd(OFF + (i << 45)))
        hi = Decimal(_asd(OFF + ((i + 1) << 45)))
        centre = (lo + hi) / 2
        if centre < 1:
            invc = Decimal(round(N / centre)) / N
        else:
            invc = Decimal(round(2 * N / centre)) / (2 * N)
        logc_exact = -invc.ln()
        logc = Decimal(round(logc_exact * 2 ** 43)) / 2 ** 43
        rows.append((float(invc), float(logc), float(logc_exact - logc)))
    return rows


def main() -> None:
    out = ["// Generated by gen_pow_log_table.py -- do not edit.",
           "// log(x) table of the table-driven pow: per subinterval i = 0..127,",
           "// {1/c, logc, logctail} with logc + logctail = log(c) to ~2^-97.",
           "#pragma once",
           "#define SIGP_POW_LOG_TAB_INIT {                                        \\"]
    for a, b, c in table():
        out.append("    %s, %s, %s, \\" % (a.hex(), b.hex(), c.hex()))
    out.append("}")
    with open(__file__.replace("gen_pow_log_table.py", "sigp_pow_log_table.h"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
