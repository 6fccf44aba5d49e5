"""Writes twiddle_table.h: exp(-2 pi i m / 192), m = 0..191, as double
literals.  Every in-register DFT length used by the FFT kernels divides 192
(2..64 and 3 * 2^k up to 192), so their constant twiddles are entries of this
table.  Run from this directory: python gen_twiddles.py"""
import math
import os

L = 192
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "twiddle_table.h"), "w") as f:
    f.write("// Generated by gen_twiddles.py: exp(-2 pi i m / 192), m = 0..191.\n")
    f.write("#pragma once\n\nnamespace capow {\n\n")
    f.write("constexpr int TWC_LEN = %d;\n" % L)
    for name, fn in (("TWC_RE", lambda m: math.cos(2 * math.pi * m / L)),
                     ("TWC_IM", lambda m: -math.sin(2 * math.pi * m / L))):
        f.write("__device__ constexpr double %s[%d] = {\n" % (name, L))
        vals = []
        for m in range(L):
            v = fn(m)
            # exact zeros/ones where the value is exact
            if m % (L // 4) == 0:
                v = float(round(v))
            vals.append("%.17g" % v)
        for i in range(0, L, 4):
            f.write("    " + ", ".join(vals[i:i + 4]) + ",\n")
        f.write("};\n")
    f.write("\n}  // namespace capow\n")
