"""Generate erfcx_coef.h: polynomial approximations of the scaled
complementary error function erfcx(x) = exp(x^2) erfc(x) on x in [0, XMAX],
in the variable t = (x - K) / (x + K) (the mapping of [0, inf) onto [-1, 1)
used in Chebyshev expansions of erfcx, e.g. Schonfelder, Math. Comp. 32
(1978) 1232-1240). The pair kernels then form erfc(alpha r) as
exp(-(alpha r)^2) * erfcx(alpha r), sharing the exponential that the Ewald
real-space term already needs.

Method: least-squares Chebyshev fit in t on Chebyshev nodes against
scipy.special.erfcx (double precision), converted to monomial coefficients
for Horner evaluation; the script asserts the maximum relative error of the
monomial form, evaluated in double precision, on a dense grid.

Usage: python gen_erfcx.py  (rewrites erfcx_coef.h next to this file)
"""
import os

import numpy as np
import numpy.polynomial.chebyshev as Ch
from scipy.special import erfcx

XMAX = 12.0


def fit(k, deg, dtype):
    n = 8000
    tmax = (XMAX - k) / (XMAX + k)
    tt = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    t = -1 + (tt + 1) / 2 * (tmax + 1)
    x = k * (1 + t) / (1 - t)
    c = Ch.chebfit(t, erfcx(x), deg)
    mono = Ch.cheb2poly(c)
    xt = np.linspace(0, XMAX, 400001)
    tv = ((xt - k) / (xt + k)).astype(dtype)
    acc = np.zeros_like(tv)
    for a in mono[::-1].astype(dtype):
        acc = acc * tv + a
    rel = np.abs(acc.astype(np.float64) - erfcx(xt)) / erfcx(xt)
    return mono, float(rel.max())


def main():
    k64, d64 = 4.0, 19
    m64, e64 = fit(k64, d64, np.float64)
    assert e64 < 5e-14, e64
    k32, d32 = 4.0, 9
    m32, e32 = fit(k32, d32, np.float32)
    assert e32 < 2e-6, e32
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "erfcx_coef.h")
    with open(out, "w") as f:
        f.write("// Generated by gen_erfcx.py: erfcx(x) = exp(x^2) erfc(x) ~ P((x-K)/(x+K)), x in [0, %g].\n" % XMAX)
        f.write("// Max relative error on a 400001-point grid: FP64 %.2e, FP32 %.2e.\n" % (e64, e32))
        f.write("#pragma once\n\nnamespace mds {\n\n")
        f.write("constexpr double kErfcxK64 = %r;\nconstexpr int kErfcxDeg64 = %d;\n" % (k64, d64))
        # __constant__ (not constexpr): the kernels load the coefficients with
        # scalar loads into SGPRs instead of holding 64-bit literals in VGPRs
        f.write("__constant__ double kErfcx64[%d] = {\n" % (d64 + 1))
        for a in m64:
            f.write("    %s,\n" % np.format_float_scientific(a, unique=True))
        f.write("};\n\n")
        f.write("constexpr float kErfcxK32 = %rf;\nconstexpr int kErfcxDeg32 = %d;\n" % (k32, d32))
        # constexpr: the FP32 Horner steps take them as instruction literals
        f.write("constexpr float kErfcx32[%d] = {\n" % (d32 + 1))
        for a in m32.astype(np.float32):
            f.write("    %sf,\n" % np.format_float_scientific(a, unique=True))
        f.write("};\n\n}  // namespace mds\n")
    print(out, e64, e32)


if __name__ == "__main__":
    main()
