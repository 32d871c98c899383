"""Generates the binary32 erf coefficients of sm2d_math.h (run once; prints C++).

|x| <= 1      : Maclaurin series erf(x) = 2/sqrt(pi) * sum_n (-1)^n x^(2n+1) / (n! (2n+1)),
                12 terms (truncation < 1e-10), coefficients rounded to binary32.
1 < |x| < 3.95: erf(x) = 1 - exp(-x^2) * erfcx(x); erfcx = exp(x^2) erfc(x) on [1, 2] and
                [2, 3.95] as Chebyshev series (least squares on 4000 Chebyshev nodes of
                scipy.special.erfcx, coefficients rounded to binary32), evaluated with Clenshaw's
                recurrence.
|x| >= 3.95   : erfc(3.95) < 3e-8, half a unit in the last place of binary32 below 1, so erf
                rounds to +-1.
"""
import math

import numpy as np
from numpy.polynomial import chebyshev as Ch
from scipy.special import erfcx

INTERVALS = [(1.0, 2.0, 10), (2.0, 3.95, 10)]


def main():
    c = [2.0 / math.sqrt(math.pi) * (-1) ** n / (math.factorial(n) * (2 * n + 1)) for n in range(12)]
    print("// Maclaurin coefficients of erf(x)/x in x^2")
    print("constexpr float ERF_SMALL[12] = {" + ", ".join(repr(float(np.float32(v))) + "f" for v in c) + "};")
    for i, (a, b, deg) in enumerate(INTERVALS):
        k = np.arange(4000)
        t = np.cos(np.pi * (k + 0.5) / 4000)
        x = (t * (b - a) + (a + b)) / 2
        co = Ch.chebfit(t, erfcx(x), deg).astype(np.float32)
        xt = np.linspace(a, b, 100001)
        err = np.max(np.abs(Ch.chebval((2 * xt - (a + b)) / (b - a), co.astype(np.float64)) / erfcx(xt) - 1))
        print(f"// erfcx on [{a}, {b}], degree {deg}, relative error of the rounded series {err:.2e}")
        print(f"constexpr float ERFCX_{i}[{deg + 1}] = {{" + ", ".join(repr(float(v)) + "f" for v in co) + "};")


if __name__ == "__main__":
    main()
