"""Generates the erf coefficients of sm2d_math.h (run once; prints C++).

|x| <= 1     : Maclaurin series erf(x) = 2/sqrt(pi) * sum_n (-1)^n x^(2n+1) / (n! (2n+1)),
               18 terms (truncation < 1e-17), coefficients rounded to double.
1 < |x| < 6  : erf(x) = 1 - exp(-x^2) * erfcx(x); erfcx = exp(x^2) erfc(x) on [1,2], [2,3.5],
               [3.5,6] as degree-16 Chebyshev series (least squares on 4000 Chebyshev nodes of
               scipy.special.erfcx; relative error < 2e-14), evaluated with Clenshaw's recurrence.
"""
import math

import numpy as np
from numpy.polynomial import chebyshev as Ch
from scipy.special import erfcx

DEG = 16
INTERVALS = [(1.0, 2.0), (2.0, 3.5), (3.5, 6.0)]


def main():
    c = [2.0 / math.sqrt(math.pi) * (-1) ** n / (math.factorial(n) * (2 * n + 1)) for n in range(18)]
    print("// Maclaurin coefficients of erf(x)/x in x^2")
    print("constexpr double ERF_SMALL[18] = {" + ", ".join(repr(v) for v in c) + "};")
    for i, (a, b) in enumerate(INTERVALS):
        k = np.arange(4000)
        t = np.cos(np.pi * (k + 0.5) / 4000)
        x = (t * (b - a) + (a + b)) / 2
        co = Ch.chebfit(t, erfcx(x), DEG)
        xt = np.linspace(a, b, 100001)
        err = np.max(np.abs(Ch.chebval((2 * xt - (a + b)) / (b - a), co) / erfcx(xt) - 1))
        print(f"// erfcx on [{a}, {b}], relative error {err:.2e}")
        print(f"constexpr double ERFCX_{i}[{DEG + 1}] = {{" + ", ".join(repr(float(v)) for v in co) + "};")


if __name__ == "__main__":
    main()
