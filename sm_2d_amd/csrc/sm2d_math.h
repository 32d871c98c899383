// sm2d_math.h — the project's own elementary functions for the transport step.
//
// Every transcendental the transport physics needs (exp, log, sin, cos, erf) is
// evaluated here in IEEE binary64 from textbook expansions and then rounded once
// to binary32.  The same source is compiled by g++ for the CPU oracle and by
// hipcc for gfx950, and it uses only operations that IEEE 754 defines exactly
// (+, -, *, /, fma, rint, frexp, ldexp, conversions).  With floating-point
// contraction disabled on both compilers (-ffp-contract=off) the CPU and the GPU
// therefore return bit-identical results, which is what makes bit-exact parity
// between the HIP kernels and the oracle possible.
//
// Methods (none is taken from a vendor library):
//   exp  : Cody–Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor series of
//          e^r to degree 14 (truncation < 2e-18 relative), scaling by ldexp.
//   log  : x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
//          s = (m-1)/(m+1), |s| <= 0.1716, odd series to s^25.
//   sin/cos : reduction by pi/2 with a three-part Cody–Waite constant, Taylor
//          series on |r| <= pi/4 (degree 19 / 20).
//   erf  : erf x = 2/sqrt(pi) e^{-x^2} sum_n 2^n x^{2n+1} / (2n+1)!!  (all terms
//          positive, so no cancellation), |x| >= 6 returns +-1.
// The double results are accurate to a few units in the last double place, so
// the rounded float equals the correctly rounded float except for inputs whose
// exact result lies within ~1e-16 relative of a float rounding boundary.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SM2D_HD __host__ __device__
#else
#define SM2D_HD
#endif

namespace sm2d {

SM2D_HD inline double exp_d(double x) {
    if (x < -745.0) return 0.0;
    if (x > 709.0) return HUGE_VAL;
    const double inv_ln2 = 1.44269504088896338700e+00;
    // ln2 split so that k * ln2_hi is exact for |k| < 2^20.
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double k = __builtin_rint(x * inv_ln2);
    double r = (x - k * ln2_hi) - k * ln2_lo;
    // Taylor coefficients 1/n!, n = 14 .. 0 (Horner, fused).
    double p = 1.0 / 87178291200.0;
    p = __builtin_fma(p, r, 1.0 / 6227020800.0);
    p = __builtin_fma(p, r, 1.0 / 479001600.0);
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}

SM2D_HD inline double log_d(double x) {
    if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
    if (x == HUGE_VAL) return x;
    int e = 0;
    double m = __builtin_frexp(x, &e);   // m in [0.5, 1)
    const double sqrt_half = 7.07106781186547524401e-01;
    if (m < sqrt_half) {
        m = m * 2.0;
        e = e - 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    // 2 * (1 + s2/3 + s2^2/5 + ... + s2^12/25)
    double p = 2.0 / 25.0;
    p = __builtin_fma(p, s2, 2.0 / 23.0);
    p = __builtin_fma(p, s2, 2.0 / 21.0);
    p = __builtin_fma(p, s2, 2.0 / 19.0);
    p = __builtin_fma(p, s2, 2.0 / 17.0);
    p = __builtin_fma(p, s2, 2.0 / 15.0);
    p = __builtin_fma(p, s2, 2.0 / 13.0);
    p = __builtin_fma(p, s2, 2.0 / 11.0);
    p = __builtin_fma(p, s2, 2.0 / 9.0);
    p = __builtin_fma(p, s2, 2.0 / 7.0);
    p = __builtin_fma(p, s2, 2.0 / 5.0);
    p = __builtin_fma(p, s2, 2.0 / 3.0);
    p = __builtin_fma(p, s2, 2.0);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double de = (double)e;
    return de * ln2_hi + (de * ln2_lo + s * p);
}

// sin and cos of a reduced argument |r| <= pi/4.
SM2D_HD inline double sin_poly(double r) {
    double r2 = r * r;
    double p = -1.0 / 121645100408832000.0;            // -1/19!
    p = __builtin_fma(p, r2, 1.0 / 355687428096000.0);  //  1/17!
    p = __builtin_fma(p, r2, -1.0 / 1307674368000.0);   // -1/15!
    p = __builtin_fma(p, r2, 1.0 / 6227020800.0);       //  1/13!
    p = __builtin_fma(p, r2, -1.0 / 39916800.0);        // -1/11!
    p = __builtin_fma(p, r2, 1.0 / 362880.0);           //  1/9!
    p = __builtin_fma(p, r2, -1.0 / 5040.0);            // -1/7!
    p = __builtin_fma(p, r2, 1.0 / 120.0);              //  1/5!
    p = __builtin_fma(p, r2, -1.0 / 6.0);               // -1/3!
    return __builtin_fma(r * r2, p, r);
}

SM2D_HD inline double cos_poly(double r) {
    double r2 = r * r;
    double p = 1.0 / 2432902008176640000.0;             //  1/20!
    p = __builtin_fma(p, r2, -1.0 / 6402373705728000.0);  // -1/18!
    p = __builtin_fma(p, r2, 1.0 / 20922789888000.0);     //  1/16!
    p = __builtin_fma(p, r2, -1.0 / 87178291200.0);       // -1/14!
    p = __builtin_fma(p, r2, 1.0 / 479001600.0);          //  1/12!
    p = __builtin_fma(p, r2, -1.0 / 3628800.0);           // -1/10!
    p = __builtin_fma(p, r2, 1.0 / 40320.0);              //  1/8!
    p = __builtin_fma(p, r2, -1.0 / 720.0);               // -1/6!
    p = __builtin_fma(p, r2, 1.0 / 24.0);                 //  1/4!
    p = __builtin_fma(p, r2, -0.5);                       // -1/2!
    return __builtin_fma(r2, p, 1.0);
}

// Reduces x by multiples of pi/2; valid (to double accuracy) for |x| < 2^19.
SM2D_HD inline double reduce_pio2(double x, int& quadrant) {
    const double two_over_pi = 6.36619772367581382433e-01;
    // pi/2 = pio2_1 + pio2_2 + pio2_3, pio2_1 and pio2_2 with 33 significant bits.
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_2 = 6.07710050630396597660e-11;
    const double pio2_3 = 2.02226624879595063154e-21;
    double q = __builtin_rint(x * two_over_pi);
    double r = ((x - q * pio2_1) - q * pio2_2) - q * pio2_3;
    quadrant = ((int)q) & 3;
    return r;
}

// sin and cos together: one reduction, both polynomials, the quadrant picks signs and roles
// (the values are those of sin_d and cos_d; no branch on the quadrant).
SM2D_HD inline void sincos_d(double x, double& sn, double& cs) {
    int q = 0;
    const double r = reduce_pio2(x, q);
    const double sp = sin_poly(r), cp = cos_poly(r);
    sn = q == 0 ? sp : q == 1 ? cp : q == 2 ? -sp : -cp;
    cs = q == 0 ? cp : q == 1 ? -sp : q == 2 ? -cp : sp;
}

SM2D_HD inline double sin_d(double x) {
    int q = 0;
    double r = reduce_pio2(x, q);
    switch (q) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}

SM2D_HD inline double cos_d(double x) {
    int q = 0;
    double r = reduce_pio2(x, q);
    switch (q) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}

// erf in FP64 with a fixed cost: |x| <= 1 Maclaurin series (18 terms); 1 < |x| < 6
// erf = 1 - exp(-x^2) erfcx(|x|) with erfcx as degree-16 Chebyshev series on [1,2], [2,3.5],
// [3.5,6] (Clenshaw recurrence); |x| >= 6 gives +-1 (erfc(6) < 2.2e-17).  Coefficients from
// gen_erf_coeffs.py.
// Maclaurin coefficients of erf(x)/x in x^2
constexpr double ERF_SMALL[18] = {1.1283791670955126, -0.37612638903183754, 0.11283791670955126, -0.026866170645131252, 0.005223977625442188, -0.0008548327023450852, 0.00012055332981789664, -1.492565035840625e-05, 1.6462114365889246e-06, -1.6365844691234924e-07, 1.4807192815879218e-08, -1.2290555301717926e-09, 9.422759064650411e-11, -6.7113668551641105e-12, 4.4632242632864775e-13, -2.7835162072109212e-14, 1.6342614095367152e-15, -9.063970842808673e-17};
// erfcx on [1.0, 2.0], relative error 1.11e-15
constexpr double ERFCX_0[17] = {0.33142728163945206, -0.08500213770674318, 0.009951127341466127, -0.0010810888102208297, 0.0001102324363119606, -1.0636830518200183e-05, 9.775012680520952e-07, -8.597922867975393e-08, 7.267708404039144e-09, -5.92354376740777e-10, 4.668392193897692e-11, -3.5661104057662214e-12, 2.645897641899542e-13, -1.9121009342342266e-14, 1.3381078217890974e-15, -9.434725441780485e-17, -4.368034708275173e-18};
// erfcx on [2.0, 3.5], relative error 6.66e-16
constexpr double ERFCX_1[17] = {0.19943078169681194, -0.049380025359927374, 0.00584049638189269, -0.0006632226933018013, 7.259178935511782e-05, -7.68295276597933e-06, 7.883915794572118e-07, -7.861599623216689e-08, 7.632681891791811e-09, -7.22722091081603e-10, 6.683956871903844e-11, -6.045468160968323e-12, 5.353825538120602e-13, -4.64767559382121e-14, 3.949204052985312e-15, -3.3675011888776724e-16, 2.081224107920472e-17};
// erfcx on [3.5, 6.0], relative error 8.88e-16
constexpr double ERFCX_2[17] = {0.12011119518612724, -0.030775083678439782, 0.003865397162805734, -0.00047649553104630075, 5.7703715281466306e-05, -6.8706380375185965e-06, 8.049536979224912e-07, -9.285957122895145e-08, 1.0554537585057767e-08, -1.1826607953411862e-09, 1.3071368727667955e-10, -1.4257236555423402e-11, 1.5353302013861495e-12, -1.633113461668731e-13, 1.715853345016225e-14, -1.7882945488732093e-15, 1.8428163989931123e-16};

SM2D_HD inline double cheb16(const double* c, double t) {
    const double t2 = 2.0 * t;
    double b1 = 0.0, b2 = 0.0;
    for (int k = 16; k >= 1; --k) {
        const double b0 = __builtin_fma(t2, b1, c[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    return __builtin_fma(t, b1, c[0] - b2);
}

SM2D_HD inline double erf_d(double x) {
    const double ax = x < 0.0 ? -x : x;
    if (ax >= 6.0) return x < 0.0 ? -1.0 : 1.0;
    if (ax == 0.0) return x;
    const double x2 = ax * ax;
    double r;
    if (ax <= 1.0) {
        double p = ERF_SMALL[17];
        for (int n = 16; n >= 0; --n) p = __builtin_fma(p, x2, ERF_SMALL[n]);
        r = ax * p;
    } else {
        double cx;
        if (ax <= 2.0) cx = cheb16(ERFCX_0, 2.0 * ax - 3.0);
        else if (ax <= 3.5) cx = cheb16(ERFCX_1, (2.0 * ax - 5.5) / 1.5);
        else cx = cheb16(ERFCX_2, (2.0 * ax - 9.5) / 2.5);
        r = 1.0 - exp_d(-x2) * cx;
    }
    if (r > 1.0) r = 1.0;
    return x < 0.0 ? -r : r;
}

SM2D_HD inline float expf_(float x) { return (float)exp_d((double)x); }
SM2D_HD inline float logf_(float x) { return (float)log_d((double)x); }
SM2D_HD inline float sinf_(float x) { return (float)sin_d((double)x); }
SM2D_HD inline float cosf_(float x) { return (float)cos_d((double)x); }
SM2D_HD inline void sincosf_(float x, float& sn, float& cs) {
    double s, c;
    sincos_d((double)x, s, c);
    sn = (float)s;
    cs = (float)c;
}
SM2D_HD inline float erff_(float x) { return (float)erf_d((double)x); }
SM2D_HD inline float sqrtf_(float x) { return (float)__builtin_sqrt((double)x); }

}  // namespace sm2d
