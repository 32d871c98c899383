// sm2d_math.h — the project's own elementary functions for the transport step.
//
// Every transcendental the per-component physics needs (exp, log, sin, cos, erf) is evaluated
// here in IEEE binary32 from textbook expansions.  The same source is compiled by g++ for the
// CPU oracle and by hipcc for gfx950, and it uses only operations that IEEE 754 defines exactly
// (+, -, *, /, fma, sqrt, rint, floor, frexp, ldexp, conversions).  With floating-point
// contraction disabled on both compilers (-ffp-contract=off) the CPU and the GPU therefore
// return bit-identical results, which is what makes bit-exact parity between the HIP kernels
// and the oracle possible.  No binary64 arithmetic is involved, so the per-bin physics runs at
// the binary32 rate of the vector units.
//
// Methods (none is taken from a vendor library):
//   exp  : Cody–Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split in two binary32
//          parts (the first with 16 significant bits, so k * ln2_hi is exact); Taylor series of
//          e^r to degree 7 (truncation < 5e-9 relative) by Horner's rule with fused
//          multiply-adds; scaling by ldexp.
//   log  : x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f);
//          log(1 + f) = f - (f^2/2 - s (f^2/2 + R(s^2))) with R the odd-series remainder of
//          2 atanh(s) to s^9 (truncation < 2e-9 relative); e ln2 added in two parts.
//   sin/cos : reduction by pi/2 with a three-part Cody–Waite constant (exact products for
//          |x| < 2^15), Taylor series on |r| <= pi/4 to degree 9 (sin) and 10 (cos).
//   erf  : |x| <= 1 Maclaurin series (12 terms); 1 < |x| < 3.95 erf = 1 - e^{-x^2} erfcx(|x|)
//          with erfcx as degree-10 Chebyshev series on [1, 2] and [2, 3.95] (Clenshaw);
//          |x| >= 3.95 returns +-1 (erfc there is below half a unit in the last place).
//          Coefficients from gen_erf_coeffs.py.
// Accuracy against the correctly rounded binary32 result is stated and checked in
// tests/test_math.py (at most 2 units in the last place on the ranges the physics uses).
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

SM2D_HD inline float expf_(float x) {
    if (!(x == x)) return x;
    if (x > 88.72284f) return HUGE_VALF;
    if (x < -103.97208f) return 0.0f;
    const float inv_ln2 = 1.44269504088896341f;
    const float ln2_hi = 0.693145751953125f;       // 16 significant bits
    const float ln2_lo = 1.42860676533018720e-06f;  // ln2 - ln2_hi
    const float k = __builtin_rintf(x * inv_ln2);
    float r = __builtin_fmaf(-k, ln2_hi, x);
    r = __builtin_fmaf(-k, ln2_lo, r);
    float p = 1.0f / 5040.0f;
    p = __builtin_fmaf(p, r, 1.0f / 720.0f);
    p = __builtin_fmaf(p, r, 1.0f / 120.0f);
    p = __builtin_fmaf(p, r, 1.0f / 24.0f);
    p = __builtin_fmaf(p, r, 1.0f / 6.0f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_ldexpf(p, (int)k);
}

SM2D_HD inline float logf_(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -HUGE_VALF : (x - x) / (x - x);
    if (x == HUGE_VALF) return x;
    int e = 0;
    float m = __builtin_frexpf(x, &e);  // m in [0.5, 1), subnormals included
    const float sqrt_half = 0.70710678118654752f;
    if (m < sqrt_half) {
        m = m * 2.0f;
        e = e - 1;
    }
    const float f = m - 1.0f;  // exact (Sterbenz)
    const float s = f / (2.0f + f);
    const float z = s * s;
    // R(z) = z (2/3 + z (2/5 + z (2/7 + z 2/9))): 2 atanh(s) = 2 s + s R(z)
    float R = 2.0f / 9.0f;
    R = __builtin_fmaf(R, z, 2.0f / 7.0f);
    R = __builtin_fmaf(R, z, 2.0f / 5.0f);
    R = __builtin_fmaf(R, z, 2.0f / 3.0f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float lm = f - (hfsq - s * (hfsq + R));
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.42860676533018720e-06f;
    const float fe = (float)e;
    return __builtin_fmaf(fe, ln2_hi, __builtin_fmaf(fe, ln2_lo, lm));
}

// sin and cos of a reduced argument |r| <= pi/4.
SM2D_HD inline float sin_poly(float r) {
    const float r2 = r * r;
    float p = 1.0f / 362880.0f;
    p = __builtin_fmaf(p, r2, -1.0f / 5040.0f);
    p = __builtin_fmaf(p, r2, 1.0f / 120.0f);
    p = __builtin_fmaf(p, r2, -1.0f / 6.0f);
    return __builtin_fmaf(r * r2, p, r);
}

SM2D_HD inline float cos_poly(float r) {
    const float r2 = r * r;
    float p = -1.0f / 3628800.0f;
    p = __builtin_fmaf(p, r2, 1.0f / 40320.0f);
    p = __builtin_fmaf(p, r2, -1.0f / 720.0f);
    p = __builtin_fmaf(p, r2, 1.0f / 24.0f);
    p = __builtin_fmaf(p, r2, -0.5f);
    return __builtin_fmaf(r2, p, 1.0f);
}

// Reduces x by multiples of pi/2 (pi/2 = P1 + P2 + P3; P1 and P2 with few significant bits, so
// the products q * P1 and q * P2 are exact for |q| < 2^15).
SM2D_HD inline float reduce_pio2(float x, int& quadrant) {
    const float two_over_pi = 0.636619772367581343f;
    const float P1 = 1.5703125f;
    const float P2 = 4.837512969970703125e-4f;
    const float P3 = 7.54978995489188216e-8f;
    const float q = __builtin_rintf(x * two_over_pi);
    float r = __builtin_fmaf(-q, P1, x);
    r = __builtin_fmaf(-q, P2, r);
    r = __builtin_fmaf(-q, P3, r);
    quadrant = ((int)q) & 3;
    return r;
}

// sin and cos together: one reduction, both polynomials, the quadrant picks signs and roles
// (the values are those of sinf_ and cosf_).
SM2D_HD inline void sincosf_(float x, float& sn, float& cs) {
    int q = 0;
    const float r = reduce_pio2(x, q);
    const float sp = sin_poly(r), cp = cos_poly(r);
    sn = q == 0 ? sp : q == 1 ? cp : q == 2 ? -sp : -cp;
    cs = q == 0 ? cp : q == 1 ? -sp : q == 2 ? -cp : sp;
}

SM2D_HD inline float sinf_(float x) {
    float s, c;
    sincosf_(x, s, c);
    return s;
}

SM2D_HD inline float cosf_(float x) {
    float s, c;
    sincosf_(x, s, c);
    return c;
}

// Maclaurin coefficients of erf(x)/x in x^2
constexpr float ERF_SMALL[12] = {1.128379225730896f, -0.37612637877464294f, 0.11283791810274124f, -0.02686617150902748f, 0.005223977845162153f, -0.0008548327023163438f, 0.00012055332626914605f, -1.4925650248187594e-05f, 1.6462114444948384e-06f, -1.6365844146548625e-07f, 1.4807192805221803e-08f, -1.2290555240213052e-09f};
// erfcx on [1.0, 2.0], degree 10, relative error of the rounded series 2.60e-08
constexpr float ERFCX_0[11] = {0.33142727613449097f, -0.08500213921070099f, 0.00995112769305706f, -0.0010810887906700373f, 0.00011023243860108778f, -1.0636830666044261e-05f, 9.775012586032972e-07f, -8.597923084607828e-08f, 7.267708213021251e-09f, -5.92354387762839e-10f, 4.668393102646995e-11f};
// erfcx on [2.0, 3.95], degree 10, relative error of the rounded series 2.81e-08
constexpr float ERFCX_1[11] = {0.18853557109832764f, -0.057192061096429825f, 0.008320423774421215f, -0.0011659585870802402f, 0.00015791664191056043f, -2.073007453873288e-05f, 2.643795824042172e-06f, -3.2823416518112936e-07f, 3.97396000550998e-08f, -4.69900518496047e-09f, 5.433894445516785e-10f};

SM2D_HD inline float cheb10(const float* c, float t) {
    const float t2 = 2.0f * t;
    float b1 = 0.0f, b2 = 0.0f;
    for (int k = 10; k >= 1; --k) {
        const float b0 = __builtin_fmaf(t2, b1, c[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    return __builtin_fmaf(t, b1, c[0] - b2);
}

SM2D_HD inline float erff_(float x) {
    const float ax = x < 0.0f ? -x : x;
    if (!(ax < 3.95f)) return (x == x) ? (x < 0.0f ? -1.0f : 1.0f) : x;
    if (ax == 0.0f) return x;
    const float x2 = ax * ax;
    float r;
    if (ax <= 1.0f) {
        // x (c0 + x^2 P(x^2)) with c0 = 2/sqrt(pi) in two parts and one final rounding
        const float c0_lo = -5.8635383e-08f;  // 2/sqrt(pi) - ERF_SMALL[0]
        float p = ERF_SMALL[11];
        for (int n = 10; n >= 1; --n) p = __builtin_fmaf(p, x2, ERF_SMALL[n]);
        r = __builtin_fmaf(ax, ERF_SMALL[0], ax * __builtin_fmaf(p, x2, c0_lo));
    } else {
        const float cx = ax <= 2.0f ? cheb10(ERFCX_0, 2.0f * ax - 3.0f)
                                    : cheb10(ERFCX_1, (2.0f * ax - 5.95f) / 1.95f);
        r = 1.0f - expf_(-x2) * cx;
    }
    if (r > 1.0f) r = 1.0f;
    return x < 0.0f ? -r : r;
}

// Correctly rounded binary32 square root (the binary64 root of an exactly representable value,
// rounded once, is the correctly rounded binary32 root).
SM2D_HD inline float sqrtf_(float x) { return (float)__builtin_sqrt((double)x); }

}  // namespace sm2d
