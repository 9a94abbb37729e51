// gs_math.h - deterministic scalar maths shared by host and device code.
//
// GLSL leaves log/exp/sin/cos implementation-defined.  This pipeline
// evaluates them with one fixed double-precision algorithm built only from
// IEEE double operations (no FMA: the library is compiled with
// -ffp-contract=off), then rounds to float.  The same algorithm gives the
// same bits on the host and on gfx950, so the projected splats and the
// compositor's discard decisions are reproducible bit for bit on the CPU.
// The GPS point sampler (a log and a sin/cos per point) and the
// StochasticSplats fragment cutoff (a log per staged splat) use the binary32
// fused-multiply-add forms further down instead (equally reproducible).
//
//   log: x = 2^e m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//        s = (m-1)/(m+1), series to s^21.
//   exp: Cody-Waite reduction by ln 2 (two-part constant), Taylor to r^13.
//   sin/cos: reduction by pi/2 with a three-part constant, Taylor series on
//        |r| <= pi/4 to r^17 / r^18.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GS_HD __host__ __device__ __forceinline__

namespace gs {

GS_HD uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
GS_HD double bitsd(uint64_t b) { return __builtin_bit_cast(double, b); }
GS_HD uint32_t fbits(float x) { return __builtin_bit_cast(uint32_t, x); }
GS_HD float bitsf(uint32_t b) { return __builtin_bit_cast(float, b); }

constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kTwoOverPi = 6.36619772367581382433e-01;
constexpr double kPio2_1 = 1.57079632673412561417e+00;
constexpr double kPio2_2 = 6.07710050630396597660e-11;
constexpr double kPio2_3 = 2.02226624871116645580e-21;

GS_HD double log_d(double x) {
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {
        x *= 18014398509481984.0;
        b = dbits(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = bitsd((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    return 2.0 * s * p + (double)e * kLn2;
}

GS_HD double exp_d(double x) {
    if (x != x) return x;
    if (x > 709.0) return __builtin_inf();
    if (x < -745.0) return 0.0;
    double k = __builtin_floor(x * kInvLn2 + 0.5);
    double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int ki = (int)k;
    int k1 = ki / 2, k2 = ki - k1;
    p *= bitsd((uint64_t)(k1 + 1023) << 52);
    p *= bitsd((uint64_t)(k2 + 1023) << 52);
    return p;
}

GS_HD double sin_poly(double r) {
    double z = r * r;
    double p = -1.0 / 355687428096000.0;
    p = p * z + 1.0 / 1307674368000.0;
    p = p * z - 1.0 / 6227020800.0;
    p = p * z + 1.0 / 39916800.0;
    p = p * z - 1.0 / 362880.0;
    p = p * z + 1.0 / 5040.0;
    p = p * z - 1.0 / 120.0;
    p = p * z + 1.0 / 6.0;
    return r - r * z * p;
}

GS_HD double cos_poly(double r) {
    double z = r * r;
    double p = 1.0 / 6402373705728000.0;
    p = p * z - 1.0 / 20922789888000.0;
    p = p * z + 1.0 / 87178291200.0;
    p = p * z - 1.0 / 479001600.0;
    p = p * z + 1.0 / 3628800.0;
    p = p * z - 1.0 / 40320.0;
    p = p * z + 1.0 / 720.0;
    p = p * z - 1.0 / 24.0;
    p = p * z + 0.5;
    return 1.0 - z * p;
}

GS_HD double reduce_pio2(double x, int* q) {
    double k = __builtin_floor(x * kTwoOverPi + 0.5);
    double r = ((x - k * kPio2_1) - k * kPio2_2) - k * kPio2_3;
    long long ki = (long long)k;
    *q = (int)(ki & 3);
    return r;
}

GS_HD double sin_d(double x) {
    if (x != x || x == __builtin_inf() || x == -__builtin_inf()) return __builtin_nan("");
    int q;
    double r = reduce_pio2(x, &q);
    double s = sin_poly(r), c = cos_poly(r);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}

GS_HD double cos_d(double x) {
    if (x != x || x == __builtin_inf() || x == -__builtin_inf()) return __builtin_nan("");
    int q;
    double r = reduce_pio2(x, &q);
    double s = sin_poly(r), c = cos_poly(r);
    return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}

// sin and cos of one argument with a shared reduction.
GS_HD void sincos_f(float x, float* sn, float* cs) {
    double xd = (double)x;
    int q;
    double r = reduce_pio2(xd, &q);
    double s = sin_poly(r), c = cos_poly(r);
    double sv = q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
    double cv = q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
    *sn = (float)sv;
    *cs = (float)cv;
}

GS_HD float logf_(float x) { return (float)log_d((double)x); }
GS_HD float expf_(float x) { return (float)exp_d((double)x); }
GS_HD float sinf_(float x) { return (float)sin_d((double)x); }
GS_HD float cosf_(float x) { return (float)cos_d((double)x); }

// Binary32 log, sin and cos with explicit fused multiply-adds (one rounding
// each, the same on the host and on gfx950), a third of the instructions of
// the double-precision forms: the GPS point sampler's radius and angle, and
// the StochasticSplats fragment cutoff log(1 / (255 o)).
// Coefficients: minimax fits by Lawson's iteratively reweighted least
// squares on Chebyshev nodes, each coefficient rounded to float in turn and
// the rest refitted (oracle/fit_gps_math.py).
//   logf_fma: positive normal x; x = 2^k m, m in [sqrt(1/2), sqrt(2)),
//     log m = f + f^2 Q(f), f = m - 1, Q of degree 8 with Q(0) = -1/2.
//     Within 1 ulp of the correctly rounded log for every normal float up
//     to 1e30 (checked exhaustively).
//   gps_sincosf: |x| < 1e4; reduction by pi/2 with a three-part constant,
//     sin r = r + r^3 S(r^2), cos r = 1 - r^2/2 + r^4 C(r^2), S and C of
//     degree 2 on |r| <= 0.801.  For every float in [0, 2 pi): within 1 ulp
//     (sin), 2 ulp (cos) and 9.3e-8 absolute.
constexpr float kLogFmaQ[9] = {-0.5f, 0.3333333134651184f, -0.2500081956386566f, 0.2000124156475067f, -0.166234090924263f, 0.14201579988002777f, -0.13159610331058502f, 0.12762217223644257f, -0.07636439055204391f};
constexpr float kGpsSinS[3] = {-0.16666653752326965f, 0.008332093246281147f, -0.00019505145610310137f};
constexpr float kGpsCosC[3] = {0.04166664183139801f, -0.0013887130189687014f, 2.4412080165348016e-05f};

GS_HD float logf_fma(float x) {
    uint32_t ix = fbits(x) + (0x3f800000u - 0x3f3504f3u);
    const int k = (int)(ix >> 23) - 127;
    ix = (ix & 0x007fffffu) + 0x3f3504f3u;
    const float f = bitsf(ix) - 1.0f;
    float q = kLogFmaQ[8];
    for (int i = 7; i >= 0; --i) q = __builtin_fmaf(q, f, kLogFmaQ[i]);
    const float r = __builtin_fmaf(f * f, q, f);
    const float fk = (float)k;
    return __builtin_fmaf(fk, 0.693147182f, __builtin_fmaf(fk, -1.90465421e-09f, r));
}

GS_HD void gps_sincosf(float x, float* sn, float* cs) {
    const float q = __builtin_rintf(x * 0.636619747f);
    const int iq = (int)q;
    float r = __builtin_fmaf(-q, 1.57079637f, x);
    r = __builtin_fmaf(-q, -4.37113883e-08f, r);
    r = __builtin_fmaf(-q, -1.71512451e-15f, r);
    const float z = r * r;
    const float ps = __builtin_fmaf(__builtin_fmaf(kGpsSinS[2], z, kGpsSinS[1]), z, kGpsSinS[0]);
    const float pc = __builtin_fmaf(__builtin_fmaf(kGpsCosC[2], z, kGpsCosC[1]), z, kGpsCosC[0]);
    const float s = __builtin_fmaf(r * z, ps, r);
    const float c = __builtin_fmaf(z * z, pc, __builtin_fmaf(-0.5f, z, 1.0f));
    const int k = iq & 3;
    *sn = k == 0 ? s : k == 1 ? c : k == 2 ? -s : -c;
    *cs = k == 0 ? c : k == 1 ? -s : k == 2 ? -c : s;
}

GS_HD float minf(float a, float b) { return a < b ? a : b; }
GS_HD float maxf(float a, float b) { return a > b ? a : b; }
GS_HD float clampf(float x, float lo, float hi) { return minf(maxf(x, lo), hi); }
GS_HD float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
GS_HD float fractf(float x) { return x - __builtin_floorf(x); }
GS_HD float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// GLSL uint(float) for the non-negative values of this pipeline, saturating
// at 2^32 the way the hardware conversion does.
GS_HD uint32_t f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

// Li2 fit used by GPS (shaders/gps.glsl:102-109, shaders/cull.glsl:116-123).
GS_HD float dilog(float x) {
    float s = 1.0f - x;
    float y = (((((((-1.068681974f * x + 3.334685126f) * x - 4.173996483f) * x +
                   2.567860600f) * x - 0.884150470f) * x - 0.123550674f) * x +
                   1.992765336f) * x + 6.74195669e-05f);
    y += (s > 0.0f) ? (s * logf_(s)) : 0.0f;
    return y;
}

// Inverse fit (shaders/gps.glsl:225-230).
GS_HD float inv_dilog(float x) {
    return ((((((((((-0.322192871f * x + 2.562830719f) * x - 8.669815892f) * x +
                    16.243561492f) * x - 18.394710417f) * x + 12.891925084f) * x -
                    5.501796628f) * x + 1.361194673f) * x - 0.417517201f) * x +
                    1.008071981f) * x - 6.39879974e-05f);
}

// Integer hash of GPS work-count rounding (shaders/gps.glsl:111-118).
GS_HD uint32_t hash_bits(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

GS_HD float hash01(uint32_t x) { return (float)hash_bits(x) * (1.0f / 4294967296.0f); }

// Point-seed mix of gps_splat_cs (shaders/gps.glsl:262-268).
GS_HD void pcg2d(uint32_t& x, uint32_t& y) {
    x = 1664525u * x + 1013904223u;
    y = 1664525u * y + 1013904223u;
    x += 1664525u * y;
    y += 1664525u * x;
    x ^= x >> 16;
    y ^= y >> 16;
    x += 1664525u * y;
    y += 1664525u * x;
    x ^= x >> 16;
    y ^= y >> 16;
}

}  // namespace gs
