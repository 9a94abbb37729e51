/*
 * gs_oracle.c - CPU restatement of gamesplat's splat frame pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (see gs_oracle.h).  Built by oracle/Makefile with
 * -ffp-contract=off so every float operation below is one IEEE binary32
 * operation in source order, as the GLSL it restates.
 */
#include "gs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================
 * Deterministic transcendental functions.
 *
 * GLSL's log/exp/sin/cos are implementation-defined.  The oracle and the HIP
 * path evaluate them with the same double-precision algorithm (IEEE double
 * operations only, no FMA) and round to float, so both produce identical
 * bits.  Accuracy is pinned against libm in tests/test_oracle_math.py.
 * ====================================================================== */

static inline uint64_t dbits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }
static inline double bitsd(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }
static inline uint32_t fbits(float x) { uint32_t b; memcpy(&b, &x, 4); return b; }
static inline float bitsf(uint32_t b) { float x; memcpy(&x, &b, 4); return x; }

#define OR_LN2 0.69314718055994530942
#define OR_LN2_HI 6.93147180369123816490e-01
#define OR_LN2_LO 1.90821492927058770002e-10
#define OR_INV_LN2 1.44269504088896338700e+00
#define OR_TWO_OVER_PI 6.36619772367581382433e-01
#define OR_PIO2_1 1.57079632673412561417e+00
#define OR_PIO2_2 6.07710050630396597660e-11
#define OR_PIO2_3 2.02226624871116645580e-21

/* log(x) = e ln2 + 2 atanh(s), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)). */
double or_log_d(double x) {
    if (x != x) return x;
    if (x < 0.0) return (x - x) / 0.0;          /* NaN */
    if (x == 0.0) return -1.0 / 0.0;            /* -inf */
    if (x == 1.0 / 0.0) return x;
    uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {                               /* subnormal: scale by 2^54 */
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
    return 2.0 * s * p + (double)e * OR_LN2;
}

/* exp(x) = 2^k exp(r), r = x - k ln2 (two-part ln2), Taylor to r^13. */
double or_exp_d(double x) {
    if (x != x) return x;
    if (x > 709.0) return 1.0 / 0.0;
    if (x < -745.0) return 0.0;
    double k = floor(x * OR_INV_LN2 + 0.5);
    double r = (x - k * OR_LN2_HI) - k * OR_LN2_LO;
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
    /* scale in two steps so 2^ki never leaves the normal range */
    int k1 = ki / 2, k2 = ki - k1;
    p *= bitsd((uint64_t)(k1 + 1023) << 52);
    p *= bitsd((uint64_t)(k2 + 1023) << 52);
    return p;
}

/* sin r = r - r z Q(z), Q = 1/3! - z/5! + ... - z^7/17!, |r| <= pi/4 */
static double or_sin_poly(double r) {
    double z = r * r;
    double p = -1.0 / 355687428096000.0;              /* -1/17! */
    p = p * z + 1.0 / 1307674368000.0;                /* 1/15! */
    p = p * z - 1.0 / 6227020800.0;                   /* -1/13! */
    p = p * z + 1.0 / 39916800.0;                     /* 1/11! */
    p = p * z - 1.0 / 362880.0;                       /* -1/9! */
    p = p * z + 1.0 / 5040.0;                         /* 1/7! */
    p = p * z - 1.0 / 120.0;                          /* -1/5! */
    p = p * z + 1.0 / 6.0;                            /* 1/3! */
    return r - r * z * p;
}

static double or_cos_poly(double r) {
    double z = r * r;
    double p = 1.0 / 6402373705728000.0;              /* 1/18! */
    p = p * z - 1.0 / 20922789888000.0;               /* -1/16! */
    p = p * z + 1.0 / 87178291200.0;                  /* 1/14! */
    p = p * z - 1.0 / 479001600.0;
    p = p * z + 1.0 / 3628800.0;
    p = p * z - 1.0 / 40320.0;
    p = p * z + 1.0 / 720.0;
    p = p * z - 1.0 / 24.0;
    p = p * z + 0.5;
    return 1.0 - z * p;
}

static double or_reduce_pio2(double x, int* q) {
    double k = floor(x * OR_TWO_OVER_PI + 0.5);
    double r = ((x - k * OR_PIO2_1) - k * OR_PIO2_2) - k * OR_PIO2_3;
    long long ki = (long long)k;
    *q = (int)(ki & 3);
    return r;
}

double or_sin_d(double x) {
    if (x != x || x == 1.0 / 0.0 || x == -1.0 / 0.0) return (x - x) / (x - x);
    int q;
    double r = or_reduce_pio2(x, &q);
    switch (q) {
        case 0: return or_sin_poly(r);
        case 1: return or_cos_poly(r);
        case 2: return -or_sin_poly(r);
        default: return -or_cos_poly(r);
    }
}

double or_cos_d(double x) {
    if (x != x || x == 1.0 / 0.0 || x == -1.0 / 0.0) return (x - x) / (x - x);
    int q;
    double r = or_reduce_pio2(x, &q);
    switch (q) {
        case 0: return or_cos_poly(r);
        case 1: return -or_sin_poly(r);
        case 2: return -or_cos_poly(r);
        default: return or_sin_poly(r);
    }
}

float or_logf(float x) { return (float)or_log_d((double)x); }
float or_expf(float x) { return (float)or_exp_d((double)x); }
float or_sinf(float x) { return (float)or_sin_d((double)x); }
float or_cosf(float x) { return (float)or_cos_d((double)x); }

/* The GPS point sampler's log and sin/cos (the log, cos and sin of
 * gps_splat_cs, shaders/gps.glsl:270-276) and the StochasticSplats fragment
 * cutoff's log (shaders/splat.glsl:107-146): binary32 arithmetic with explicit
 * fmaf (one rounding each; glibc's fmaf is exact), minimax polynomial
 * coefficients from oracle/fit_gps_math.py.  log: x = 2^k m with m in
 * [sqrt(1/2), sqrt(2)), log m = f + f^2 Q(f), f = m - 1.  sin/cos: reduction
 * by pi/2 in three parts, sin r = r + r^3 S(r^2), cos r = 1 - r^2/2 +
 * r^4 C(r^2).  Accuracy against libm is pinned in tests/test_oracle_math.py. */
static const float kLogFmaQ[9] = {-0.5f, 0.3333333134651184f, -0.2500081956386566f, 0.2000124156475067f, -0.166234090924263f, 0.14201579988002777f, -0.13159610331058502f, 0.12762217223644257f, -0.07636439055204391f};
static const float kGpsSinS[3] = {-0.16666653752326965f, 0.008332093246281147f, -0.00019505145610310137f};
static const float kGpsCosC[3] = {0.04166664183139801f, -0.0013887130189687014f, 2.4412080165348016e-05f};

float or_logf_fma(float x) {
    uint32_t ix = fbits(x) + (0x3f800000u - 0x3f3504f3u);
    int k = (int)(ix >> 23) - 127;
    ix = (ix & 0x007fffffu) + 0x3f3504f3u;
    float f = bitsf(ix) - 1.0f;
    float q = kLogFmaQ[8];
    for (int i = 7; i >= 0; --i) q = fmaf(q, f, kLogFmaQ[i]);
    float r = fmaf(f * f, q, f);
    float fk = (float)k;
    return fmaf(fk, 0.693147182f, fmaf(fk, -1.90465421e-09f, r));
}

void or_gps_sincosf(float x, float* sn, float* cs) {
    float q = rintf(x * 0.636619747f);
    int iq = (int)q;
    float r = fmaf(-q, 1.57079637f, x);
    r = fmaf(-q, -4.37113883e-08f, r);
    r = fmaf(-q, -1.71512451e-15f, r);
    float z = r * r;
    float ps = fmaf(fmaf(kGpsSinS[2], z, kGpsSinS[1]), z, kGpsSinS[0]);
    float pc = fmaf(fmaf(kGpsCosC[2], z, kGpsCosC[1]), z, kGpsCosC[0]);
    float s = fmaf(r * z, ps, r);
    float c = fmaf(z * z, pc, fmaf(-0.5f, z, 1.0f));
    switch (iq & 3) {
        case 0: *sn = s; *cs = c; break;
        case 1: *sn = c; *cs = -s; break;
        case 2: *sn = -s; *cs = -c; break;
        default: *sn = -c; *cs = s; break;
    }
}

void or_logf_fma_n(const float* x, float* y, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) y[i] = or_logf_fma(x[i]);
}

void or_gps_sincosf_n(const float* x, float* sn, float* cs, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) or_gps_sincosf(x[i], sn + i, cs + i);
}

/* dilog fit, shaders/cull.glsl:116-123 and shaders/gps.glsl:102-109. */
float or_dilog(float x) {
    float s = 1.0f - x;
    float y = (((((((-1.068681974f * x + 3.334685126f) * x - 4.173996483f) * x +
                   2.567860600f) * x - 0.884150470f) * x - 0.123550674f) * x +
                   1.992765336f) * x + 6.74195669e-05f);
    y += (s > 0.0f) ? (s * or_logf(s)) : 0.0f;
    return y;
}

/* inverse fit, shaders/gps.glsl:225-230. */
float or_inv_dilog(float x) {
    return ((((((((((-0.322192871f * x + 2.562830719f) * x - 8.669815892f) * x +
                    16.243561492f) * x - 18.394710417f) * x + 12.891925084f) * x -
                    5.501796628f) * x + 1.361194673f) * x - 0.417517201f) * x +
                    1.008071981f) * x - 6.39879974e-05f);
}

/* hash01, shaders/gps.glsl:111-118. */
uint32_t or_hash01_bits(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

float or_hash01(uint32_t x) {
    return (float)or_hash01_bits(x) * (1.0f / 4294967296.0f);
}

/* point-seed mix, shaders/gps.glsl:262-268. */
void or_pcg2d(uint32_t v[2]) {
    v[0] = 1664525u * v[0] + 1013904223u;
    v[1] = 1664525u * v[1] + 1013904223u;
    v[0] += 1664525u * v[1];
    v[1] += 1664525u * v[0];
    v[0] ^= v[0] >> 16;
    v[1] ^= v[1] >> 16;
    v[0] += 1664525u * v[1];
    v[1] += 1664525u * v[0];
    v[0] ^= v[0] >> 16;
    v[1] ^= v[1] >> 16;
}

/* binary32 -> binary16, round to nearest even (RGBA16F storage). */
uint16_t or_f32_to_f16(float f) {
    uint32_t x = fbits(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
    int e = (int)exp - 127 + 15;
    if (e >= 0x1f) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        int shift = 14 - e;
        uint32_t h = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float or_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    if (exp == 0) {
        if (man == 0) return bitsf(sign);
        float v = (float)man * (1.0f / 16777216.0f);
        return sign ? -v : v;
    }
    if (exp == 0x1f) return bitsf(sign | 0x7f800000u | (man << 13));
    return bitsf(sign | ((exp - 15 + 127) << 23) | (man << 13));
}

/* ======================================================================
 * cull_cs (shaders/cull.glsl:166-424) and cull_reset_cs (:429-474)
 * ====================================================================== */

typedef struct { float m[3][3]; } mat3; /* m[column][row], as GLSL */

static mat3 mat3_mul(const mat3* a, const mat3* b) {
    mat3 o;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            o.m[c][r] = (a->m[0][r] * b->m[c][0] + a->m[1][r] * b->m[c][1]) + a->m[2][r] * b->m[c][2];
    return o;
}

static mat3 mat3_transpose(const mat3* a) {
    mat3 o;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) o.m[c][r] = a->m[r][c];
    return o;
}

static inline float or_mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static inline float or_clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float or_fract(float x) { return x - floorf(x); }
static inline float or_smoothstep(float e0, float e1, float x) {
    float t = or_clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
static inline float or_hash11(float p) { return or_fract(or_sinf(p * 127.1f) * 43758.5453123f); }

/* GLSL uint(float) for values this pipeline produces: saturate at 2^32. */
static inline uint32_t or_f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

static const float SH_C0 = 0.28209479177387814f;
static const float SH_C1 = 0.4886025119029199f;
static const float SH_C2_0 = 1.0925484305920792f;
static const float SH_C2_1 = -1.0925484305920792f;
static const float SH_C2_2 = 0.31539156525252005f;
static const float SH_C2_3 = -1.0925484305920792f;
static const float SH_C2_4 = 0.5462742152960396f;
static const float SH_C3_0 = -0.5900435899266435f;
static const float SH_C3_1 = 2.890611442640554f;
static const float SH_C3_2 = -0.4570457994644658f;
static const float SH_C3_3 = 0.3731763325901154f;
static const float SH_C3_4 = -0.4570457994644658f;
static const float SH_C3_5 = 1.445305721320277f;
static const float SH_C3_6 = -0.5900435899266435f;

/* animation_coefficient (cull.glsl:125-132): the int8 delta of coefficient j
 * of row `row` in `state`, from the words as the renderer packs them (four
 * signed bytes per word, src/renderer.cpp:852-855). */
static float or_anim_state(const OrAnimation* an, float base, uint32_t row, int j, int state) {
    if (state == 0) return base;
    uint32_t offset = ((uint32_t)(state - 1) * an->count + row) * 48u + (uint32_t)j;
    uint32_t word = an->words[offset / 4u];
    int q = (int)((word >> ((offset % 4u) * 8u)) & 255u);
    if (q >= 128) q -= 256;
    return base + (float)q * an->scales[j];
}

/* fetch_vec4's delta branch (cull.glsl:150-157): state_a, lerped to state_b. */
static float or_anim_coef(const OrAnimation* an, float base, uint32_t row, int j) {
    float a = or_anim_state(an, base, row, j, an->state_a);
    if (an->weight > 0.0f) {
        float b = or_anim_state(an, base, row, j, an->state_b);
        a = (1.0f - an->weight) * a + an->weight * b;
    }
    return a;
}

/* fetch_vec4's layer branch (cull.glsl:143-149): half coefficient j of row
 * `row` in `layer`, two halves per word (src/renderer.cpp:849-851). */
static float or_layer_coef(const OrAnimation* an, int layer, uint32_t row, int j) {
    uint32_t offset = ((uint32_t)layer * an->count + row) * 48u + (uint32_t)j;
    uint32_t pair = an->words[offset / 2u];
    return or_f16_to_f32((uint16_t)((pair >> ((offset % 2u) * 16u)) & 65535u));
}

/* Channel-major coefficient index of rest value r (RGB triples, r = 3 k + c):
 * j = 3 + 15 c + k (cull.glsl:141). */
static inline int or_rest_j(int r) { return 3 + (r % 3) * 15 + r / 3; }

/* SH result of one channel before the +0.5 offset (cull.glsl:354-408). */
static float or_sh_result(const float* dc, const float* sh, int ch, int eff, float dx, float dy, float dz);

static void or_cull_one(const GsGaussian* gs, uint32_t idx, const GsCameraUniforms* cam,
                        const GsSplatEffectParams* fx, int sh_degree, float sh_lod_distance,
                        int collect_stats, const OrAnimation* an, uint32_t* keys, uint32_t* ids,
                        float* projected, uint32_t* visible_count, uint32_t diag[8]) {
    const float* V = cam->view;   /* V[c*4 + r] = view[c][r] */
    const float* P = cam->proj;
    float position[3] = {gs->position[0], gs->position[1], gs->position[2]};
    float opacity = gs->opacity;
    float scale[3] = {gs->scale[0], gs->scale[1], gs->scale[2]};
    float qw = gs->rotation[0], qx = gs->rotation[1], qy = gs->rotation[2], qz = gs->rotation[3];
    float dc[3] = {gs->color[0], gs->color[1], gs->color[2]};
    float* rec = projected + (size_t)idx * 16;

    /* effect (cull.glsl:183-207) */
    float effect_wave = 0.0f;
    if (fx && fx->params[3] > 0.5f) {
        float duration = fmaxf(fx->params[1], 0.001f);
        float progress = or_clampf(fx->params[0] / duration, 0.0f, 1.0f);
        float scene_radius = fmaxf(fx->center_radius[3], 0.001f);
        float to_splat[3], jitter[3];
        for (int k = 0; k < 3; ++k) to_splat[k] = position[k] - fx->center_radius[k];
        float jscale = scene_radius * 0.015f;
        float fidx = (float)idx;
        for (int k = 0; k < 3; ++k) jitter[k] = (or_hash11(fidx + (float)(k + 1)) * 2.0f - 1.0f) * jscale;
        float dist = sqrtf((to_splat[0] * to_splat[0] + to_splat[1] * to_splat[1]) + to_splat[2] * to_splat[2]);
        float wave_radius = progress * scene_radius * 1.15f;
        float band_width = scene_radius * 0.085f;
        float band = (dist - wave_radius) / band_width;
        float wave = or_expf(-band * band);
        float attack = or_smoothstep(0.0f, 0.08f, progress);
        float release = 1.0f - or_smoothstep(0.82f, 1.0f, progress);
        effect_wave = wave * attack * release;
        float d[3];
        d[0] = (to_splat[0] + jitter[0]) + 0.0f;
        d[1] = (to_splat[1] + jitter[1]) + scene_radius * 0.01f;
        d[2] = (to_splat[2] + jitter[2]) + 0.0f;
        float inv = 1.0f / sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) d[k] = d[k] * inv;
        float signed_ripple = or_sinf(progress * 18.849556f + dist * 8.0f / scene_radius);
        float push = effect_wave * fx->params[2] * scene_radius;
        for (int k = 0; k < 3; ++k) position[k] = position[k] + d[k] * push;
        float ripple = effect_wave * signed_ripple * 0.012f * scene_radius;
        for (int k = 0; k < 3; ++k) position[k] = position[k] + d[k] * ripple;
        float sgrow = 1.0f + effect_wave * 0.75f;
        for (int k = 0; k < 3; ++k) scale[k] = scale[k] * sgrow;
        opacity = opacity * (1.0f - effect_wave * 0.20f);
    }

    /* p_view = view * vec4(position, 1) (cull.glsl:209-210) */
    float pv[3];
    for (int r = 0; r < 3; ++r)
        pv[r] = ((V[0 * 4 + r] * position[0] + V[1 * 4 + r] * position[1]) + V[2 * 4 + r] * position[2]) + V[3 * 4 + r] * 1.0f;

    /* visibility (cull.glsl:212-220) */
    int visible = pv[2] <= -0.2f;
    if (visible) {
        float inv_z = -1.0f / pv[2];
        float o = cam->orthographic;
        float ndc_x = (P[0] * pv[0]) * inv_z * (1.0f - o) + P[0] * pv[0] * o;
        float ndc_y = (P[5] * pv[1]) * inv_z * (1.0f - o) + P[5] * pv[1] * o;
        visible = fabsf(ndc_x) <= 1.3f && fabsf(ndc_y) <= 1.3f;
    }
    if (!visible) {
        memset(rec, 0, 16 * sizeof(float));
        rec[8] = 1.0f; rec[9] = 0.0f; rec[10] = 1.0f; rec[11] = 1.0f;
        return;
    }

    /* visible_count, id and far-to-near key (cull.glsl:230-232, 91-99) */
#ifdef _OPENMP
#pragma omp atomic
#endif
    (*visible_count)++;
    ids[idx] = idx;
    float depth = fmaxf(-pv[2], 0.0f);
    keys[idx] = 0xFFFFFFFFu - fbits(depth);

    /* 3D covariance (cull.glsl:234-254) */
    mat3 R;
    R.m[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz);
    R.m[0][1] = 2.0f * (qx * qy + qw * qz);
    R.m[0][2] = 2.0f * (qx * qz - qw * qy);
    R.m[1][0] = 2.0f * (qx * qy - qw * qz);
    R.m[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz);
    R.m[1][2] = 2.0f * (qy * qz + qw * qx);
    R.m[2][0] = 2.0f * (qx * qz + qw * qy);
    R.m[2][1] = 2.0f * (qy * qz - qw * qx);
    R.m[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
    mat3 S;
    memset(&S, 0, sizeof(S));
    S.m[0][0] = scale[0]; S.m[1][1] = scale[1]; S.m[2][2] = scale[2];
    mat3 M = mat3_mul(&R, &S);
    mat3 Mt = mat3_transpose(&M);
    mat3 cov3d = mat3_mul(&M, &Mt);

    /* EWA projection (cull.glsl:256-277) */
    float z = pv[2], x = pv[0], y = pv[1];
    float fx_p = cam->persp_focal, fy_p = cam->persp_focal;
    float fx_o = cam->ortho_focal, fy_o = cam->ortho_focal;
    float ortho = cam->orthographic;
    float J00 = or_mix(fx_p / z, -fx_o, ortho);
    float J11 = or_mix(fy_p / z, -fy_o, ortho);
    float J02 = or_mix(-fx_p * x / (z * z), 0.0f, ortho);
    float J12 = or_mix(-fy_p * y / (z * z), 0.0f, ortho);
    mat3 W;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) W.m[c][r] = V[c * 4 + r];
    mat3 Wt = mat3_transpose(&W);
    mat3 T1 = mat3_mul(&W, &cov3d);
    mat3 WcW = mat3_mul(&T1, &Wt);
    float a = J00 * J00 * WcW.m[0][0] + 2.0f * J00 * J02 * WcW.m[0][2] + J02 * J02 * WcW.m[2][2];
    float b = J00 * J11 * WcW.m[0][1] + J00 * J12 * WcW.m[0][2] + J02 * J11 * WcW.m[1][2] + J02 * J12 * WcW.m[2][2];
    float c = J11 * J11 * WcW.m[1][1] + 2.0f * J11 * J12 * WcW.m[1][2] + J12 * J12 * WcW.m[2][2];
    a += 0.3f;
    c += 0.3f;
    float det = a * c - b * b;
    if (det < 1e-6f) det = 1e-6f;
    float conic[3] = {c / det, -b / det, a / det};

    /* raster centre (cull.glsl:279-291) */
    float pcx = fx_p * x / z + cam->viewport[0] * 0.5f;
    float pcy = cam->viewport[1] * 0.5f + fy_p * y / z;
    float ocx = -fx_o * x + cam->viewport[0] * 0.5f;
    float ocy = cam->viewport[1] * 0.5f - fy_o * y;
    float cx = or_mix(pcx, ocx, ortho);
    float cy = or_mix(pcy, ocy, ortho);
    float y_flip = (1.0f - cam->clip_y_sign) * 0.5f;
    float rcy = or_mix(cy, cam->viewport[1] - cy, y_flip);
    float rconic_y = conic[1] * cam->clip_y_sign;
    float radius_x = ceilf(3.0f * sqrtf(a));
    float radius_y = ceilf(3.0f * sqrtf(c));

    /* diagnostics (cull.glsl:296-318) */
    if (collect_stats) {
        float quad_area = fmaxf(0.0f, 4.0f * radius_x * radius_y);
        uint32_t qpx = or_f2u_sat(fminf(quad_area, 4294967295.0f));
        float gps_points = 6.28318530718f * sqrtf(fmaxf(det, 0.0f)) * or_dilog(or_clampf(opacity, 0.0f, 1.0f));
        uint32_t gpc = or_f2u_sat(fminf(ceilf(gps_points), 4294967295.0f));
#ifdef _OPENMP
#pragma omp critical(or_diag)
#endif
        {
            diag[0] += (qpx + 1023u) >> 10u;
            if (qpx > diag[1]) diag[1] = qpx;
            if (qpx > 1024u) diag[2]++;
            if (qpx > 16384u) diag[3]++;
            diag[4] += (gpc + 1023u) >> 10u;
            if (gpc > diag[5]) diag[5] = gpc;
            if (gpc > 1024u) diag[6]++;
            if (gpc > 16384u) diag[7]++;
        }
    }

    /* ndc depth (cull.glsl:320-321) */
    float ndc_z = (P[2 * 4 + 2] * z + P[3 * 4 + 2]) / (-z);
    ndc_z = or_mix(ndc_z, ndc_z * 0.5f + 0.5f, cam->clip_z_01);

    /* view direction (cull.glsl:323-324): cam = -transpose(W) * view[3].xyz */
    float cpos[3];
    for (int r = 0; r < 3; ++r)
        cpos[r] = -((Wt.m[0][r] * V[12] + Wt.m[1][r] * V[13]) + Wt.m[2][r] * V[14]);
    float dv[3] = {position[0] - cpos[0], position[1] - cpos[1], position[2] - cpos[2]};
    float dinv = 1.0f / sqrtf((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]);
    float dx = dv[0] * dinv, dy = dv[1] * dinv, dz = dv[2] * dinv;

    /* SH (cull.glsl:348-416) */
    int eff = sh_degree;
    if (sh_lod_distance > 0.0f && (-z) > sh_lod_distance) eff = 0;
    const int animated = an && an->count > 0;
    const int layers = animated ? an->layer_count : 0;
    float sh[45];
    for (int r = 0; r < 45; ++r) sh[r] = gs->sh_rest[r];
    if (animated && layers == 0) {
        for (int ch = 0; ch < 3; ++ch) dc[ch] = or_anim_coef(an, dc[ch], idx, ch);
        for (int r = 0; r < 45; ++r) sh[r] = or_anim_coef(an, sh[r], idx, or_rest_j(r));
    }
    float color[3] = {0.0f, 0.0f, 0.0f};
    for (int layer = 0; layer < (layers > 1 ? layers : 1); ++layer) {
        if (layers > 0) {
            for (int ch = 0; ch < 3; ++ch) dc[ch] = or_layer_coef(an, layer, idx, ch);
            for (int r = 0; r < 45; ++r) sh[r] = or_layer_coef(an, layer, idx, or_rest_j(r));
        }
        for (int ch = 0; ch < 3; ++ch) {
            float res = or_sh_result(dc, sh, ch, eff, dx, dy, dz);
            if (layers == 0) res += 0.5f;
            const float w = layers > 0 ? an->layer_weights[layer * 3 + ch] : 1.0f;
            color[ch] = color[ch] + fmaxf(res, 0.0f) * w;
        }
    }
    for (int ch = 0; ch < 3; ++ch) {
        float ecol = fx ? fx->color[ch] : 0.0f;
        float ea = fx ? fx->color[3] : 0.0f;
        color[ch] += ecol * (effect_wave * ea);
    }

    /* projected record (cull.glsl:420-423) */
    rec[0] = color[0]; rec[1] = color[1]; rec[2] = color[2]; rec[3] = opacity;
    rec[4] = cx; rec[5] = rcy; rec[6] = radius_x; rec[7] = radius_y;
    rec[8] = conic[0]; rec[9] = rconic_y; rec[10] = conic[2]; rec[11] = ndc_z;
    rec[12] = a; rec[13] = b * cam->clip_y_sign; rec[14] = c; rec[15] = det;
}

static float or_sh_result(const float* dc, const float* sh, int ch, int eff, float dx, float dy, float dz) {
    float res = SH_C0 * dc[ch];
    if (eff >= 1) {
        res += -SH_C1 * dy * sh[0 * 3 + ch];
        res += SH_C1 * dz * sh[1 * 3 + ch];
        res += -SH_C1 * dx * sh[2 * 3 + ch];
        if (eff >= 2) {
            float xx = dx * dx, yy = dy * dy, zz = dz * dz;
            float xy = dx * dy, yz = dy * dz, xz = dz * dx;
            res += SH_C2_0 * xy * sh[3 * 3 + ch];
            res += SH_C2_1 * yz * sh[4 * 3 + ch];
            res += SH_C2_2 * (2.0f * zz - xx - yy) * sh[5 * 3 + ch];
            res += SH_C2_3 * xz * sh[6 * 3 + ch];
            res += SH_C2_4 * (xx - yy) * sh[7 * 3 + ch];
            if (eff >= 3) {
                res += SH_C3_0 * dy * (3.0f * xx - yy) * sh[8 * 3 + ch];
                res += SH_C3_1 * xy * dz * sh[9 * 3 + ch];
                res += SH_C3_2 * dy * (4.0f * zz - xx - yy) * sh[10 * 3 + ch];
                res += SH_C3_3 * dz * (2.0f * zz - 3.0f * xx - 3.0f * yy) * sh[11 * 3 + ch];
                res += SH_C3_4 * dx * (4.0f * zz - xx - yy) * sh[12 * 3 + ch];
                res += SH_C3_5 * dz * (xx - yy) * sh[13 * 3 + ch];
                res += SH_C3_6 * dx * (xx - 3.0f * yy) * sh[14 * 3 + ch];
            }
        }
    }
    return res;
}

void or_cull(const GsGaussian* g, uint32_t n, const GsCameraUniforms* cam,
             const GsSplatEffectParams* effect, int sh_degree, float sh_lod_distance,
             int collect_stats, uint32_t* keys, uint32_t* ids, float* projected,
             uint32_t* visible_count, uint32_t diag[8], int threads) {
    or_cull_anim(g, n, cam, effect, sh_degree, sh_lod_distance, collect_stats, NULL, keys, ids, projected,
                 visible_count, diag, threads);
}

void or_cull_anim(const GsGaussian* g, uint32_t n, const GsCameraUniforms* cam,
                  const GsSplatEffectParams* effect, int sh_degree, float sh_lod_distance,
                  int collect_stats, const OrAnimation* animation, uint32_t* keys, uint32_t* ids,
                  float* projected, uint32_t* visible_count, uint32_t diag[8], int threads) {
    /* cull_reset_cs (cull.glsl:456-473) */
    *visible_count = 0;
    for (int k = 0; k < 8; ++k) diag[k] = 0;
    for (uint32_t i = 0; i < n; ++i) { keys[i] = 0xFFFFFFFFu; ids[i] = 0xFFFFFFFFu; }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
    for (long long i = 0; i < (long long)n; ++i)
        or_cull_one(g + i, (uint32_t)i, cam, effect, sh_degree, sh_lod_distance, collect_stats, animation,
                    keys, ids, projected, visible_count, diag);
    (void)threads;
}

/* ======================================================================
 * radix_hist_cs / radix_prefix_cs / radix_scatter_cs (shaders/sort.glsl:4-164)
 * as dispatched by renderer_sort_gaussians_gpu (src/renderer.cpp:1884-1961).
 * ====================================================================== */
void or_radix_sort(uint32_t* keys, uint32_t* ids, uint32_t n) {
    if (n <= 1) return;
    uint32_t groups = (n + 255u) / 256u;
    uint32_t* hist = (uint32_t*)malloc((size_t)groups * 256u * sizeof(uint32_t));
    uint32_t* tk = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* ti = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* kin = keys; uint32_t* iin = ids; uint32_t* kout = tk; uint32_t* iout = ti;
    for (int pass = 0; pass < 4; ++pass) {
        int shift = pass * 8;
        /* radix_hist_cs */
        memset(hist, 0, (size_t)groups * 256u * sizeof(uint32_t));
        for (uint32_t gid = 0; gid < n; ++gid)
            hist[(gid / 256u) * 256u + ((kin[gid] >> shift) & 255u)]++;
        /* radix_prefix_cs: group-major offsets per bin */
        uint32_t totals[256];
        for (uint32_t bin = 0; bin < 256; ++bin) {
            uint32_t t = 0;
            for (uint32_t g = 0; g < groups; ++g) t += hist[g * 256u + bin];
            totals[bin] = t;
        }
        uint32_t running = 0;
        for (uint32_t bin = 0; bin < 256; ++bin) {
            uint32_t r = running;
            running += totals[bin];
            for (uint32_t g = 0; g < groups; ++g) {
                uint32_t cnt = hist[g * 256u + bin];
                hist[g * 256u + bin] = r;
                r += cnt;
            }
        }
        /* radix_scatter_cs: rank = lower lanes of the group with the same digit */
        for (uint32_t g = 0; g < groups; ++g) {
            uint32_t seen[256];
            memset(seen, 0, sizeof(seen));
            uint32_t end = (g + 1u) * 256u < n ? (g + 1u) * 256u : n;
            for (uint32_t gid = g * 256u; gid < end; ++gid) {
                uint32_t bin = (kin[gid] >> shift) & 255u;
                uint32_t out = hist[g * 256u + bin] + seen[bin]++;
                kout[out] = kin[gid];
                iout[out] = iin[gid];
            }
        }
        uint32_t* t;
        t = kin; kin = kout; kout = t;
        t = iin; iin = iout; iout = t;
    }
    /* four passes: result is back in keys/ids */
    free(hist); free(tk); free(ti);
}

/* ======================================================================
 * Sorted-alpha composite: splat_vs/splat_fs (shaders/splat.glsl:6-105),
 * blend ONE / ONE_MINUS_SRC_ALPHA into RGBA16F cleared to 0
 * (src/renderer.cpp:277-299, 2534-2541), quad coverage of the two triangles
 * centre +- radius with pixel centres at +0.5 and 1/256-pixel vertex snapping.
 * ====================================================================== */

static int32_t or_snap(float w) {
    float s = w * 256.0f;
    if (s > 1073741824.0f) return 1073741824;
    if (s < -1073741824.0f) return -1073741824;
    return (int32_t)rintf(s);
}

void or_quad_fixed(const float* rec, float width, float height, float jitter_x, float jitter_y, int32_t out[4]) {
    float cx = rec[4], cy = rec[5], rx = rec[6], ry = rec[7];
    for (int k = 0; k < 2; ++k) {
        float corner = k ? 1.0f : -1.0f;
        float px = (cx + corner * rx) + jitter_x;  /* splat_vs: centre + corner * radius + jitter */
        float py = (cy + corner * ry) + jitter_y;
        float nx = 2.0f * px / width - 1.0f;
        float ny = 2.0f * py / height - 1.0f;
        float wx = (nx + 1.0f) * (0.5f * width);
        float wy = (ny + 1.0f) * (0.5f * height);
        out[k] = or_snap(wx);
        out[2 + k] = or_snap(wy);
    }
}

static inline int32_t or_ceil_div256(int32_t a) { return -((-a) >> 8); }

/* splat_vs emits gl_Position = (ndc, ndc_z, 1) (shaders/splat.glsl:63-67) and
 * the pipelines enable no depth clamp (src/renderer.cpp:277-320), so the
 * rasteriser clips the whole quad (all four corners share ndc_z) unless
 * -w <= z <= w, i.e. -1 <= ndc_z <= 1 for GL clip space, 0 <= ndc_z <= 1 when
 * clip_z_01 selects the [0, 1] depth range.  With the camera's far plane at
 * 100 (src/camera.cpp:12) this drops splats whose view depth exceeds 100. */
int or_clipped(float ndc_z, float clip_z_01) {
    const float lo = clip_z_01 > 0.5f ? 0.0f : -1.0f;
    return !(ndc_z <= 1.0f && ndc_z >= lo);
}

void or_composite(const float* projected, const uint32_t* sorted_ids, uint32_t n,
                  int width, int height, float clip_y_sign, float clip_z_01, int mode, double* out,
                  int threads) {
    const int band = 16;
    const int nbands = (height + band - 1) / band;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int bi = 0; bi < nbands; ++bi) {
        int j0 = bi * band, j1 = j0 + band < height ? j0 + band : height;
        size_t npx = (size_t)(j1 - j0) * width;
        double* acc = (double*)calloc(npx * 5, sizeof(double)); /* rgba (+T for front-to-back) */
        float* acc16 = (float*)calloc(npx * 4, sizeof(float));
        if (mode == 1)
            for (size_t p = 0; p < npx; ++p) acc[p * 5 + 4] = 1.0;
        for (uint32_t s = 0; s < n; ++s) {
            uint32_t id = sorted_ids[mode == 1 ? n - 1 - s : s];
            if (id == 0xFFFFFFFFu) continue;
            const float* rec = projected + (size_t)id * 16;
            if (or_clipped(rec[11], clip_z_01)) continue;
            int32_t q[4];
            or_quad_fixed(rec, (float)width, (float)height, 0.0f, 0.0f, q);
            int32_t i_lo = or_ceil_div256(q[0] - 128), i_hi = or_ceil_div256(q[1] - 128);
            int32_t y_lo = or_ceil_div256(q[2] - 128), y_hi = or_ceil_div256(q[3] - 128);
            if (i_lo < 0) i_lo = 0;
            if (i_hi > width) i_hi = width;
            if (y_lo < j0) y_lo = j0;
            if (y_hi > j1) y_hi = j1;
            if (i_lo >= i_hi || y_lo >= y_hi) continue;
            float o = rec[3];
            float cutoff = or_logf((1.0f / 255.0f) / fmaxf(o, 1e-8f));
            float cx = rec[4], cy = rec[5];
            float k0 = rec[8], k1 = rec[9], k2 = rec[10];
            for (int32_t j = y_lo; j < y_hi; ++j) {
                float dy = ((float)j + 0.5f) - cy;
                for (int32_t i = i_lo; i < i_hi; ++i) {
                    float dx = ((float)i + 0.5f) - cx;
                    float power = -0.5f * (k0 * dx * dx + 2.0f * k1 * dx * dy + k2 * dy * dy);
                    if (power > 0.0f) continue;
                    if (power < cutoff) continue;
                    size_t p = (size_t)(j - j0) * width + i;
                    if (mode == 2) {
                        float alpha = fminf(o * or_expf(power), 0.99f);
                        float src[4] = {rec[0] * alpha, rec[1] * alpha, rec[2] * alpha, alpha};
                        for (int ch = 0; ch < 4; ++ch) {
                            float d = acc16[p * 4 + ch];
                            float v = src[ch] + d * (1.0f - alpha);
                            acc16[p * 4 + ch] = or_f16_to_f32(or_f32_to_f16(v));
                        }
                    } else {
                        double alpha = (double)o * or_exp_d((double)power);
                        if (alpha > 0.99) alpha = 0.99;
                        if (mode == 0) {
                            for (int ch = 0; ch < 3; ++ch)
                                acc[p * 5 + ch] = (double)rec[ch] * alpha + acc[p * 5 + ch] * (1.0 - alpha);
                            acc[p * 5 + 3] = alpha + acc[p * 5 + 3] * (1.0 - alpha);
                        } else {
                            double T = acc[p * 5 + 4];
                            for (int ch = 0; ch < 3; ++ch) acc[p * 5 + ch] += (double)rec[ch] * alpha * T;
                            acc[p * 5 + 4] = T * (1.0 - alpha);
                        }
                    }
                }
            }
        }
        for (int32_t j = j0; j < j1; ++j) {
            int row = clip_y_sign > 0.0f ? height - 1 - j : j;
            for (int i = 0; i < width; ++i) {
                size_t p = (size_t)(j - j0) * width + i;
                double* dst = out + ((size_t)row * width + i) * 4;
                if (mode == 2) {
                    for (int ch = 0; ch < 4; ++ch) dst[ch] = acc16[p * 4 + ch];
                } else if (mode == 0) {
                    for (int ch = 0; ch < 4; ++ch) dst[ch] = acc[p * 5 + ch];
                } else {
                    for (int ch = 0; ch < 3; ++ch) dst[ch] = acc[p * 5 + ch];
                    dst[3] = 1.0 - acc[p * 5 + 4];
                }
            }
        }
        free(acc);
        free(acc16);
    }
    (void)threads;
}

/* ======================================================================
 * GPS (shaders/gps.glsl) and accumulation (shaders/accum.glsl)
 * ====================================================================== */

void or_gps_expand(const float* projected, const uint32_t* ids, uint32_t n,
                   uint32_t max_work_items, float ss_in, float clip_z_01, int linear_color,
                   uint32_t* prepared, uint32_t* work, uint32_t stats[2]) {
    uint32_t work_count = 0, overflow = 0;
    uint32_t capacity = max_work_items;
    for (uint32_t idx = 0; idx < n; ++idx) {
        uint32_t splat_id = ids[idx];
        if (splat_id == 0xFFFFFFFFu) continue;
        const float* rec = projected + (size_t)splat_id * 16;
        float o = rec[3];
        float ss = fmaxf(ss_in, 1.0f);
        float cov_a = rec[12] + 0.3f;
        float cov_b = rec[13];
        float cov_c = rec[14] + 0.3f;
        float det = cov_a * cov_c - cov_b * cov_b;
        float expected_points = ss * ss * 6.28318530718f * sqrtf(fmaxf(det, 0.0f)) *
            or_dilog(or_clampf(o, 0.0f, 1.0f));
        float ppw = fmaxf(ss * ss, 1.0f);
        float scaled = fmaxf(expected_points / ppw, 0.0f);
        float fl = or_clampf(floorf(scaled), 0.0f, 4294967295.0f);
        uint32_t count = or_f2u_sat(fl);
        if (or_hash01(splat_id + 0x9e3779b9u) < or_fract(scaled) && count < 0xFFFFFFFFu) count += 1u;
        if (count == 0u) continue;
        uint32_t base = work_count;
        work_count += count;
        if (base >= capacity || count > capacity - base) overflow = 1u;
        if (base >= capacity) continue;
        uint32_t write_count = count < capacity - base ? count : capacity - base;
        float l00 = sqrtf(fmaxf(cov_a, 1.0e-5f));
        float l10 = cov_b / l00;
        float l11 = sqrtf(fmaxf(cov_c - l10 * l10, 1.0e-5f));
        float alpha = or_clampf(o, 1.0e-6f, 1.0f);
        float dilog_alpha = or_dilog(alpha);
        float ndc_depth = rec[11];
        float z = or_mix(ndc_depth * 0.5f + 0.5f, ndc_depth, clip_z_01);
        uint32_t depth_key = or_f2u_sat(or_clampf(z, 0.0f, 1.0f) * 4294967294.0f);
        uint32_t rgb[3];
        for (int ch = 0; ch < 3; ++ch) rgb[ch] = (uint32_t)rintf(or_clampf(rec[ch], 0.0f, 1.0f) * 255.0f);
        uint32_t packed = rgb[0] | (rgb[1] << 8) | (rgb[2] << 16) | 0xFF000000u;
        if (linear_color) packed = splat_id;
        uint32_t* pr = prepared + (size_t)idx * 12;
        float f[8] = {rec[4], rec[5], l00, l10, l11, alpha, dilog_alpha, 0.0f};
        memcpy(pr, f, sizeof(f));
        pr[8] = depth_key; pr[9] = packed; pr[10] = splat_id; pr[11] = 0u;
        for (uint32_t i = 0; i < write_count; ++i) {
            work[(size_t)(base + i) * 2] = idx;
            work[(size_t)(base + i) * 2 + 1] = i;
        }
    }
    stats[0] = work_count;
    stats[1] = overflow;
}

void or_gps_splat(const uint32_t* prepared, const uint32_t* work, uint32_t work_count,
                  int gps_width, int gps_height, uint32_t frame_seed, float ss,
                  uint64_t* cells) {
    uint32_t fs = (uint32_t)(float)frame_seed;  /* float UBO field, uint(frame_seed) */
    int ppw = (int)rintf(ss * ss);
    if (ppw < 1) ppw = 1;
    float ssm = fmaxf(ss, 1.0f);
    for (uint32_t w = 0; w < work_count; ++w) {
        uint32_t idx = work[(size_t)w * 2], sample_idx = work[(size_t)w * 2 + 1];
        const uint32_t* pr = prepared + (size_t)idx * 12;
        float f[8];
        memcpy(f, pr, sizeof(f));
        float mx = f[0], my = f[1], l00 = f[2], l10 = f[3], l11 = f[4], alpha = f[5], dla = f[6];
        uint32_t depth_key = pr[8], packed = pr[9], splat_id = pr[10];
        uint32_t seed0 = splat_id * 747796405u + fs * 2891336453u;
        uint32_t seed1 = splat_id * 277803737u + 0x9E3779B9u;
        uint64_t value = ((uint64_t)depth_key << 32) | packed;
        for (int p = 0; p < ppw; ++p) {
            uint32_t v[2] = {seed0 + (sample_idx * (uint32_t)ppw + (uint32_t)p) * 374761393u, seed1};
            or_pcg2d(v);
            float r0 = (float)v[0] * 2.32830643654e-10f;
            float r1 = (float)v[1] * 2.32830643654e-10f;
            float arg = or_inv_dilog((1.0f - r0) * dla) / alpha;
            arg = or_clampf(arg, 1.0e-7f, 1.0f);
            float radius = sqrtf(-2.0f * or_logf_fma(arg));
            float az = 6.28318530718f * r1;
            float sn, cs;
            or_gps_sincosf(az, &sn, &cs);
            float lx = cs * radius, ly = sn * radius;
            float ptx = (mx + l00 * lx) * ssm;
            float pty = (my + (l10 * lx + l11 * ly)) * ssm;
            int px = (int)floorf(ptx + 0.5f), py = (int)floorf(pty + 0.5f);
            if (px >= 0 && py >= 0 && px < gps_width && py < gps_height) {
                uint64_t* cell = cells + (size_t)py * gps_width + px;
                if (value < *cell) *cell = value;
            }
        }
    }
}

void or_gps_resolve(const uint64_t* cells, int width, int height, int ss_in, float background,
                    int linear_color, const float* projected, float* color, float* depth) {
    int ss = ss_in < 1 ? 1 : (ss_in > 4 ? 4 : ss_in);
    int gw = width * ss, gh = height * ss;
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            float sum[3] = {0.0f, 0.0f, 0.0f};
            float cov = 0.0f;
            uint32_t min_key = 0xFFFFFFFFu;
            for (int yy = 0; yy < ss; ++yy) {
                for (int xx = 0; xx < ss; ++xx) {
                    int sx = x * ss + xx, sy = y * ss + yy;
                    if (sx >= gw || sy >= gh) continue;
                    uint64_t cell = cells[(size_t)sy * gw + sx];
                    uint32_t key = (uint32_t)(cell >> 32);
                    if (key != 0xFFFFFFFFu) {
                        uint32_t payload = (uint32_t)cell;
                        if (linear_color) {
                            const float* rec = projected + (size_t)payload * 16;
                            for (int ch = 0; ch < 3; ++ch) sum[ch] += rec[ch];
                        } else {
                            sum[0] += (float)(payload & 255u) / 255.0f;
                            sum[1] += (float)((payload >> 8) & 255u) / 255.0f;
                            sum[2] += (float)((payload >> 16) & 255u) / 255.0f;
                        }
                        cov += 1.0f;
                        if (key < min_key) min_key = key;
                    }
                }
            }
            float* c = color + ((size_t)y * width + x) * 4;
            if (cov <= 0.0f) {
                c[0] = c[1] = c[2] = background;
                c[3] = 0.0f;
                depth[(size_t)y * width + x] = 1.0f;
                continue;
            }
            float sub = (float)(ss * ss);
            for (int ch = 0; ch < 3; ++ch) c[ch] = sum[ch] / sub;
            c[3] = cov / sub;
            depth[(size_t)y * width + x] = (float)min_key / 4294967294.0f;
        }
    }
}

/* hash13 of shaders/splat.glsl:121-125, float operations in source order
 * (dot product summed left to right). */
float or_hash13(float x, float y, float z) {
    float px = or_fract(x * 0.1031f), py = or_fract(y * 0.1031f), pz = or_fract(z * 0.1031f);
    float d = (px * (py + 33.33f) + py * (pz + 33.33f)) + pz * (px + 33.33f);
    px = px + d;
    py = py + d;
    pz = pz + d;
    return or_fract((px + py) * pz);
}

/* 24-bit window depth of an ndc depth: GL depth range [0, 1]
 * (z_w = 0.5 z + 0.5, or z itself for clip_z_01), clamped, to a D24 value
 * with round to nearest. */
uint32_t or_depth24(float ndc_z, float clip_z_01) {
    float d = clip_z_01 > 0.5f ? ndc_z : ndc_z * 0.5f + 0.5f;
    d = fminf(fmaxf(d, 0.0f), 1.0f);
    return (uint32_t)rint((double)d * 16777215.0);
}

/* StochasticSplats sample pass (shaders/splat.glsl:107-146 with the splat_vs
 * quad, pipeline src/renderer.cpp:302-320, pass src/renderer.cpp:2331-2352,
 * draw order src/renderer.cpp:1688-1716): the visible splats are drawn in
 * index order (ids[i] = i or 0xFFFFFFFF); a covered fragment that passes the
 * alpha test is kept when hash13(gl_FragCoord.xy, id + frame_seed) < alpha,
 * and written as (rgb, 1) when its 24-bit depth is LESS than the stored one
 * (cleared to 1.0); the colour target holds half floats (RGBA16F).  Clear
 * colour (bg, bg, bg, 0).  color: W*H*4, depth: W*H D24 depths of the kept
 * splat / (2^24 - 1) (1 where nothing was kept), framebuffer row order. */
void or_stochastic_splats(const float* projected, const uint32_t* ids, uint32_t n, int width, int height,
                          float clip_z_01, float frame_seed, float background, float jitter_x, float jitter_y,
                          float* color, float* depth, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int j = 0; j < height; ++j) {
        uint32_t* best = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)width);
        for (int i = 0; i < width; ++i) {
            best[i] = 16777215u;
            float* c = color + ((size_t)j * width + i) * 4;
            c[0] = c[1] = c[2] = background;
            c[3] = 0.0f;
            depth[(size_t)j * width + i] = 1.0f;
        }
        for (uint32_t s = 0; s < n; ++s) {
            const uint32_t id = ids[s];
            if (id == 0xFFFFFFFFu) continue;
            const float* rec = projected + (size_t)id * 16;
            if (or_clipped(rec[11], clip_z_01)) continue;
            /* the quad corners and the fragment centre move by the TAA jitter
             * (shaders/splat.glsl:61,72) */
            int32_t q[4];
            or_quad_fixed(rec, (float)width, (float)height, jitter_x, jitter_y, q);
            int32_t i_lo = or_ceil_div256(q[0] - 128), i_hi = or_ceil_div256(q[1] - 128);
            int32_t y_lo = or_ceil_div256(q[2] - 128), y_hi = or_ceil_div256(q[3] - 128);
            if (j < y_lo || j >= y_hi) continue;
            if (i_lo < 0) i_lo = 0;
            if (i_hi > width) i_hi = width;
            const float o = rec[3];
            const float cutoff = or_logf_fma((1.0f / 255.0f) / fmaxf(o, 1e-8f));
            const float cx = rec[4] + jitter_x, cy = rec[5] + jitter_y;
            const float k0 = rec[8], k1 = rec[9], k2 = rec[10];
            const uint32_t dq = or_depth24(rec[11], clip_z_01);
            const float fy = (float)j + 0.5f;
            const float dy = fy - cy;
            const float seed = (float)id + frame_seed;
            for (int32_t i = i_lo; i < i_hi; ++i) {
                const float fx = (float)i + 0.5f;
                const float dx = fx - cx;
                const float power = -0.5f * (k0 * dx * dx + 2.0f * k1 * dx * dy + k2 * dy * dy);
                if (power > 0.0f || power < cutoff) continue;
                const float alpha = fminf(o * or_expf(power), 0.99f);
                if (or_hash13(fx, fy, seed) >= alpha) continue;
                if (!(dq < best[i])) continue;
                best[i] = dq;
                float* c = color + ((size_t)j * width + i) * 4;
                for (int ch = 0; ch < 3; ++ch) c[ch] = or_f16_to_f32(or_f32_to_f16(rec[ch]));
                c[3] = 1.0f;
                depth[(size_t)j * width + i] = (float)((double)dq / 16777215.0);
            }
        }
        free(best);
    }
    (void)threads;
}

/* layers_display_fs (shaders/accum.glsl:197-240) over `count` pixels of
 * linear RGBA floats; lut: size^3 RGB uint16 triples, red fastest. */
static float or_lut_entry(const uint16_t* lut, int size, int x, int y, int z, int c) {
    int offset = ((z * size + y) * size + x) * 3 + c;
    return (float)lut[offset] / 65535.0f;
}

void or_layers_display(const float* src, uint32_t count, const uint16_t* lut, int size, float log2_min,
                       float log2_max, float exposure_ev, float* out) {
    for (uint32_t i = 0; i < count; ++i) {
        int lo[3], hi[3];
        float f[3];
        for (int c = 0; c < 3; ++c) {
            float rad = src[(size_t)i * 4 + c] * exp2f(exposure_ev);
            float s = or_clampf((log2f(fmaxf(rad, exp2f(log2_min))) - log2_min) / (log2_max - log2_min), 0.0f, 1.0f);
            float p = s * (float)(size - 1);
            lo[c] = (int)floorf(p);
            hi[c] = lo[c] + 1 < size - 1 ? lo[c] + 1 : size - 1;
            f[c] = or_fract(p);
        }
        for (int c = 0; c < 3; ++c) {
            float a = or_mix(or_mix(or_lut_entry(lut, size, lo[0], lo[1], lo[2], c),
                                    or_lut_entry(lut, size, hi[0], lo[1], lo[2], c), f[0]),
                             or_mix(or_lut_entry(lut, size, lo[0], hi[1], lo[2], c),
                                    or_lut_entry(lut, size, hi[0], hi[1], lo[2], c), f[0]), f[1]);
            float b = or_mix(or_mix(or_lut_entry(lut, size, lo[0], lo[1], hi[2], c),
                                    or_lut_entry(lut, size, hi[0], lo[1], hi[2], c), f[0]),
                             or_mix(or_lut_entry(lut, size, lo[0], hi[1], hi[2], c),
                                    or_lut_entry(lut, size, hi[0], hi[1], hi[2], c), f[0]), f[1]);
            out[(size_t)i * 4 + c] = or_mix(a, b, f[2]);
        }
        out[(size_t)i * 4 + 3] = 1.0f;
    }
}

/* halton (src/renderer.cpp:246-255). */
float or_halton(uint32_t index, uint32_t base) {
    float f = 1.0f, r = 0.0f;
    while (index > 0) {
        f /= (float)base;
        r += f * (float)(index % base);
        index /= base;
    }
    return r;
}

/* GLSL mat4 * vec4, column-major, summed left to right. */
static void or_mat4_vec4(const float* m, const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = ((m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1]) + m[2 * 4 + r] * v[2]) + m[3 * 4 + r] * v[3];
}

static void or_texel(const float* img, int w, int h, int x, int y, float out[4]) {
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    memcpy(out, img + ((size_t)y * w + x) * 4, 16);
}

/* sample_history_*_bilinear (shaders/accum.glsl:120-152). */
static void or_bilinear(const float* img, int w, int h, float u, float v, float out[4]) {
    float px = u * (float)w - 0.5f, py = v * (float)h - 0.5f;
    int x0 = (int)floorf(px), y0 = (int)floorf(py);
    float fx = or_fract(px), fy = or_fract(py);
    float c00[4], c10[4], c01[4], c11[4];
    or_texel(img, w, h, x0, y0, c00);
    or_texel(img, w, h, x0 + 1, y0, c10);
    or_texel(img, w, h, x0, y0 + 1, c01);
    or_texel(img, w, h, x0 + 1, y0 + 1, c11);
    for (int k = 0; k < 4; ++k)
        out[k] = or_mix(or_mix(c00[k], c10[k], fx), or_mix(c01[k], c11[k], fx), fy);
}

/* taa_accum_fs (shaders/accum.glsl:47-179) over a W x H frame; images are
 * RGBA (W*H*4) or R (W*H) floats in framebuffer row order, u = TaaAccumUBO
 * as 52 floats (inv_view, prev_view, viewport_orthographic, focal_clip,
 * prev_focal_clip, proj_z, history_valid, view_changed, pad, pad). */
void or_taa_accumulate(const float* cur_color, const float* cur_depth, const float* hist_color,
                       const float* hist_xyz, int w, int h, const float* u, float* out_color, float* out_xyz) {
    const float* inv_view = u;
    const float* prev_view = u + 16;
    const float* vo = u + 32;   /* viewport xy, orthographic now, orthographic before */
    const float* fc = u + 36;   /* persp focal, ortho focal, clip_y_sign, clip_z_01 */
    const float* pfc = u + 40;  /* previous persp focal, ortho focal, clip_y_sign */
    const float* pz = u + 44;   /* proj[2][2], proj[3][2] */
    const float history_valid = u[48], view_changed = u[49];
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) {
            const size_t p = (size_t)j * w + i;
            const float uvx = ((float)i + 0.5f) / (float)w, uvy = ((float)j + 0.5f) / (float)h;
            const float* cc = cur_color + p * 4;
            const float depth = cur_depth[p];
            /* reconstruct_view_pos */
            const float z_ndc = depth * 2.0f - 1.0f;
            const float view_z = -pz[1] / (z_ndc + pz[0]);
            float pxx = uvx * vo[0], pxy = uvy * vo[1];
            const float unflip = (1.0f - fc[2]) * 0.5f;
            pxy = or_mix(pxy, vo[1] - pxy, unflip);
            const float cx = pxx - vo[0] * 0.5f, cy = pxy - vo[1] * 0.5f;
            const float denom = or_mix(fc[0] / view_z, -fc[1], vo[2]);
            const float view[4] = {cx / denom, cy / denom, view_z, 1.0f};
            float world[4];
            or_mat4_vec4(inv_view, view, world);
            /* project_prev_uv */
            const float wv[4] = {world[0], world[1], world[2], 1.0f};
            float pv[4];
            or_mat4_vec4(prev_view, wv, pv);
            float pu = -1.0f, pvv = -1.0f;
            if (!(pv[2] >= -0.001f)) {
                const float t = vo[3];
                const float ppx = pfc[0] * pv[0] / pv[2] + vo[0] * 0.5f;
                const float ppy = vo[1] * 0.5f + pfc[0] * pv[1] / pv[2];
                const float opx = -pfc[1] * pv[0] + vo[0] * 0.5f;
                const float opy = vo[1] * 0.5f - pfc[1] * pv[1];
                float qx = or_mix(ppx, opx, t), qy = or_mix(ppy, opy, t);
                const float y_flip = (1.0f - pfc[2]) * 0.5f;
                qy = or_mix(qy, vo[1] - qy, y_flip);
                pu = qx / vo[0];
                pvv = qy / vo[1];
            }
            int valid = history_valid != 0.0f && pu >= 0.0f && pu <= 1.0f && pvv >= 0.0f && pvv <= 1.0f;
            float hc[4] = {0, 0, 0, 0}, hx[4] = {0, 0, 0, 0};
            if (valid) {
                or_bilinear(hist_color, w, h, pu, pvv, hc);
                or_bilinear(hist_xyz, w, h, pu, pvv, hx);
                const float dx = world[0] - hx[0], dy = world[1] - hx[1], dz = world[2] - hx[2];
                valid = hc[3] > 0.0f && (view_changed == 0.0f || sqrtf((dx * dx + dy * dy) + dz * dz) < 100.0f);
            }
            float* oc = out_color + p * 4;
            float* ox = out_xyz + p * 4;
            if (valid) {
                const float n = fminf(hc[3], 128.0f);
                for (int k = 0; k < 3; ++k) oc[k] = (hc[k] * n + cc[k]) / (n + 1.0f);
                oc[3] = fminf(n + 1.0f, 128.0f);
                for (int k = 0; k < 3; ++k) ox[k] = (hx[k] * n + world[k]) / (n + 1.0f);
                ox[3] = 1.0f;
            } else {
                oc[0] = cc[0]; oc[1] = cc[1]; oc[2] = cc[2]; oc[3] = 1.0f;
                ox[0] = world[0]; ox[1] = world[1]; ox[2] = world[2]; ox[3] = 1.0f;
            }
        }
}

void or_accumulate(const float* history, const float* sample, float sample_count,
                   float* out, uint32_t count) {
    float n = fmaxf(sample_count, 1.0f);
    for (uint32_t i = 0; i < count; ++i) out[i] = history[i] + (sample[i] - history[i]) / n;
}
