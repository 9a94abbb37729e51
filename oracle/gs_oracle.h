/*
 * gs_oracle.h - CPU restatement of gamesplat's splat frame pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  Nothing in gamesplat_amd/ may call it.
 *
 * Each function restates one reference kernel (file:line given per function
 * in gs_oracle.c).  Floating-point policy: IEEE binary32 operations in the
 * order the GLSL source writes them, no FMA contraction, correctly rounded
 * division and sqrt.  GLSL's log/exp/sin/cos are evaluated with the
 * deterministic double-precision routines or_log/or_exp/or_sin/or_cos and
 * rounded to float.
 *
 * Parity pinning (tests/test_oracle_*.py): libm for the math routines, numpy
 * for float16 rounding and stable argsort, scipy for the dilogarithm and the
 * real spherical harmonics, closed forms for projection and compositing,
 * generic inversion of the integer hashes.  The reference itself (GLSL under
 * a GL driver) does not build or run in this environment.
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H

#include <stdint.h>
#include "../include/gs_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* math */
double or_log_d(double x);
double or_exp_d(double x);
double or_sin_d(double x);
double or_cos_d(double x);
float or_logf(float x);
float or_expf(float x);
float or_sinf(float x);
float or_cosf(float x);
float or_dilog(float x);
float or_inv_dilog(float x);
float or_logf_fma(float x);
void or_gps_sincosf(float x, float* sn, float* cs);
void or_logf_fma_n(const float* x, float* y, uint32_t n);
void or_gps_sincosf_n(const float* x, float* sn, float* cs, uint32_t n);
uint32_t or_hash01_bits(uint32_t x);
float or_hash01(uint32_t x);
void or_pcg2d(uint32_t v[2]);
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);

/* cull: dense reference semantics.  keys/ids/projected are n-sized and
 * indexed by Gaussian; culled rows keep the 0xFFFFFFFF sentinel and the
 * reference's culled projected record.  effect may be NULL. */
/* AnimationUBO + the packed animation buffer (cull.glsl:70-85): count = rows
 * (0 = no animation); words: int8 deltas four per word, or half layer
 * coefficients two per word; layer_weights: 3 per layer. */
typedef struct OrAnimation {
    uint32_t count;
    int32_t state_a, state_b;
    float weight;
    float scales[48];
    int32_t layer_count;
    const uint32_t* words;
    const float* layer_weights;
} OrAnimation;
void or_cull_anim(const GsGaussian* g, uint32_t n, const GsCameraUniforms* cam,
                  const GsSplatEffectParams* effect, int sh_degree, float sh_lod_distance,
                  int collect_stats, const OrAnimation* animation, uint32_t* keys, uint32_t* ids,
                  float* projected, uint32_t* visible_count, uint32_t diag[8], int threads);
void or_cull(const GsGaussian* g, uint32_t n, const GsCameraUniforms* cam,
             const GsSplatEffectParams* effect, int sh_degree, float sh_lod_distance,
             int collect_stats, uint32_t* keys, uint32_t* ids, float* projected,
             uint32_t* visible_count, uint32_t diag[8], int threads);

/* 4-pass 8-bit LSD radix sort of all n (key, id) pairs, in place. */
void or_radix_sort(uint32_t* keys, uint32_t* ids, uint32_t n);

/* Whether the rasteriser clips a splat quad at this ndc depth (outside
 * [-1, 1], or [0, 1] with clip_z_01). */
int or_clipped(float ndc_z, float clip_z_01);
/* Sorted-alpha composite.  mode 0: fp64 back-to-front, 1: fp64 front-to-back,
 * 2: fp32 shader + fp16 storage after every blend (RGBA16F target).  Quads
 * outside the clip volume's depth range are not drawn.
 * out: W*H*4 doubles, top row first (clip_y_sign > 0 flips rows). */
void or_composite(const float* projected, const uint32_t* sorted_ids, uint32_t n,
                  int width, int height, float clip_y_sign, float clip_z_01, int mode, double* out,
                  int threads);

/* Quad coverage of one projected splat in 24.8 fixed point window
 * coordinates: x0,x1,y0,y1; pixel i is covered iff x0 <= 256 i + 128 < x1. */
void or_quad_fixed(const float* rec, float width, float height, float jitter_x, float jitter_y, int32_t out[4]);

/* GPS expand (shaders/gps.glsl:120-174); items are reserved in index order.
 * prepared: 12 words per Gaussian (mean_basis, sampling, packed) indexed by
 * position; work: 2 words per item; stats: requested, overflow. */
void or_gps_expand(const float* projected, const uint32_t* ids, uint32_t n,
                   uint32_t max_work_items, float ss, float clip_z_01, int linear_color,
                   uint32_t* prepared, uint32_t* work, uint32_t stats[2]);
/* GPS splat (shaders/gps.glsl:232-285) into 64-bit depth<<32|payload cells
 * initialised to 0xFFFFFFFF00000000. */
void or_gps_splat(const uint32_t* prepared, const uint32_t* work, uint32_t work_count,
                  int gps_width, int gps_height, uint32_t frame_seed, float ss,
                  uint64_t* cells);
/* GPS resolve (shaders/gps.glsl:344-385).  color: W*H*4, depth: W*H; both in
 * framebuffer row order (row 0 = first row of the cell buffer). */
void or_gps_resolve(const uint64_t* cells, int width, int height, int ss, float background,
                    int linear_color, const float* projected, float* color, float* depth);
/* hash13 (shaders/splat.glsl:121-125) and the D24 window depth. */
float or_hash13(float x, float y, float z);
uint32_t or_depth24(float ndc_z, float clip_z_01);
/* StochasticSplats sample (shaders/splat.glsl:107-146): splats drawn in index
 * order (quads outside the clip volume's depth range are not drawn), kept when hash13 < alpha, nearest 24-bit depth wins (LESS), colour
 * stored as half floats; clear (bg, bg, bg, 0).  color W*H*4, depth W*H,
 * framebuffer row order. */
void or_stochastic_splats(const float* projected, const uint32_t* ids, uint32_t n, int width, int height,
                          float clip_z_01, float frame_seed, float background, float jitter_x, float jitter_y,
                          float* color, float* depth, int threads);
/* layers_display_fs (shaders/accum.glsl:197-240). */
void or_layers_display(const float* src, uint32_t count, const uint16_t* lut, int size, float log2_min,
                       float log2_max, float exposure_ev, float* out);
/* halton (src/renderer.cpp:246-255). */
float or_halton(uint32_t index, uint32_t base);
/* taa_accum_fs (shaders/accum.glsl:47-179); u = TaaAccumUBO as 52 floats. */
void or_taa_accumulate(const float* cur_color, const float* cur_depth, const float* hist_color,
                       const float* hist_xyz, int w, int h, const float* u, float* out_color, float* out_xyz);
/* accum_fs (shaders/accum.glsl:39-44), elementwise over count floats. */
void or_accumulate(const float* history, const float* sample, float sample_count,
                   float* out, uint32_t count);

#ifdef __cplusplus
}
#endif

#endif
