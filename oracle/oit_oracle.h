/*
 * oit_oracle.h - CPU oracle for the OIT fragment-store + composite path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product (vk_order_independent_transparency_amd/) never links or calls it.
 *
 * It restates, single-threaded and in draw (fragment-id) order, the
 * reference's GLSL shaders:
 *   shaders/oitSimple.frag.glsl, oitLinkedList.frag.glsl, oitLoop64.frag.glsl,
 *   oitSpinlock.frag.glsl (critical section as in oitInterlock.frag.glsl:
 *   103-148), oitWeighted.frag.glsl, with the helpers of
 *   oitCompositeDefines.glsl and shaderCommon.glsl, the clear values of
 *   oitRender.cpp:156-425 and the framebuffer blend equations of
 *   main.cpp:540-592 / oit.h:52-60.
 * Each function in oit_oracle.c cites the lines it follows.
 *
 * Fixed semantics where the reference leaves them to hardware (each is
 * checked by tests/test_oracle_*.py against closed forms, README worked
 * examples, brute force or high-precision arithmetic):
 *  - depth order: order-preserving u32 key of the float bits, ties broken by
 *    fragment id (the reference compares floats and ties by arrival);
 *  - sRGB8 codec: encode is correctly rounded (threshold table from the
 *    IEC 61966-2-1 curve), decode a correctly rounded 256-entry table,
 *    alpha q = floor(clamp(a,0,1)*255 + 0.5) in binary32;
 *  - tail order: colour-pass tail fragments blend over the background in
 *    ascending fragment id ("latest on top", the ROP's primitive order);
 *    for Loop64/Spinlock an evicted fragment blends at its own id with its
 *    packed sRGB8 colour (ORC_TAIL_CANONICAL), or at the evicting fragment's
 *    position as a sequential reading of the shader gives
 *    (ORC_TAIL_SEQUENTIAL, for reporting only, where a Spinlock fragment
 *    that is never stored blends its full-precision colour as
 *    oitSpinlock.frag.glsl:126-128 does; whether a fragment is rejected or
 *    stored and evicted later depends on the arrival order, so the canonical
 *    order uses the packed colour for both);
 *  - colour buffer: binary32 RGBA, no sRGB8 round trip between blends.
 */
#ifndef OIT_ORACLE_H
#define OIT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SIMPLE = 0, ORC_LINKEDLIST = 1, ORC_LOOP = 2, ORC_LOOP64 = 3, ORC_SPINLOCK = 4, ORC_INTERLOCK = 5,
       ORC_WEIGHTED = 6 };
enum { ORC_TAIL_CANONICAL = 0, ORC_TAIL_SEQUENTIAL = 1 };

typedef struct orc_config {
  uint32_t algorithm;
  uint32_t layers;       /* K */
  uint32_t tail_blend;
  uint32_t width, height, samples;
  uint32_t tail_order;   /* ORC_TAIL_* */
  uint64_t ll_capacity;  /* LL elements including terminator 0 */
  float background[4];
  float wboit_view_near, wboit_view_span;
  uint32_t coverage_samples; /* OIT_COVERAGE_SHADING: MSAA samples per pixel (2, 4, 8), 0 or 1 = off */
  /* Opaque pass of the frame (orc_opaque_pass), NULL when the frame has none:
   * depth key and colour per sample target; with coverage shading sample s
   * of pixel p is entry s * W * H + p. Transparent fragments are depth
   * tested against it (LESS, samples of the mask) and the composite blends
   * over its colour instead of the background. */
  const uint32_t* opaque_key;
  const float*    opaque_rgba;
} orc_config;

typedef struct orc_result {
  float*    out;          /* [T*4] final premultiplied RGBA */
  float*    composite;    /* [T*4] composite-pass colour C (optional) */
  float*    colorbuf;     /* [T*4] colour buffer after the colour pass (optional) */
  uint32_t* count;        /* [T] algorithm count (optional) */
  uint32_t* frontk_key;   /* [T*K] sorted front-K depth keys (optional) */
  uint32_t* frontk_val;   /* [T*K] ids, or packed colours for Loop64 (optional) */
  uint8_t*  tail_flag;    /* [F] 1 if the fragment was tail-blended in the colour pass (optional) */
  float*    wb_accum;     /* [T*4] WBOIT accumulation (optional) */
  float*    wb_reveal;    /* [T] WBOIT revealage (optional) */
  uint64_t  ll_overflow;
  uint64_t  tail_fragments;
} orc_result;

/* codec and keys */
void     orc_tables(float enc_threshold[256], float dec_lut[256], float alpha_lut[256]);
uint32_t orc_depth_key(float d);
uint32_t orc_encode(const float premult[4]);
void     orc_decode_premult(uint32_t packed, float out[4]);
double   orc_srgb_encode_exact(double c);   /* S(c) of the IEC curve as written in shaderCommon.glsl:60-69 */
double   orc_srgb_decode_exact(double f);   /* shaderCommon.glsl:84-94 */

/* Sequential run of one frame. Returns 0 or a negative error. */
int orc_run(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
            const float* rgba, orc_result* res);
/* The same with per-fragment sample masks (gl_SampleMaskIn, required when
 * cfg->coverage_samples > 1; Simple, Linked List and Spinlock only). */
int orc_run_cov(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
                const float* rgba, const uint32_t* mask, orc_result* res);

/* Opaque pass (oitRender.cpp:113-122; BlendMode::NONE pipeline with depth
 * test LESS and depth write, main.cpp:526-533): the planes are cleared to
 * depth 1.0 and the background colour (oitRender.cpp:89-91), then every
 * fragment in draw order writes its depth key and colour to each sample
 * target (its target; with coverage shading the samples of its mask) whose
 * stored key is larger. Outputs [cov * W * H] or [W * H * samples] entries. */
int orc_opaque_pass(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
                    const float* rgba, const uint32_t* mask, uint32_t* key_out, float* rgba_out);

/* Display image: the average of `samples` sample planes and of each
 * ss x ss block of a (width*ss) x (height*ss) premultiplied image (the MSAA
 * resolve, main.cpp:700-714, and the supersample downsample, :715-731),
 * then sRGB8 rgb (exact curve, round to nearest) and linear 8-bit alpha,
 * RGBA byte order. */
int orc_finalize_srgb8(const float* in, uint32_t width, uint32_t height, uint32_t samples, uint32_t ss,
                       uint8_t* out);

/* FP64 mirror: per target, all fragments sorted by (depth key, id) and blended
 * front to back in double (quantize=1 uses the packed sRGB8 colour, 0 the
 * stream colour); no K limit. Writes out[T*4] = result over background.
 * For OIT_WEIGHTED computes the accumulation, revealage and resolve in double. */
int orc_run_fp64(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
                 const float* rgba, int quantize, double* out, double* wb_accum, double* wb_reveal);

/* Composite of given A-buffer contents (the replay oracle): for each target,
 * n = min(count, K) elements (colour, key, id) at [layer][target], sorted
 * and blended (oitSimple.frag.glsl:115-169), then over the tail fold of the
 * fragments whose tail_flag is set, in id order, over the background. */
int orc_replay_kbuffer(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* rgba,
                       const uint32_t* count, const uint32_t* ab_color, const uint32_t* ab_key,
                       const uint32_t* ab_id, const uint8_t* tail_flag, int tail_quantized, float* out);
/* With coverage masks (fragment masks and the A-buffer's mask plane). */
int orc_replay_kbuffer_cov(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* rgba,
                           const uint32_t* mask, const uint32_t* count, const uint32_t* ab_color,
                           const uint32_t* ab_key, const uint32_t* ab_id, const uint32_t* ab_mask,
                           const uint8_t* tail_flag, int tail_quantized, float* out);

#ifdef __cplusplus
}
#endif
#endif
