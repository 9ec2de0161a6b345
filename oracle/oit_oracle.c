/*
 * oit_oracle.c - CPU oracle for the OIT fragment-store + composite path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Single-threaded restatement of the reference's shaders, processing
 * fragments one at a time in draw (fragment id) order. Build flags:
 * -O2 -ffp-contract=off (every float operation rounds once, as in the
 * HIP kernels, which are built the same way).
 */
#include "oit_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EMPTY_U32 0xFFFFFFFFu
#define EMPTY_U64 0xFFFFFFFFFFFFFFFFull

/* ------------------------------------------------------------------ */
/* sRGB8 codec                                                          */
/* ------------------------------------------------------------------ */

static float g_thr[256];   /* g_thr[q] = smallest float c with Q(c) >= q (q >= 1) */
static float g_dec[256];   /* linear value of sRGB code u */
static float g_alpha[256]; /* u / 255 */
static int   g_init = 0;

/* unPremultLinearToSRGB, shaderCommon.glsl:60-69 (evaluated in double). */
double orc_srgb_encode_exact(double c) {
  if (c < (double)0.0031308f) return c * 12.92;
  return pow(c, 1.0 / 2.4) * 1.055 - 0.055;
}

/* unPremultSRGBToLinear, shaderCommon.glsl:84-94 (evaluated in double). */
double orc_srgb_decode_exact(double f) {
  if (f < (double)0.04045f) return f / 12.92;
  return pow((f + 0.055) / 1.055, 2.4);
}

/* packUnorm4x8's per-channel code for an sRGB-encoded value:
 * round(clamp(S(c), 0, 1) * 255) with halves rounded up. */
static int q_exact(float c) {
  double s = orc_srgb_encode_exact((double)c);
  if (!(s > 0.0)) s = 0.0;
  if (s > 1.0) s = 1.0;
  return (int)floor(s * 255.0 + 0.5);
}

static float bits_to_float(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t float_to_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

static void init_tables(void) {
  if (g_init) return;
  g_thr[0] = -INFINITY;
  for (int q = 1; q < 256; ++q) {
    /* binary search over the ordered bit patterns of [0, 1] */
    uint32_t lo = 0u, hi = 0x3F800000u; /* q_exact(lo) < q <= q_exact(hi) */
    while (hi - lo > 1u) {
      uint32_t mid = lo + (hi - lo) / 2u;
      if (q_exact(bits_to_float(mid)) >= q) hi = mid; else lo = mid;
    }
    g_thr[q] = bits_to_float(hi);
  }
  for (int u = 0; u < 256; ++u) {
    g_dec[u] = (float)orc_srgb_decode_exact((double)u / 255.0);
    g_alpha[u] = (float)u / 255.0f;
  }
  g_init = 1;
}

void orc_tables(float enc_threshold[256], float dec_lut[256], float alpha_lut[256]) {
  init_tables();
  memcpy(enc_threshold, g_thr, sizeof g_thr);
  memcpy(dec_lut, g_dec, sizeof g_dec);
  memcpy(alpha_lut, g_alpha, sizeof g_alpha);
}

/* Number of thresholds <= c: the sRGB8 code of unpremultiplied linear c. */
static uint32_t code_rgb(float c) {
  if (!(c >= g_thr[1])) return 0u; /* also NaN */
  int a = 1, b = 255;
  while (a < b) {
    int m = (a + b + 1) / 2;
    if (c >= g_thr[m]) a = m; else b = m - 1;
  }
  return (uint32_t)a;
}

/* shading() output -> unPremultLinearToSRGB -> packUnorm4x8
 * (oitLinkedList.frag.glsl:75-77, oitSimple.frag.glsl:55,70,
 * oitLoop64.frag.glsl:70,79, oitSpinlock.frag.glsl:54,60). The stream holds
 * premultiplied colour, so rgb is first divided by alpha. */
uint32_t orc_encode(const float p[4]) {
  init_tables();
  float a = p[3];
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if (a > 0.0f) { r = p[0] / a; g = p[1] / a; b = p[2] / a; }
  float ac = fminf(fmaxf(a, 0.0f), 1.0f);
  uint32_t qa = (uint32_t)floorf(ac * 255.0f + 0.5f);
  return code_rgb(r) | (code_rgb(g) << 8) | (code_rgb(b) << 16) | (qa << 24);
}

/* unpackUnorm4x8 -> unPremultSRGBToLinear -> premultiply, the first half of
 * doBlendPacked (shaderCommon.glsl:117-124) and of the tail colours at
 * oitLoop64.frag.glsl:134-136 / oitSpinlock.frag.glsl:112,127. */
void orc_decode_premult(uint32_t c, float o[4]) {
  init_tables();
  float a = g_alpha[c >> 24];
  o[0] = g_dec[c & 255u] * a;
  o[1] = g_dec[(c >> 8) & 255u] * a;
  o[2] = g_dec[(c >> 16) & 255u] * a;
  o[3] = a;
}

/* Order-preserving key of a float depth (replaces floatBitsToUint, which
 * orders only non-negative depths: oitLoop64.frag.glsl:29-31). 0xFFFFFFFF
 * stays free as the empty sentinel (oitRender.cpp:307). */
uint32_t orc_depth_key(float d) {
  uint32_t b = float_to_bits(d);
  uint32_t k = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return k == EMPTY_U32 ? 0xFFFFFFFEu : k;
}

/* ------------------------------------------------------------------ */
/* Blending                                                             */
/* ------------------------------------------------------------------ */

/* doBlend (shaderCommon.glsl:108-112): color = color over base. The same
 * formula is the PREMULTIPLIED framebuffer blend (main.cpp:548-558,
 * oit.h:55) with src on top: dst = src + (1 - src.a) * dst. */
static void over_into(float top[4], const float base[4]) {
  float t = 1.0f - top[3];
  top[0] = top[0] + t * base[0];
  top[1] = top[1] + t * base[1];
  top[2] = top[2] + t * base[2];
  top[3] = top[3] + t * base[3];
}

static void rop_blend(float dst[4], const float src[4]) {
  float t = 1.0f - src[3];
  dst[0] = src[0] + t * dst[0];
  dst[1] = src[1] + t * dst[1];
  dst[2] = src[2] + t * dst[2];
  dst[3] = src[3] + t * dst[3];
}

/* doBlendPacked (shaderCommon.glsl:117-124). */
static void blend_packed(float color[4], uint32_t packed) {
  float f[4];
  orc_decode_premult(packed, f);
  over_into(color, f);
}

/* ------------------------------------------------------------------ */
/* Element helpers: (key, id) total order                              */
/* ------------------------------------------------------------------ */

typedef struct elem { uint32_t key, id, col, mask; } elem;

static int elem_lt(elem a, elem b) { return a.key < b.key || (a.key == b.key && a.id < b.id); }

/* Front-to-back blend of the n sorted elements (oitLinkedList.frag.glsl:
 * 162-166); with coverage shading, per sample over the elements whose mask
 * holds it, summed in sample order and divided by the sample count
 * (:147-160, oitSimple.frag.glsl:142-159). */
static void blend_front(float c[4], const elem* a, int n, uint32_t cov) {
  if (cov <= 1) {
    for (int i = 0; i < n; ++i) blend_packed(c, a[i].col);
    return;
  }
  float sum[4] = {0, 0, 0, 0};
  for (uint32_t s = 0; s < cov; ++s) {
    float sc[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i)
      if ((a[i].mask >> s) & 1u) blend_packed(sc, a[i].col);
    for (int k = 0; k < 4; ++k) sum[k] = sum[k] + sc[k];
  }
  for (int k = 0; k < 4; ++k) c[k] = sum[k] / (float)cov;
}

/* bubbleSort, oitCompositeDefines.glsl:51-68 (swap on >=). */
static void bubble_sort(elem* a, int n) {
  for (int i = n - 2; i >= 0; --i)
    for (int j = 0; j <= i; ++j)
      if (!elem_lt(a[j], a[j + 1])) { elem t = a[j + 1]; a[j + 1] = a[j]; a[j] = t; }
}

/* insertionSort, oitCompositeDefines.glsl:94-109. */
static void insertion_sort(elem* a, int K, elem x) {
  for (int i = 0; i < K; ++i)
    if (elem_lt(x, a[i])) {
      for (int j = K - 1; j > i; --j) a[j] = a[j - 1];
      a[i] = x;
      return;
    }
}

/* insertionSortTail, oitCompositeDefines.glsl:116-139. */
static elem insertion_sort_tail(elem* a, int K, elem x) {
  elem last = x;
  if (elem_lt(x, a[K - 1]))
    for (int i = 0; i < K; ++i)
      if (elem_lt(x, a[i])) {
        last = a[K - 1];
        for (int j = K - 1; j > i; --j) a[j] = a[j - 1];
        a[i] = x;
        break;
      }
  return last;
}

/* ------------------------------------------------------------------ */
/* Colour-pass tail events                                              */
/* ------------------------------------------------------------------ */

typedef struct tail_ev {
  uint32_t pos;     /* blend position: id of the fragment the colour belongs to (canonical) or of the evictor (sequential) */
  uint32_t target;
  uint32_t mask;    /* samples the ROP writes (coverage shading); all ones otherwise */
  float    c[4];    /* premultiplied colour blended by the ROP */
} tail_ev;

typedef struct tail_list { tail_ev* ev; uint64_t n, cap; } tail_list;

static int tail_push_m(tail_list* L, uint32_t pos, uint32_t target, uint32_t mask, const float c[4]);
static int tail_push(tail_list* L, uint32_t pos, uint32_t target, const float c[4]) {
  return tail_push_m(L, pos, target, 0xFFFFFFFFu, c);
}
static int tail_push_m(tail_list* L, uint32_t pos, uint32_t target, uint32_t mask, const float c[4]) {
  if (L->n == L->cap) {
    uint64_t nc = L->cap ? L->cap * 2 : 1024;
    tail_ev* p = (tail_ev*)realloc(L->ev, nc * sizeof(tail_ev));
    if (!p) return -2;
    L->ev = p; L->cap = nc;
  }
  tail_ev* e = &L->ev[L->n++];
  e->pos = pos; e->target = target; e->mask = mask; memcpy(e->c, c, sizeof e->c);
  return 0;
}

static int ev_cmp(const void* a, const void* b) {
  const tail_ev* x = (const tail_ev*)a; const tail_ev* y = (const tail_ev*)b;
  return (x->pos > y->pos) - (x->pos < y->pos);
}

/* Blends the tail events over the colour buffer in ascending position:
 * the ROP applies PREMULTIPLIED in primitive order (main.cpp:548-558).
 * Sequential order: one ROP blend per event, oldest first. Canonical order
 * (fragment ids): the same composite (over is associative), associated
 * newest first: acc = e_n over e_(n-1) over ... over e_1, then acc over the
 * background. The reference blends into an 8-bit target, so the float
 * association is not fixed by it; this one lets the kernels fold a
 * newest-first list in one walk. */
static int fold_tail_cov(tail_list* L, float* colorbuf, uint64_t T, int canonical, uint32_t cov,
                         const float* sbase);
static int fold_tail(tail_list* L, float* colorbuf, uint64_t T, int canonical) {
  return fold_tail_cov(L, colorbuf, T, canonical, 0, NULL);
}
/* Coverage shading: the ROP blends each event into the samples of its mask;
 * per sample the same fold, then the colour buffer of a target with tail
 * events is the sum of its sample values in sample order divided by the
 * sample count (the MSAA resolve). `sbase` (an opaque pass): the sample
 * colours under the fold, entry s * T + t; every target is then resolved
 * from its samples. Without it each sample starts at colorbuf[t] and a
 * target without events keeps colorbuf[t]. */
static int fold_tail_cov(tail_list* L, float* colorbuf, uint64_t T, int canonical, uint32_t cov,
                         const float* sbase) {
  qsort(L->ev, L->n, sizeof(tail_ev), ev_cmp); /* positions are unique per target in canonical order */
  if (cov > 1) {
    float* acc = (float*)calloc(T * cov * 4, sizeof(float));
    unsigned char* seen = (unsigned char*)calloc(T, 1);
    if (!acc || !seen) { free(acc); free(seen); return -2; }
    if (sbase) memset(seen, 1, T);
    if (canonical) {
      for (uint64_t i = L->n; i-- > 0;) {
        const uint64_t t = L->ev[i].target;
        for (uint32_t s = 0; s < cov; ++s)
          if ((L->ev[i].mask >> s) & 1u) over_into(&acc[(t * cov + s) * 4], L->ev[i].c);
        seen[t] = 1;
      }
      for (uint64_t t = 0; t < T; ++t) {
        if (!seen[t]) continue;
        for (uint32_t s = 0; s < cov; ++s)
          over_into(&acc[(t * cov + s) * 4], sbase ? &sbase[(s * T + t) * 4] : &colorbuf[4 * t]);
      }
    } else {
      for (uint64_t t = 0; t < T; ++t)
        for (uint32_t s = 0; s < cov; ++s)
          memcpy(&acc[(t * cov + s) * 4], sbase ? &sbase[(s * T + t) * 4] : &colorbuf[4 * t], 4 * sizeof(float));
      for (uint64_t i = 0; i < L->n; ++i) {
        const uint64_t t = L->ev[i].target;
        for (uint32_t s = 0; s < cov; ++s)
          if ((L->ev[i].mask >> s) & 1u) rop_blend(&acc[(t * cov + s) * 4], L->ev[i].c);
        seen[t] = 1;
      }
    }
    for (uint64_t t = 0; t < T; ++t) {
      if (!seen[t]) continue;
      float sum[4] = {0, 0, 0, 0};
      for (uint32_t s = 0; s < cov; ++s)
        for (int k = 0; k < 4; ++k) sum[k] = sum[k] + acc[(t * cov + s) * 4 + k];
      for (int k = 0; k < 4; ++k) colorbuf[4 * t + k] = sum[k] / (float)cov;
    }
    free(acc);
    free(seen);
    return 0;
  }
  if (!canonical) {
    for (uint64_t i = 0; i < L->n; ++i) rop_blend(&colorbuf[4 * (uint64_t)L->ev[i].target], L->ev[i].c);
    return 0;
  }
  float* acc = (float*)calloc(T * 4, sizeof(float));
  unsigned char* seen = (unsigned char*)calloc(T, 1);
  if (!acc || !seen) { free(acc); free(seen); return -2; }
  for (uint64_t i = L->n; i-- > 0;) {
    const uint64_t t = L->ev[i].target;
    over_into(&acc[4 * t], L->ev[i].c);
    seen[t] = 1;
  }
  for (uint64_t t = 0; t < T; ++t) {
    if (!seen[t]) continue;
    float o[4];
    memcpy(o, &acc[4 * t], sizeof o);
    over_into(o, &colorbuf[4 * t]);
    memcpy(&colorbuf[4 * t], o, sizeof o);
  }
  free(acc);
  free(seen);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Frame                                                                */
/* ------------------------------------------------------------------ */

static int valid_layers(uint32_t k) { return k == 1 || k == 2 || k == 4 || k == 8 || k == 16 || k == 32 || k == 64; }

static void write_frontk(orc_result* res, uint64_t t, uint32_t K, const elem* a, int n, int loop64) {
  if (res->frontk_key)
    for (uint32_t i = 0; i < K; ++i) res->frontk_key[t * K + i] = (int)i < n ? a[i].key : EMPTY_U32;
  if (res->frontk_val)
    for (uint32_t i = 0; i < K; ++i) res->frontk_val[t * K + i] = (int)i < n ? (loop64 ? a[i].col : a[i].id) : EMPTY_U32;
}

/* Early depth test of a transparent fragment against the frame's opaque
 * pass (early_fragment_tests with post_depth_coverage,
 * oitColorDepthDefines.glsl:24-37; VK_COMPARE_OP_LESS, main.cpp:526-528,
 * no depth write for the transparent pipelines, :541-543): the samples of m
 * whose opaque depth key is larger than key. 0: the fragment does not exist
 * (its shader does not run). */
static uint32_t opq_visible(const orc_config* cfg, uint64_t T, uint32_t cov, uint32_t t, uint32_t key, uint32_t m) {
  if (!cfg->opaque_key) return m;
  if (cov) {
    uint32_t r = 0;
    for (uint32_t s = 0; s < cov; ++s)
      if (((m >> s) & 1u) && key < cfg->opaque_key[(uint64_t)s * T + t]) r |= 1u << s;
    return r;
  }
  return key < cfg->opaque_key[t] ? m : 0u;
}

int orc_opaque_pass(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
                    const float* rgba, const uint32_t* mask, uint32_t* key_out, float* rgba_out) {
  init_tables();
  if (!cfg || !key_out || !rgba_out) return -1;
  const uint64_t T = (uint64_t)cfg->width * cfg->height * cfg->samples;
  const uint32_t cov = cfg->coverage_samples > 1 ? cfg->coverage_samples : 0;
  if (T == 0 || (cov && (cfg->samples != 1 || !mask))) return -1;
  if (cov && cov != 2 && cov != 4 && cov != 8) return -1;
  const uint32_t planes = cov ? cov : 1;
  const uint32_t clear = orc_depth_key(1.0f); /* depth clear, oitRender.cpp:91 */
  for (uint64_t j = 0; j < T * planes; ++j) {
    key_out[j] = clear;
    memcpy(&rgba_out[4 * j], cfg->background, 4 * sizeof(float)); /* colour clear, :90 */
  }
  for (uint64_t i = 0; i < F; ++i) if (target[i] >= T) return -1;
  for (uint64_t i = 0; i < F; ++i) {
    const uint32_t key = orc_depth_key(depth[i]);
    const uint32_t m = cov ? mask[i] & ((1u << cov) - 1u) : 1u;
    for (uint32_t s = 0; s < planes; ++s) {
      if (!((m >> s) & 1u)) continue;
      const uint64_t j = (uint64_t)s * T + target[i];
      if (key < key_out[j]) { /* LESS; a later fragment at an equal depth fails */
        key_out[j] = key;
        memcpy(&rgba_out[4 * j], &rgba[4 * i], 4 * sizeof(float));
      }
    }
  }
  return 0;
}

int orc_run(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
            const float* rgba, orc_result* res) {
  return orc_run_cov(cfg, F, target, depth, rgba, NULL, res);
}

int orc_run_cov(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
                const float* rgba, const uint32_t* mask, orc_result* res) {
  init_tables();
  const uint32_t K = cfg->layers;
  const uint64_t T = (uint64_t)cfg->width * cfg->height * cfg->samples;
  if (T == 0 || !res || !res->out) return -1;
  const uint32_t cov = cfg->coverage_samples > 1 ? cfg->coverage_samples : 0;
  if (cov) {
    if (cov != 2 && cov != 4 && cov != 8) return -1;
    if (cfg->samples != 1 || !mask) return -1;
    if (cfg->algorithm != ORC_SIMPLE && cfg->algorithm != ORC_LINKEDLIST && cfg->algorithm != ORC_SPINLOCK &&
        cfg->algorithm != ORC_INTERLOCK)
      return -1; /* Loop and Loop64 have no coverage shading (README.md:33-34) */
  }
  const uint32_t full = cov ? (1u << cov) - 1u : 0xFFFFFFFFu;
  if (cfg->algorithm != ORC_WEIGHTED && !valid_layers(K)) return -1;
  for (uint64_t i = 0; i < F; ++i) if (target[i] >= T) return -1;
  if (F >= 0xFFFFFFFFull) return -1;
  if (!cfg->opaque_key != !cfg->opaque_rgba) return -1;

  float* colorbuf = (float*)malloc(T * 4 * sizeof(float));
  if (!colorbuf) return -2;
  for (uint64_t t = 0; t < T; ++t) memcpy(&colorbuf[4 * t], cfg->background, 4 * sizeof(float));
  /* with an opaque pass the colour buffer starts as its colour (per sample with coverage shading) */
  const float* sbase = cov ? cfg->opaque_rgba : NULL;
  if (cfg->opaque_rgba && !cov) memcpy(colorbuf, cfg->opaque_rgba, T * 4 * sizeof(float));
  tail_list tails = {0, 0, 0};
  res->ll_overflow = 0;
  res->tail_fragments = 0;
  if (res->tail_flag) memset(res->tail_flag, 0, F);
  int rc = 0;
  /* the ordered interlock blends its tail in primitive order by construction
   * (oitInterlock.frag.glsl:52-79 ordered critical section + ROP order) */
  const int canonical = cfg->tail_order == ORC_TAIL_CANONICAL && cfg->algorithm != ORC_INTERLOCK;

  if (cfg->algorithm == ORC_LINKEDLIST) {
    /* ---- oitLinkedList.frag.glsl ---- */
    const uint64_t cap = cfg->ll_capacity;
    if (cap < 1) { rc = -1; goto done; }
    uint32_t* head = (uint32_t*)calloc(T, sizeof(uint32_t)); /* imgAux cleared to 0, oitRender.cpp:206-210 */
    elem*     node = (elem*)malloc((cap < F + 1 ? cap : F + 1) * sizeof(elem));
    uint32_t* next = (uint32_t*)malloc((cap < F + 1 ? cap : F + 1) * sizeof(uint32_t));
    if (!head || !node || !next) { free(head); free(node); free(next); rc = -2; goto done; }
    uint64_t counter = 0; /* imgCounter cleared to 0, oitRender.cpp:211-212 */
    for (uint64_t i = 0; i < F; ++i) {
      const uint64_t n = counter + 1; /* :55 newOffset = atomicAdd(counter, 1) + 1 */
      counter++;
      const uint32_t m = opq_visible(cfg, T, cov, target[i], orc_depth_key(depth[i]), mask ? mask[i] & full : full);
      if (!m) continue; /* no sample covered or all behind the opaque depth: the shader does not run
                           (its node stays unused) */
      if (n >= cap) { /* :59-68 out of memory: tail-blend the shaded colour */
        res->ll_overflow++;
        if (cfg->tail_blend) {
          if ((rc = tail_push_m(&tails, (uint32_t)i, target[i], m, &rgba[4 * i]))) goto ll_free;
          if (res->tail_flag) res->tail_flag[i] = 1;
          res->tail_fragments++;
        }
        continue;
      }
      const uint32_t t = target[i];
      next[n] = head[t]; /* :72 oldOffset = atomicExchange(head, newOffset) */
      head[t] = (uint32_t)n;
      node[n].col = orc_encode(&rgba[4 * i]); /* :75-82 */
      node[n].key = orc_depth_key(depth[i]);
      node[n].id = (uint32_t)i;
      node[n].mask = m; /* :79 storeMask */
    }
    if ((rc = fold_tail_cov(&tails, colorbuf, T, canonical, cov, sbase))) goto ll_free;
    /* composite, :106-172 */
    elem* a = (elem*)malloc(K * sizeof(elem));
    if (!a) { rc = -2; goto ll_free; }
    for (uint64_t t = 0; t < T; ++t) {
      int frags = 0;
      uint32_t p = head[t];
      uint32_t len = 0;
      while (p != 0 && frags < (int)K) { a[frags++] = node[p]; p = next[p]; ++len; } /* :116-123 */
      bubble_sort(a, frags);                                                    /* :126 */
      float tailc[4] = {0, 0, 0, 0};
      while (p != 0) {                                                          /* :131-143 */
        if (cfg->tail_blend) {
          elem ev = insertion_sort_tail(a, (int)K, node[p]);
          blend_packed(tailc, ev.col);
        } else {
          insertion_sort(a, (int)K, node[p]);
        }
        p = next[p];
        ++len;
      }
      float c[4] = {0, 0, 0, 0};
      blend_front(c, a, frags, cov); /* :147-166 */
      over_into(c, tailc);           /* :170 */
      if (res->composite) memcpy(&res->composite[4 * t], c, sizeof c);
      if (res->count) res->count[t] = len;
      write_frontk(res, t, K, a, frags, 0);
      float o[4];
      memcpy(o, &colorbuf[4 * t], sizeof o);
      rop_blend(o, c); /* composite pipeline blends PREMULTIPLIED onto the colour buffer, oit.cpp:723-726 */
      memcpy(&res->out[4 * t], o, sizeof o);
    }
    free(a);
  ll_free:
    free(head); free(node); free(next);
  } else if (cfg->algorithm == ORC_SIMPLE || cfg->algorithm == ORC_SPINLOCK || cfg->algorithm == ORC_INTERLOCK) {
    /* Interlock (oitInterlock.frag.glsl:103-148) runs the Spinlock's critical
     * section under the fragment shader interlock: one invocation per pixel at
     * a time, in primitive order when ordered (oit.h:71) */
    const int spin = cfg->algorithm == ORC_SPINLOCK || cfg->algorithm == ORC_INTERLOCK;
    uint32_t* count = (uint32_t*)calloc(T, sizeof(uint32_t)); /* imgAux cleared to 0 */
    uint32_t* auxd = (uint32_t*)malloc(T * sizeof(uint32_t)); /* imgAuxDepth cleared to ~0, oitRender.cpp:348 */
    elem*     ab = (elem*)malloc(T * K * sizeof(elem));
    if (!count || !auxd || !ab) { free(count); free(auxd); free(ab); rc = -2; goto done; }
    for (uint64_t t = 0; t < T; ++t) auxd[t] = EMPTY_U32;
    for (uint64_t i = 0; i < F; ++i) {
      const uint32_t t = target[i];
      const uint32_t m = opq_visible(cfg, T, cov, t, orc_depth_key(depth[i]), mask ? mask[i] & full : full);
      if (!m) continue; /* no sample covered or all behind the opaque depth: the shader does not run */
      elem x = { orc_depth_key(depth[i]), (uint32_t)i, orc_encode(&rgba[4 * i]), m };
      if (!spin) {
        /* oitSimple.frag.glsl:74-90 */
        uint32_t slot = count[t]++;
        if (slot < K) { ab[(uint64_t)t * K + slot] = x; continue; }
        if (cfg->tail_blend) {
          if ((rc = tail_push_m(&tails, (uint32_t)i, t, m, &rgba[4 * i]))) goto kb_free;
          if (res->tail_flag) res->tail_flag[i] = 1;
          res->tail_fragments++;
        }
        continue;
      }
      /* oitSpinlock.frag.glsl:66-128 */
      int have_tail = 0;
      elem tail_e = x;      /* the element whose colour is tail-blended */
      int  tail_full = 1;   /* colour is the full-precision shaded colour */
      if (x.key <= auxd[t]) { /* :66-69 early depth test */
        uint32_t c = count[t];          /* :86-87 */
        count[t] = c + 1;
        if (c < K) {
          ab[(uint64_t)t * K + c] = x;  /* :89-93 */
        } else {
          int furthest = 0;             /* :96-107 */
          elem maxe = {0u, 0u, 0u, 0u};
          for (uint32_t j = 0; j < K; ++j)
            if (elem_lt(maxe, ab[(uint64_t)t * K + j])) { maxe = ab[(uint64_t)t * K + j]; furthest = (int)j; }
          if (elem_lt(x, maxe)) {       /* :109-117 replace the furthest, tail-blend it */
            tail_e = maxe; tail_full = 0; have_tail = 1;
            ab[(uint64_t)t * K + furthest] = x;
            auxd[t] = maxe.key;
          } else {
            have_tail = 1;              /* not inserted: tail-blend self (:126-128) */
          }
        }
      } else {
        have_tail = 1;
      }
      if (have_tail && cfg->tail_blend) {
        float c4[4];
        /* sequential: a fragment left out blends its own shaded colour (`color` = shading(IN),
         * :51,126-128), an evicted element its stored sRGB8 colour (:112); canonical: the packed
         * colour for both (which of the two a fragment is depends on the arrival order) */
        if (canonical || !tail_full) orc_decode_premult(tail_e.col, c4);
        else memcpy(c4, &rgba[4 * i], sizeof c4);
        uint32_t pos = canonical ? tail_e.id : (uint32_t)i;
        /* canonical: the tail fragment at its own position with its own samples;
         * sequential: the evicting invocation's output, written to its samples */
        if ((rc = tail_push_m(&tails, pos, t, canonical ? tail_e.mask : x.mask, c4))) goto kb_free;
        if (res->tail_flag) res->tail_flag[tail_e.id] = 1;
        res->tail_fragments++;
      }
    }
    if ((rc = fold_tail_cov(&tails, colorbuf, T, canonical, cov, sbase))) goto kb_free;
    {
      elem* a = (elem*)malloc(K * sizeof(elem));
      if (!a) { rc = -2; goto kb_free; }
      for (uint64_t t = 0; t < T; ++t) {
        /* composite, oitSimple.frag.glsl:115-169 == oitSpinlock.frag.glsl:153-207 */
        int n = (int)(count[t] < K ? count[t] : K);
        for (int i = 0; i < n; ++i) a[i] = ab[(uint64_t)t * K + i];
        bubble_sort(a, n);
        float c[4] = {0, 0, 0, 0};
        blend_front(c, a, n, cov); /* oitSimple.frag.glsl:142-165 */
        if (res->composite) memcpy(&res->composite[4 * t], c, sizeof c);
        if (res->count) res->count[t] = count[t];
        write_frontk(res, t, K, a, n, 0);
        float o[4];
        memcpy(o, &colorbuf[4 * t], sizeof o);
        rop_blend(o, c);
        memcpy(&res->out[4 * t], o, sizeof o);
      }
      free(a);
    }
  kb_free:
    free(count); free(auxd); free(ab);
  } else if (cfg->algorithm == ORC_LOOP) {
    /* ---- oitLoop.frag.glsl: depth pass (:57-100), colour pass (:105-180),
     * composite (:186-222); buffers cleared to all ones (depths) ---- */
    uint32_t* zs = (uint32_t*)malloc(T * K * sizeof(uint32_t));
    uint64_t* cs = (uint64_t*)calloc(T * K, sizeof(uint64_t)); /* (id + 1) << 32 | colour of the last writer */
    if (!zs || !cs) { free(zs); free(cs); rc = -2; goto done; }
    for (uint64_t j = 0; j < T * K; ++j) zs[j] = EMPTY_U32;
    for (uint64_t i = 0; i < F; ++i) { /* depth pass: the K nearest distinct depths, sorted */
      const uint64_t base = (uint64_t)target[i] * K;
      uint32_t z = orc_depth_key(depth[i]);
      if (!opq_visible(cfg, T, 0, target[i], z, 1u)) continue;
      uint32_t j = 0;
      if (z > zs[base + K - 1]) continue;    /* :75-80 pretest */
      if (z > zs[base + K / 2]) j = K / 2;   /* :82-87 */
      for (; j < K; ++j) {                   /* :89-99 atomicMin chain */
        const uint32_t old = zs[base + j];
        if (z < old) zs[base + j] = z;
        if (old == EMPTY_U32 || old == z) break;
        z = old > z ? old : z;
      }
    }
    for (uint64_t i = 0; i < F; ++i) { /* colour pass */
      const uint64_t base = (uint64_t)target[i] * K;
      const uint32_t z = orc_depth_key(depth[i]);
      if (!opq_visible(cfg, T, 0, target[i], z, 1u)) continue;
      if (zs[base + K - 1] < z) {            /* :139-150 behind the K-th depth: tail-blend itself */
        if (cfg->tail_blend) {
          if ((rc = tail_push(&tails, (uint32_t)i, target[i], &rgba[4 * i]))) { free(zs); free(cs); goto done; }
          if (res->tail_flag) res->tail_flag[i] = 1;
          res->tail_fragments++;
        }
        continue;
      }
      int start = 0, end = (int)K - 1;       /* :154-175 binary search of its depth */
      while (start < end) {
        const int mid = (start + end) / 2;
        if (zs[base + (uint32_t)mid] < z) start = mid + 1; else end = mid;
      }
      cs[base + (uint32_t)start] = ((uint64_t)(i + 1) << 32) | orc_encode(&rgba[4 * i]); /* :177, later writes win */
    }
    if ((rc = fold_tail(&tails, colorbuf, T, canonical))) { free(zs); free(cs); goto done; }
    elem* a = (elem*)malloc(K * sizeof(elem));
    if (!a) { free(zs); free(cs); rc = -2; goto done; }
    for (uint64_t t = 0; t < T; ++t) {
      float c[4] = {0, 0, 0, 0};
      int n = 0;
      for (uint32_t j = 0; j < K; ++j) {     /* :198-211 */
        if (zs[t * K + j] == EMPTY_U32) break;
        ++n;
      }
      for (int j = 0; j < n; ++j) {          /* :213-217 */
        const uint64_t s = cs[t * K + (uint32_t)j];
        blend_packed(c, (uint32_t)(s & 0xFFFFFFFFu));
        a[j].key = zs[t * K + (uint32_t)j]; a[j].col = (uint32_t)(s & 0xFFFFFFFFu);
        a[j].id = (uint32_t)(s >> 32) - 1u; a[j].mask = 0;
      }
      if (res->composite) memcpy(&res->composite[4 * t], c, sizeof c);
      if (res->count) res->count[t] = (uint32_t)n;
      write_frontk(res, t, K, a, n, 1);
      float o[4];
      memcpy(o, &colorbuf[4 * t], sizeof o);
      rop_blend(o, c);
      memcpy(&res->out[4 * t], o, sizeof o);
    }
    free(a);
    free(zs); free(cs);
  } else if (cfg->algorithm == ORC_LOOP64) {
    /* ---- oitLoop64.frag.glsl ---- */
    uint64_t* slot = (uint64_t*)malloc(T * K * sizeof(uint64_t)); /* cleared to all ones, oitRender.cpp:307 */
    uint32_t* sid = (uint32_t*)malloc(T * K * sizeof(uint32_t));  /* shadow ids: which fragment a key came from */
    if (!slot || !sid) { free(slot); free(sid); rc = -2; goto done; }
    for (uint64_t j = 0; j < T * K; ++j) { slot[j] = EMPTY_U64; sid[j] = EMPTY_U32; }
    for (uint64_t i = 0; i < F; ++i) {
      const uint64_t base = (uint64_t)target[i] * K;
      if (!opq_visible(cfg, T, 0, target[i], orc_depth_key(depth[i]), 1u)) continue;
      uint64_t zcur = ((uint64_t)orc_depth_key(depth[i]) << 32) | orc_encode(&rgba[4 * i]); /* :79 */
      uint32_t zid = (uint32_t)i;
      uint32_t j = 0;
      int can_insert = 1;
      if (zcur > slot[base + K - 1]) can_insert = 0;     /* :86-90 USE_EARLYDEPTH pretests */
      else if (zcur > slot[base + K / 2]) j = K / 2;      /* :95-99 */
      int evict = 1;
      if (can_insert) {
        for (; j < K; ++j) {                               /* :109-121 */
          const uint64_t old = slot[base + j];
          const uint32_t oldid = sid[base + j];
          /* atomicMin; equal keys are the same bits, the shadow keeps the lower id in the slot */
          if (zcur < old || (zcur == old && zid < oldid)) { slot[base + j] = zcur; sid[base + j] = zid; }
          if (old == EMPTY_U64) { evict = 0; break; }
          if (old > zcur || (old == zcur && oldid > zid)) { zcur = old; zid = oldid; }
        }
      }
      if (evict && cfg->tail_blend) {                      /* :129-137 */
        float c4[4];
        orc_decode_premult((uint32_t)(zcur & 0xFFFFFFFFu), c4);
        uint32_t pos = canonical ? zid : (uint32_t)i;
        if ((rc = tail_push(&tails, pos, target[i], c4))) { free(slot); free(sid); goto done; }
        if (res->tail_flag) res->tail_flag[zid] = 1;
        res->tail_fragments++;
      }
    }
    if ((rc = fold_tail(&tails, colorbuf, T, canonical))) { free(slot); free(sid); goto done; }
    elem* a = (elem*)malloc(K * sizeof(elem));
    if (!a) { free(slot); free(sid); rc = -2; goto done; }
    for (uint64_t t = 0; t < T; ++t) {
      float c[4] = {0, 0, 0, 0};                           /* composite :162-183 */
      int n = 0;
      for (uint32_t j = 0; j < K; ++j) {
        const uint64_t s = slot[t * K + j];
        if ((uint32_t)(s >> 32) == EMPTY_U32) break;
        blend_packed(c, (uint32_t)(s & 0xFFFFFFFFu));
        a[n].key = (uint32_t)(s >> 32); a[n].col = (uint32_t)(s & 0xFFFFFFFFu); a[n].id = sid[t * K + j];
        ++n;
      }
      if (res->composite) memcpy(&res->composite[4 * t], c, sizeof c);
      if (res->count) res->count[t] = (uint32_t)n;
      write_frontk(res, t, K, a, n, 1);
      float o[4];
      memcpy(o, &colorbuf[4 * t], sizeof o);
      rop_blend(o, c);
      memcpy(&res->out[4 * t], o, sizeof o);
    }
    free(a);
    free(slot); free(sid);
  } else if (cfg->algorithm == ORC_WEIGHTED) {
    /* ---- oitWeighted.frag.glsl; clears oitRender.cpp:394-397 ---- */
    float* acc = (float*)calloc(T * 4, sizeof(float));
    float* rev = (float*)malloc(T * sizeof(float));
    if (!acc || !rev) { free(acc); free(rev); rc = -2; goto done; }
    for (uint64_t t = 0; t < T; ++t) rev[t] = 1.0f;
    for (uint64_t i = 0; i < F; ++i) {
      if (!opq_visible(cfg, T, 0, target[i], orc_depth_key(depth[i]), 1u)) continue;
      const float* c = &rgba[4 * i];   /* premultiplied, :55-56 */
      const float view = cfg->wboit_view_near + depth[i] * cfg->wboit_view_span;
      const float depthZ = view * 10.0f;                                   /* :65 */
      const float x = depthZ / 200.0f;
      const float x2 = x * x;
      const float x4 = x2 * x2;                                            /* pow(depthZ/200, 4.0) */
      float dw = 0.03f / (1e-5f + x4);                                     /* :67 */
      dw = fminf(fmaxf(dw, 1e-2f), 3e3f);
      float aw = fminf(1.0f, fmaxf(fmaxf(c[0], c[1]), fmaxf(c[2], c[3])) * 40.0f + 0.01f); /* :69 */
      aw = aw * aw;                                                        /* :70 */
      const float w = aw * dw;                                             /* :72 */
      float* A = &acc[4 * (uint64_t)target[i]];
      for (int k = 0; k < 4; ++k) A[k] = A[k] + c[k] * w;                  /* :75, blend ONE,ONE main.cpp:563-569 */
      rev[target[i]] = rev[target[i]] * (1.0f - c[3]);                     /* :78, blend ZERO,1-SRC main.cpp:570-576 */
    }
    for (uint64_t t = 0; t < T; ++t) {
      const float* A = &acc[4 * t];
      const float r = rev[t];
      const float den = fmaxf(A[3], 1e-5f);                                /* :108 */
      float s[4] = { A[0] / den, A[1] / den, A[2] / den, r };
      const float om = 1.0f - r;                                           /* WEIGHTED_COMPOSITE main.cpp:578-588 */
      const float* dst = &colorbuf[4 * t];
      float o[4];
      for (int k = 0; k < 4; ++k) o[k] = s[k] * om + dst[k] * r;
      memcpy(&res->out[4 * t], o, sizeof o);
      if (res->wb_accum) memcpy(&res->wb_accum[4 * t], A, 4 * sizeof(float));
      if (res->wb_reveal) res->wb_reveal[t] = r;
      if (res->composite) memcpy(&res->composite[4 * t], s, sizeof s);
    }
    free(acc); free(rev);
  } else {
    rc = -5;
  }

done:
  if (rc == 0 && res->colorbuf) memcpy(res->colorbuf, colorbuf, T * 4 * sizeof(float));
  free(colorbuf);
  free(tails.ev);
  return rc;
}

/* ------------------------------------------------------------------ */
/* FP64 mirror                                                          */
/* ------------------------------------------------------------------ */

typedef struct sort_rec { uint32_t key, id; } sort_rec;

static int rec_cmp(const void* a, const void* b) {
  const sort_rec* x = (const sort_rec*)a; const sort_rec* y = (const sort_rec*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return (x->id > y->id) - (x->id < y->id);
}

int orc_run_fp64(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* depth,
                 const float* rgba, int quantize, double* out, double* wb_accum, double* wb_reveal) {
  init_tables();
  if (cfg->coverage_samples > 1) return -1; /* the exact mirror has no coverage shading */
  const uint64_t T = (uint64_t)cfg->width * cfg->height * cfg->samples;
  if (T == 0 || !out) return -1;
  if (!cfg->opaque_key != !cfg->opaque_rgba) return -1;
  for (uint64_t i = 0; i < F; ++i) if (target[i] >= T) return -1;
#define ORC_DST(t, k) ((double)(cfg->opaque_rgba ? cfg->opaque_rgba[4 * (t) + (k)] : cfg->background[k]))
  if (cfg->algorithm == ORC_WEIGHTED) {
    double* acc = (double*)calloc(T * 4, sizeof(double));
    double* rev = (double*)malloc(T * sizeof(double));
    if (!acc || !rev) { free(acc); free(rev); return -2; }
    for (uint64_t t = 0; t < T; ++t) rev[t] = 1.0;
    for (uint64_t i = 0; i < F; ++i) {
      if (!opq_visible(cfg, T, 0, target[i], orc_depth_key(depth[i]), 1u)) continue;
      const float* c = &rgba[4 * i];
      const double view = (double)cfg->wboit_view_near + (double)depth[i] * (double)cfg->wboit_view_span;
      const double x = view * 10.0 / 200.0;
      double dw = 0.03 / (1e-5 + x * x * x * x);
      dw = dw < 1e-2 ? 1e-2 : (dw > 3e3 ? 3e3 : dw);
      double m = fmax(fmax(c[0], c[1]), fmax(c[2], c[3]));
      double aw = m * 40.0 + 0.01;
      if (aw > 1.0) aw = 1.0;
      aw *= aw;
      const double w = aw * dw;
      for (int k = 0; k < 4; ++k) acc[4 * (uint64_t)target[i] + k] += (double)c[k] * w;
      rev[target[i]] *= 1.0 - (double)c[3];
    }
    for (uint64_t t = 0; t < T; ++t) {
      const double r = rev[t];
      const double den = acc[4 * t + 3] > 1e-5 ? acc[4 * t + 3] : 1e-5;
      for (int k = 0; k < 3; ++k) out[4 * t + k] = acc[4 * t + k] / den * (1.0 - r) + ORC_DST(t, k) * r;
      out[4 * t + 3] = r * (1.0 - r) + ORC_DST(t, 3) * r;
      if (wb_accum) for (int k = 0; k < 4; ++k) wb_accum[4 * t + k] = acc[4 * t + k];
      if (wb_reveal) wb_reveal[t] = r;
    }
    free(acc); free(rev);
    return 0;
  }
  /* group by target with a counting sort (stable in id), then sort each group by (key, id) */
  uint64_t* start = (uint64_t*)calloc(T + 1, sizeof(uint64_t));
  sort_rec* recs = (sort_rec*)malloc((F ? F : 1) * sizeof(sort_rec));
  if (!start || !recs) { free(start); free(recs); return -2; }
  unsigned char* live = (unsigned char*)malloc(F ? F : 1);
  if (!live) { free(start); free(recs); return -2; }
  for (uint64_t i = 0; i < F; ++i) live[i] = opq_visible(cfg, T, 0, target[i], orc_depth_key(depth[i]), 1u) != 0;
  for (uint64_t i = 0; i < F; ++i) if (live[i]) start[target[i] + 1]++;
  for (uint64_t t = 0; t < T; ++t) start[t + 1] += start[t];
  uint64_t* fill = (uint64_t*)malloc((T + 1) * sizeof(uint64_t));
  if (!fill) { free(start); free(recs); free(live); return -2; }
  memcpy(fill, start, (T + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < F; ++i) {
    if (!live[i]) continue;
    sort_rec r = { orc_depth_key(depth[i]), (uint32_t)i };
    recs[fill[target[i]]++] = r;
  }
  free(fill);
  free(live);
  for (uint64_t t = 0; t < T; ++t) {
    const uint64_t n = start[t + 1] - start[t];
    if (n > 1) qsort(&recs[start[t]], n, sizeof(sort_rec), rec_cmp);
    double c[4] = {0, 0, 0, 0};
    for (uint64_t j = start[t]; j < start[t + 1]; ++j) {
      const uint32_t id = recs[j].id;
      double f[4];
      if (quantize) {
        const uint32_t p = orc_encode(&rgba[4 * (uint64_t)id]);
        const double a = (double)(p >> 24) / 255.0;
        f[0] = orc_srgb_decode_exact((double)(p & 255u) / 255.0) * a;
        f[1] = orc_srgb_decode_exact((double)((p >> 8) & 255u) / 255.0) * a;
        f[2] = orc_srgb_decode_exact((double)((p >> 16) & 255u) / 255.0) * a;
        f[3] = a;
      } else {
        for (int k = 0; k < 4; ++k) f[k] = (double)rgba[4 * (uint64_t)id + k];
      }
      const double tt = 1.0 - c[3];
      for (int k = 0; k < 4; ++k) c[k] += tt * f[k];
    }
    const double tt = 1.0 - c[3];
    for (int k = 0; k < 4; ++k) out[4 * t + k] = c[k] + tt * ORC_DST(t, k);
  }
  free(start); free(recs);
  return 0;
#undef ORC_DST
}

/* ------------------------------------------------------------------ */
/* Replay of a given k-buffer (Simple / Spinlock composite)              */
/* ------------------------------------------------------------------ */

int orc_replay_kbuffer(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* rgba,
                       const uint32_t* count, const uint32_t* ab_color, const uint32_t* ab_key,
                       const uint32_t* ab_id, const uint8_t* tail_flag, int tail_quantized, float* out) {
  return orc_replay_kbuffer_cov(cfg, F, target, rgba, NULL, count, ab_color, ab_key, ab_id, NULL, tail_flag,
                                tail_quantized, out);
}

int orc_replay_kbuffer_cov(const orc_config* cfg, uint64_t F, const uint32_t* target, const float* rgba,
                           const uint32_t* mask, const uint32_t* count, const uint32_t* ab_color,
                           const uint32_t* ab_key, const uint32_t* ab_id, const uint32_t* ab_mask,
                           const uint8_t* tail_flag, int tail_quantized, float* out) {
  init_tables();
  const uint32_t K = cfg->layers;
  const uint64_t T = (uint64_t)cfg->width * cfg->height * cfg->samples;
  if (!valid_layers(K) || T == 0) return -1;
  const uint32_t cov = cfg->coverage_samples > 1 ? cfg->coverage_samples : 0;
  if (cov && (!mask || !ab_mask)) return -1;
  const uint32_t full = cov ? (1u << cov) - 1u : 0xFFFFFFFFu;
  float* colorbuf = (float*)malloc(T * 4 * sizeof(float));
  elem* a = (elem*)malloc(K * sizeof(elem));
  if (!colorbuf || !a) { free(colorbuf); free(a); return -2; }
  for (uint64_t t = 0; t < T; ++t) memcpy(&colorbuf[4 * t], cfg->background, 4 * sizeof(float));
  if (!cfg->opaque_key != !cfg->opaque_rgba) { free(colorbuf); free(a); return -1; }
  if (cfg->opaque_rgba && !cov) memcpy(colorbuf, cfg->opaque_rgba, T * 4 * sizeof(float));
  if ((cfg->tail_blend && tail_flag) || (cov && cfg->opaque_rgba)) {
    tail_list tails = {0, 0, 0};
    for (uint64_t i = 0; cfg->tail_blend && tail_flag && i < F; ++i) {
      if (!tail_flag[i]) continue;
      if (target[i] >= T) { free(tails.ev); free(colorbuf); free(a); return -1; }
      float c4[4];
      if (tail_quantized) orc_decode_premult(orc_encode(&rgba[4 * i]), c4);
      else memcpy(c4, &rgba[4 * i], sizeof c4);
      const uint32_t m = mask ? mask[i] & full : full;
      if (tail_push_m(&tails, (uint32_t)i, target[i], m, c4)) { free(tails.ev); free(colorbuf); free(a); return -2; }
    }
    const int frc = fold_tail_cov(&tails, colorbuf, T, 1, cov, cov ? cfg->opaque_rgba : NULL); /* id order, canonical */
    free(tails.ev);
    if (frc) { free(colorbuf); free(a); return frc; }
  }
  for (uint64_t t = 0; t < T; ++t) {
    int n = (int)(count[t] < K ? count[t] : K);
    for (int i = 0; i < n; ++i) {
      a[i].col = ab_color[t * K + i]; a[i].key = ab_key[t * K + i]; a[i].id = ab_id[t * K + i];
      a[i].mask = ab_mask ? ab_mask[t * K + i] : full;
    }
    bubble_sort(a, n);
    float c[4] = {0, 0, 0, 0};
    blend_front(c, a, n, cov);
    float o[4];
    memcpy(o, &colorbuf[4 * t], sizeof o);
    rop_blend(o, c);
    memcpy(&out[4 * t], o, sizeof o);
  }
  free(colorbuf); free(a);
  return 0;
}

int orc_finalize_srgb8(const float* in, uint32_t width, uint32_t height, uint32_t samples, uint32_t ss,
                       uint8_t* out) {
  init_tables();
  if (!in || !out || width == 0 || height == 0 || samples == 0 || ss == 0) return -1;
  const uint64_t W = (uint64_t)width * ss, H = (uint64_t)height * ss;
  const float n = (float)(samples * ss * ss);
  for (uint64_t y = 0; y < height; ++y) {
    for (uint64_t x = 0; x < width; ++x) {
      float sum[4] = {0, 0, 0, 0};
      for (uint32_t s = 0; s < samples; ++s)
        for (uint32_t dy = 0; dy < ss; ++dy)
          for (uint32_t dx = 0; dx < ss; ++dx) {
            const float* p = &in[4 * ((s * H + y * ss + dy) * W + x * ss + dx)];
            for (int k = 0; k < 4; ++k) sum[k] = sum[k] + p[k];
          }
      uint8_t* o = &out[4 * (y * width + x)];
      for (int k = 0; k < 3; ++k) o[k] = (uint8_t)code_rgb(sum[k] / n);
      const float a = fminf(fmaxf(sum[3] / n, 0.0f), 1.0f);
      o[3] = (uint8_t)floorf(a * 255.0f + 0.5f);
    }
  }
  return 0;
}
