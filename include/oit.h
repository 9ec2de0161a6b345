/*
 * oit.h - C ABI of the MI355X (gfx950) order-independent-transparency
 * fragment-store and composite path.
 *
 * The reference (nvpro-samples/vk_order_independent_transparency) has no C
 * ABI: each algorithm is a method pair on its Sample class,
 * clearTransparent<Alg>() / drawTransparent<Alg>() (oit.h:379-425), dispatched
 * by switch(m_state.algorithm) in Sample::onRender (oitRender.cpp:43-66,
 * 125-148), with all buffers owned by Sample (oit.h:173-214) and sized in
 * Sample::createFrameImages (oit.cpp:84-163). Each entry point below names
 * the reference code it replaces.
 *
 * Conventions
 *  - Every function returns an int status (OIT_OK == 0, negative on error)
 *    and never aborts. Overflow of an A-buffer is not an error: it is
 *    tail-blended as in the reference and reported by oit_stats().
 *  - Pointers in oit_fragments and output pointers are DEVICE pointers.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Calls are stream-ordered; a context is not thread-safe.
 *  - A "target" is one A-buffer list: sample * width * height + y * width + x
 *    (the reference's per-sample image-array addressing,
 *    oitColorDepthDefines.glsl:433-441, oitCompositeDefines.glsl:282-290).
 *  - Fragment ids are assigned in append order: the i-th fragment of the
 *    k-th oit_append call of a frame has id = (fragments appended earlier in
 *    the frame) + i. Ids are the canonical draw order and break depth ties.
 */
#ifndef OIT_MI355X_H
#define OIT_MI355X_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OIT_ABI_VERSION 1

/* Algorithms; values equal the reference's shaders/common.h:44-51. */
enum oit_algorithm {
  OIT_SIMPLE     = 0,
  OIT_LINKEDLIST = 1,
  OIT_LOOP       = 2, /* two geometry passes: the fragment buffers are read again by oit_resolve */
  OIT_LOOP64     = 3,
  OIT_SPINLOCK   = 4,
  OIT_INTERLOCK  = 5, /* ordered critical sections, replayed in primitive order */
  OIT_WEIGHTED   = 6
};

/* Kernel variant bits (oit_config.variant); results are identical for every
 * variant, only the kernel schedule differs. */
#define OIT_VARIANT_LL_SPLIT_APPEND  0x1u /* Linked List: separate store and link kernels */
#define OIT_VARIANT_LL_SINGLE_WALK   0x2u /* Linked List: resolve walks one node per round trip */
#define OIT_VARIANT_LL_APPEND_8      0x4u /* Linked List: fused append takes 8 chunks per wave (default 4) */
#define OIT_VARIANT_LL_KEYS_ONLY     0x8u /* Linked List: resolve keeps sort keys only and reloads colours by node
                                             (always at 32 layers) */
#define OIT_VARIANT_STORE_CHUNKS_1   0x10u /* Loop64: store one fragment per lane (default: two 64-fragment chunks per wave) */
#define OIT_VARIANT_STORE_CHUNKS_4   0x20u /* Loop64: store four 64-fragment chunks per wave */

/* Status codes. */
enum oit_status {
  OIT_OK               = 0,
  OIT_ERR_INVALID_ARG  = -1,
  OIT_ERR_OOM          = -2,
  OIT_ERR_OVER_BUDGET  = -3,
  OIT_ERR_DEVICE       = -4,
  OIT_ERR_UNSUPPORTED  = -5,
  OIT_ERR_CAPACITY     = -6,
  OIT_ERR_STATE        = -7,
  OIT_ERR_NOT_READY    = -8
};

/* Renderer configuration; mirrors the reference's State (oit.h:64-116) and
 * the per-frame SceneData contract (shaders/common.h:77-92). */
typedef struct oit_config {
  uint32_t algorithm;          /* State::algorithm (oit.h:66) */
  uint32_t layers;             /* OIT_LAYERS = State::oitLayers (oit.h:67): 1,2,4,8,16,32 */
  uint32_t tail_blend;         /* OIT_TAILBLEND = State::tailBlend (oit.h:70) */
  uint32_t width;              /* viewport width  (SceneData.viewport.x) */
  uint32_t height;             /* viewport height (SceneData.viewport.y) */
  uint32_t samples;            /* sample planes: 1, or the MSAA count with sample shading */
  uint32_t ll_nodes_per_pixel; /* State::linkedListAllocatedPerElement (oit.h:68), used when ll_pool_nodes == 0 */
  uint32_t variant;            /* kernel variant bits for tuning (OIT_VARIANT_*); 0 = library default */
  uint64_t ll_pool_nodes;      /* Linked List capacity in nodes INCLUDING terminator node 0
                                  (SceneData.linkedListAllocatedPerElement for OIT_LINKEDLIST,
                                  oit.cpp:125); 0 = ll_nodes_per_pixel * width * height * samples */
  uint64_t max_fragments;      /* per-frame fragment capacity of the tail lists; 0 = grow on demand */
  float    background[4];      /* premultiplied RGBA the colour pass starts from (opaque gray 0.5 by default) */
  float    wboit_view_near;    /* WBOIT: view distance = near + depth * span ... */
  float    wboit_view_span;    /* ... stands in for -viewZ (object.vert.glsl:35) */
  uint32_t coverage_samples;   /* OIT_COVERAGE_SHADING (oit.h:82): MSAA samples (2, 4, 8) of per-pixel
                                  fragments with sample masks; 0 or 1 = off. Needs samples == 1; Simple,
                                  Linked List and Spinlock (Loop/Loop64 have none, README.md:33-34) */
  uint32_t reserved;           /* 0 */
} oit_config;

/* A batch of fragments in draw order (structure of arrays, device memory). */
typedef struct oit_fragments {
  const uint32_t* target; /* list index, see conventions above */
  const float*    depth;  /* gl_FragCoord.z (any finite float; order-preserving key) */
  const float*    rgba;   /* premultiplied linear RGBA, 4 floats per fragment */
  uint64_t        count;
  const uint32_t* coverage; /* coverage shading: the fragment's sample mask (gl_SampleMaskIn[0],
                               oitColorDepthDefines.glsl:41-47); NULL otherwise */
} oit_fragments;

/* Memory footprint, like the reference GUI's "Object Sizes" (oitGui.cpp:327-336). */
typedef struct oit_sizes {
  uint64_t abuffer_bytes;     /* A-buffer / node pool / WBOIT accumulation targets */
  uint64_t aux_bytes;         /* per-target counters, heads, locks, depth bounds */
  uint64_t tail_bytes;        /* per-fragment tail-list structures currently allocated */
  uint64_t total_bytes;
  uint64_t ll_capacity;       /* LL nodes including the terminator */
  uint32_t bytes_per_element; /* per stored fragment */
  uint32_t targets;           /* width * height * samples */
  uint32_t reserved;
  uint64_t resident_bytes;    /* device bytes the context holds (reconfigurations reuse them) */
} oit_sizes;

/* Counters of the last frame (synchronises `stream`). */
typedef struct oit_stats {
  uint64_t fragments;         /* appended this frame (the LL counter value, oitLinkedList.frag.glsl:55) */
  uint64_t ll_overflow;       /* LL fragments that found the pool full and were tail-blended */
  uint64_t tail_fragments;    /* fragments recorded in the colour-pass tail lists */
  uint64_t appends;           /* oit_append calls this frame */
  uint64_t device_errors;     /* bits of device-side waits that gave up since the last oit_get_stats: 1 a list
                                 insert, 2 a Spinlock lock, 4 a list walk that met a non-descending link (or a
                                 colour-pass tail walk longer than the frame's fragments); the
                                 call then returns OIT_ERR_DEVICE and the affected frames' images are undefined */
} oit_stats;

typedef struct oit_ctx oit_ctx;

/* Creates a context on HIP device `device`. Replaces Sample::onAttach's
 * resource set-up (main.cpp:74-128). */
int oit_create(oit_ctx** out, int device);
/* Frees every device buffer. Replaces Sample::onDetach/destroy* (oit.cpp). */
int oit_destroy(oit_ctx* ctx);
/* Validates `cfg`, (re)allocates the A-buffer, aux targets and pool.
 * Replaces updateRendererFromState + createFrameImages (main.cpp:130-250,
 * oit.cpp:45-220). */
int oit_configure(oit_ctx* ctx, const oit_config* cfg);
/* Reports the allocation sizes of the current configuration. */
int oit_query_sizes(const oit_ctx* ctx, oit_sizes* out);
/* Frees every frame buffer (after the context's last stream-ordered call has
 * completed) and leaves the context unconfigured. No counterpart in the
 * reference; for hosts that share a device-memory budget between contexts. */
int oit_trim(oit_ctx* ctx);
/* Starts a frame: clears counters, heads, locks and k-buffers to the
 * reference's clear values. Replaces clearTransparent<Alg> (oitRender.cpp:
 * 156-174, 201-216, 303-311, 337-356) and the WBOIT clear values
 * (oitRender.cpp:394-397). */
int oit_clear(oit_ctx* ctx, void* stream);
/* Fragment-store ("color") pass for a batch of fragments. Replaces the
 * COLOR draw of drawTransparent<Alg> (oitRender.cpp:176-186, 218-227,
 * 313-322, 358-368, 384-414) running oitSimple/oitLinkedList/oitLoop64/
 * oitSpinlock/oitWeighted.frag.glsl with PASS == PASS_COLOR. The fragment
 * buffers must stay valid until oit_resolve has been enqueued when the
 * algorithm is OIT_LOOP64 or OIT_LOOP: Loop64 counts each pixel's fragments
 * here and stores them in the resolve, once every append of the frame is
 * counted (a pixel of at most K fragments keeps them all and needs no
 * atomicMin chain); Loop32 runs its second geometry pass there. */
int oit_append(oit_ctx* ctx, const oit_fragments* frags, void* stream);
/* Composite pass: writes the final premultiplied RGBA per target
 * (composite over tail-blended colour buffer over background) to
 * out_rgba[4 * targets]. Replaces the COMPOSITE full-screen draw
 * (oitRender.cpp:191-198, 232-239, 327-334, 373-381, 416-424) with its
 * PREMULTIPLIED / WEIGHTED_COMPOSITE framebuffer blend (main.cpp:548-558,
 * 578-588). */
int oit_resolve(oit_ctx* ctx, float* out_rgba, void* stream);
/* Like oit_resolve, and additionally writes per target the sorted front-K
 * depth keys and values (fragment ids; packed colours for OIT_LOOP64) to
 * frontk_key/frontk_val[targets * layers] (0xFFFFFFFF padding) and the
 * algorithm's per-target count to counts[targets] (OIT_SPINLOCK: every
 * fragment that reached the target, where the reference skips those its early
 * depth test rejects, so beyond `layers` the two differ). Any of the three
 * extra pointers may be NULL. For parity tests; not a reference entry point. */
int oit_resolve_debug(oit_ctx* ctx, float* out_rgba, uint32_t* frontk_key,
                      uint32_t* frontk_val, uint32_t* counts, void* stream);
/* Failure-handling tests only (no reference counterpart): writes `value`
 * into word `index` of one of the context's device structures, stream
 * ordered, so that a test can plant a corrupted link or a lock that is never
 * released and check that the bounded device waits end with OIT_ERR_DEVICE
 * (oit_get_stats) instead of hanging. `index` is checked against the
 * structure's allocated words (OIT_ERR_INVALID_ARG beyond it). */
enum {
  OIT_DEBUG_LL_HEAD = 0,    /* Linked List / interlock head per target (with samples > 1 the
                               heads are pixel-major: pixel * samples + sample) */
  OIT_DEBUG_LL_NEXT = 1,    /* Linked List / interlock link word per node */
  OIT_DEBUG_SPIN_LOCK = 2   /* Spinlock lock word per target (count << 1 | held) */
};
int oit_debug_set_word(oit_ctx* ctx, uint32_t what, uint64_t index, uint32_t value, void* stream);
/* Copies the frame counters to the host (synchronises `stream`). Device
 * waits are bounded: a list insert or a Spinlock section that waits past
 * its trip limit, or a list walk that meets a link that does not descend,
 * leaves its loop and marks the context; the next oit_get_stats fills *out,
 * clears the mark and returns OIT_ERR_DEVICE (a corrupted structure ends the
 * kernels with an error instead of hanging the queue). */
int oit_get_stats(oit_ctx* ctx, oit_stats* out, void* stream);

/* Display image of a resolved frame: the MSAA resolve (average of `samples`
 * sample planes, main.cpp:700-714) and the supersample downsample (average
 * of each supersample x supersample block, :715-731) of `in_rgba`
 * ([samples][height*supersample][width*supersample] premultiplied float4),
 * then sRGB8 colour (exact curve, round to nearest; the B8G8R8A8_SRGB
 * target, oit.cpp:58) and linear 8-bit alpha: out_rgba8 is
 * [height][width] RGBA bytes (device memory). */
int oit_finalize_srgb8(oit_ctx* ctx, const float* in_rgba, uint32_t width, uint32_t height, uint32_t samples,
                       uint32_t supersample, uint8_t* out_rgba8, void* stream);
/* Opaque draw of the frame (the opaque spheres drawn before the transparent
 * ones, oitRender.cpp:68-122, with the depth-writing BlendMode::NONE
 * pipeline, main.cpp:526-533). The frame's first call clears the opaque
 * planes to depth 1.0 and the background colour (oitRender.cpp:89-91);
 * each call then writes, per sample target, the depth and colour of the
 * nearest fragment (depth test LESS: the earliest of equal depths; at or
 * behind depth 1.0 nothing). With coverage shading `frags->coverage` holds
 * each fragment's sample mask and the planes are per sample. Once a frame
 * has an opaque draw, its transparent fragments are depth tested against it
 * before they are stored (early fragment tests with post_depth_coverage,
 * oitColorDepthDefines.glsl:24-37; with coverage shading the mask keeps the
 * samples in front, an empty mask drops the fragment) and the composite
 * blends over the opaque colour instead of the background. Call after
 * oit_clear and before the frame's first oit_append (OIT_ERR_STATE
 * otherwise); a count of 0 still gives the frame its depth-1.0 test.
 * Linked List nodes stay indexed by transparent fragment id: a culled
 * fragment's node is left unused. */
int oit_draw_opaque(oit_ctx* ctx, const oit_fragments* frags, void* stream);
/* Human-readable name of a status code. */
const char* oit_status_string(int status);

/* ---- Stage timing (replaces the reference's per-pass GPU timer sections,
 * e.g. "LinkedListColor" / "LinkedListComposite", oitRender.cpp:223,235).
 * Stages: 0 clear, 1 store (colour pass), 2 link (LL head insertion),
 * 3 classify (Loop64 tail classification), 4 tail link, 5 resolve. ---- */
#define OIT_PROFILE_STAGES 6
typedef struct oit_profile {
  double   ms[OIT_PROFILE_STAGES];
  uint64_t launches[OIT_PROFILE_STAGES];
} oit_profile;

/* Brackets every later kernel launch of the context with HIP events. */
int oit_profile_enable(oit_ctx* ctx, int enable);
/* Waits for the recorded events and returns the accumulated stage times. */
int oit_profile_read(oit_ctx* ctx, oit_profile* out);
/* Zeroes the accumulated stage times. */
int oit_profile_reset(oit_ctx* ctx);

/* ---- Codec checks (no reference counterpart). ------------------------- */
/* The library's sRGB8 tables: encode thresholds thr[0..256] (thr[0] = -inf,
 * thr[256] = +inf), decode dec[0..255], alpha[0..255]. Host computation. */
int oit_codec_tables(float* thr257, float* dec256, float* alpha256);
/* Runs the device's branch-free encoder against its exact table search on
 * every binary32 bit pattern in [lo_bits, hi_bits) and reports the number of
 * differing codes and the first differing pattern (0xFFFFFFFF if none).
 * Synchronous. */
int oit_selftest_codec(oit_ctx* ctx, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches,
                       uint32_t* first_bad);

/* ---- Device plumbing for hosts without their own HIP binding (no
 * reference counterpart: the reference allocates through VMA, oit.h:173-214,
 * and times with its profiler). Thin wrappers of hipMalloc/hipMemcpy/
 * streams/events on the library's own HIP runtime. ---------------------- */
int oit_device_count(int* out);
int oit_set_device(int device);
int oit_device_alloc(void** out, uint64_t bytes);
int oit_device_free(void* ptr);
/* Stream-ordered copies (`stream` may be NULL); host memory is pageable, so
 * they return after the copy has been staged. */
int oit_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream);
int oit_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream);
int oit_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream);
int oit_stream_create(void** out);
int oit_stream_destroy(void* stream);
int oit_stream_synchronize(void* stream);
int oit_event_create(void** out);
int oit_event_destroy(void* event);
int oit_event_record(void* event, void* stream);
int oit_event_synchronize(void* event);
int oit_event_elapsed_ms(float* out, void* start, void* stop);

/* ---- Synthetic scene (stands in for the reference's rasterised sphere
 * mesh, main.cpp:334-417): analytic spheres, orthographic, each covered
 * pixel (or sample) centre emits a front and a back fragment. ---------- */
typedef struct oit_scene_params {
  uint64_t seed;
  uint32_t width, height;
  uint32_t samples;           /* 1, 2, 4 or 8 (the standard sample positions) */
  uint32_t spheres;
  float    radius_min, radius_max;
  float    alpha_min, alpha_max;
  float    cluster_fraction;  /* fraction of spheres packed in a cluster_extent-wide square */
  float    cluster_extent;
} oit_scene_params;

/* Number of fragments the scene emits (host computation). */
int oit_scene_count(const oit_scene_params* p, uint64_t* out_count);
/* Writes the fragment stream in draw order into HOST arrays of `capacity`
 * fragments (target[], depth[], rgba[4*]). */
int oit_scene_generate_host(const oit_scene_params* p, uint32_t* target, float* depth,
                            float* rgba, uint64_t capacity, uint64_t* out_count);

/* Device generator of the same stream (bit-identical to
 * oit_scene_generate_host, same order). Replaces the reference's scene
 * set-up and rasterisation of the transparent spheres as the producer of
 * fragments (main.cpp:334-417, object.vert.glsl:32-39,
 * shaderCommon.glsl:48-56). Create builds the sphere table on the host and
 * counts the fragments on `device` (synchronously); emit writes the stream
 * into DEVICE arrays (target 8-byte aligned, depth 8-byte aligned, rgba
 * 16-byte aligned) on `stream`, and may be called any number of times. */
typedef struct oit_scene_gen oit_scene_gen;
int oit_scene_gen_create(oit_scene_gen** out, const oit_scene_params* p, int device);
int oit_scene_gen_destroy(oit_scene_gen* g);
int oit_scene_gen_count(const oit_scene_gen* g, uint64_t* out_count);
int oit_scene_gen_emit(oit_scene_gen* g, uint32_t* target, float* depth, float* rgba,
                       uint64_t capacity, void* stream);

/* ---- Batched Linked List frames ------------------------------------------
 * No counterpart in the reference, whose onRender draws one frame per command
 * buffer (oitRender.cpp:28-154); this is the batched form of the Linked List
 * pair clearTransparent/drawTransparent<LinkedList> (oitRender.cpp:201-227)
 * for callers with many frames in flight. Frames of any size, depth
 * complexity and K share one resident node pool sized from their exact
 * fragment counts; one append launch covers every frame and one resolve
 * launch covers each (K, tail_blend) group. Each frame's image equals the
 * single-frame Linked List result with a pool of at least count + 1 nodes,
 * bit for bit. */
typedef struct oit_batch_frame {
  uint32_t width, height, samples;
  uint32_t layers;            /* K: 1, 2, 4, 8, 16 or 32 */
  uint32_t tail_blend;        /* OIT_TAILBLEND */
  uint32_t reserved;          /* 0 */
  float    background[4];     /* premultiplied RGBA under the composite */
  oit_fragments frags;        /* device SoA in draw order, frame-local targets; rgba 16-byte aligned */
  float*   out;               /* device [width*height*samples][4], 16-byte aligned */
} oit_batch_frame;

enum {
  OIT_BATCH_CONCURRENT = 0x1, /* resolve groups run on forked streams (join before return) */
  OIT_BATCH_GRAPH      = 0x2, /* a run is captured once per frame set and replayed as a hipGraph */
  OIT_BATCH_FRAME_ORDER = 0x4 /* append frame after frame; by default the frames' fragment slices are
                                 interleaved so that each frame advances in draw order with a share of
                                 the waves in flight (fewer same-pixel inserts at once, same results) */
};

/* Kernel choices of a batch: append chunks per wave (4 or 8) and, per
 * (K, tail_blend) group, the resolve variant bits (OIT_VARIANT_LL_SINGLE_WALK,
 * OIT_VARIANT_LL_KEYS_ONLY) and a cap on resolve blocks (0 = one block per
 * tile). Index [k][tail] with k = log2(K). */
typedef struct oit_batch_tuning {
  uint32_t append_chunks;
  uint32_t resolve_variant[6][2];
  uint32_t resolve_blocks[6][2];
  float    bin_density;       /* frames with at least this many fragments per target are binned: a
                                 counting sort gathers each target's fragments into one bucket and the
                                 composite replays it in descending id order (sorted in LDS), in place of
                                 the list append and walk; the same image. 0 = every frame; default 3e38
                                 (none), which oit_batch_tune may lower */
} oit_batch_tuning;

typedef struct oit_batch_sizes {
  uint64_t pool_nodes;        /* nodes the frames need, terminator included */
  uint64_t head_entries;      /* padded targets */
  uint64_t bytes;             /* device bytes for those (nodes 12 B, heads 4 B) */
  uint32_t groups;            /* resolve launches */
  uint32_t frames;
} oit_batch_sizes;

typedef struct oit_batch oit_batch;
int oit_batch_create(oit_batch** out, int device, uint32_t flags);
int oit_batch_destroy(oit_batch* b);
/* Pool and head sizes `frames` need (no device work). */
int oit_batch_query(const oit_batch* b, const oit_batch_frame* frames, uint32_t n, oit_batch_sizes* out);
/* Grows the resident pool to at least these sizes (never shrinks). */
int oit_batch_reserve(oit_batch* b, uint64_t pool_nodes, uint64_t head_entries);
/* Appends and resolves every frame on `stream`. Grows the pool when needed
 * (the only allocation). Fragment and output buffers must stay valid until
 * the stream reaches the end of the run. */
int oit_batch_run(oit_batch* b, const oit_batch_frame* frames, uint32_t n, void* stream);
/* Times the kernel variants on these frames (synchronous, runs the batch
 * several times) and keeps the fastest per group for later runs. */
int oit_batch_tune(oit_batch* b, const oit_batch_frame* frames, uint32_t n, void* stream);
int oit_batch_get_tuning(const oit_batch* b, oit_batch_tuning* out);
int oit_batch_set_tuning(oit_batch* b, const oit_batch_tuning* t);

/* ---- One Linked List frame without allocation -----------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. Renders `frame` (its fragments in draw order, frame-local
 * targets, no coverage masks) exactly as oit_clear + oit_append + oit_resolve
 * of a context configured for OIT_LINKEDLIST with the frame's layers,
 * tail_blend, background and ll_pool_nodes = count + 1 (so no fragment
 * overflows): the same image bit for bit. The heads and node pool live in a
 * library-owned device array of OIT_NOALLOC_SCRATCH_BYTES that each GPU the
 * library runs on holds until the process ends; a frame needs about
 * 4 * targets + 12 * (count + 1) bytes of it (plus 4 KiB) and a larger one
 * returns OIT_ERR_CAPACITY. The call issues no hipMalloc/hipFree/hipMemcpy/
 * hipMemset, only kernel launches on `stream` (the first call on a device
 * looks up the array's address). Every call on one device shares its array:
 * calls that overlap on one device race (serialise them, e.g. on one stream).
 * `frame->out` must be 16-byte aligned; `frame->samples` planes of
 * width * height targets as in oit_config. */
#define OIT_NOALLOC_SCRATCH_BYTES (256ull << 20)
int oit_ll_render_noalloc(int device, const oit_batch_frame* frame, void* stream);

/* ---- Request loop ---------------------------------------------------------
 * No counterpart in the reference (one interactive frame loop,
 * Sample::onRender, oitRender.cpp:28-154). Requests (a seeded scene, an
 * algorithm, K) are taken from any thread; up to `slots` of them run at
 * once, each slot on its own stream driven by a thread of its own
 * (continuous batching: a
 * request that arrives while others run starts as soon as a slot and memory
 * are free, and its kernels overlap theirs). Each request's fragments are
 * generated and counted on the device; its A-buffer is sized from that exact
 * count (Linked List pool = fragments + 1 nodes) and granted against
 * `budget_bytes` in arrival order; contexts keep their buffers between
 * requests and idle ones are trimmed when the budget needs room; a request
 * larger than the whole budget completes with OIT_ERR_OVER_BUDGET. Each
 * result equals the single-frame result of the same request. */
typedef struct oit_server_config {
  uint32_t slots;             /* requests in flight, 1..16 */
  uint32_t reserved;          /* 0 */
  uint64_t budget_bytes;      /* device bytes for A-buffers, pools, aux and tail arrays, and the slots' fragment-stream and output buffers */
} oit_server_config;

typedef struct oit_request {
  oit_scene_params scene;
  uint32_t algorithm;         /* any oit_algorithm */
  uint32_t layers;            /* K */
  uint32_t tail_blend;
  uint32_t reserved;          /* 0 */
  float*   out;               /* device [W*H*S][4] (16-byte aligned), or NULL for the slot's own buffer */
  uint64_t tag;               /* returned in the completion */
} oit_request;

typedef struct oit_completion {
  uint64_t tag;
  int32_t  status;            /* OIT_OK or the first error */
  uint32_t slot;
  uint64_t fragments;         /* exact count of the request's scene */
  uint64_t abuffer_bytes;     /* bytes granted from the budget */
  const float* out;           /* the image (a slot's own buffer stays valid until the slot's next request) */
  double   t_submit_us;       /* steady clock, microseconds since the server was created */
  double   t_start_us;        /* memory granted, kernels enqueued */
  double   t_done_us;         /* completion seen by the slot's thread */
  double   device_us;         /* device time of the request's work (HIP events: the count, then generation
                                 and the OIT passes; host gaps between them excluded); 0 on error */
} oit_completion;

typedef struct oit_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t held_bytes;        /* device bytes the contexts hold now */
  uint64_t peak_held_bytes;
  uint32_t peak_in_flight;
  uint32_t slots;
} oit_server_stats;

typedef struct oit_server oit_server;
int oit_server_create(oit_server** out, int device, const oit_server_config* cfg);
/* Finishes every submitted request, then frees everything. */
int oit_server_destroy(oit_server* s);
/* Queues a request (stamps its arrival time); returns at once. */
int oit_server_submit(oit_server* s, const oit_request* r);
/* Moves up to `max` completions of oit_server_submit requests into out[],
 * *n of them (oit_server_replay collects its own requests' completions). */
int oit_server_poll(oit_server* s, oit_completion* out, uint32_t max, uint32_t* n);
/* Waits until every submitted request has completed (OIT_ERR_NOT_READY on timeout). */
int oit_server_drain(oit_server* s, uint32_t timeout_ms);
/* Load driver: submits reqs[i] at arrival_us[i] (non-decreasing, relative to
 * the call) from the calling thread, waits for all, and writes out[i] for
 * reqs[i] (with the caller's tag). Its requests are marked with the call, so
 * completions of other callers' requests are never taken for them. */
int oit_server_replay(oit_server* s, const oit_request* reqs, const double* arrival_us, uint32_t n,
                      uint32_t timeout_ms, oit_completion* out);
int oit_server_get_stats(oit_server* s, oit_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* OIT_MI355X_H */
