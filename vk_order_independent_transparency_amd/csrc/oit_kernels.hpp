// oit_kernels.hpp - host-side launch interface of the gfx950 OIT kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oit_device.hpp"

namespace oit {

// Device view of one appended fragment batch.
struct FragView {
  const uint32_t* target;
  const float* depth;
  const float4* rgba;
  uint32_t count;    // fragments in this batch
  uint32_t id_base;  // frame-local id of the first fragment
  const uint32_t* mask = nullptr;  // coverage shading: sample mask per fragment (gl_SampleMaskIn)
};

// Colour-pass tail lists: per target, unordered (the composite restores id
// order). Tail queue regions: 1024 frame-local fragment ids each.
constexpr uint32_t kTailRegionShift = 10;
inline uint64_t tail_regions(uint64_t cap) { return (cap + (1u << kTailRegionShift) - 1) >> kTailRegionShift; }

struct TailBuf {
  uint32_t* head;     // [T] last linked tail value, 0 = empty
  uint32_t* next;     // [Fcap] link words, indexed by frame-local id
  float4* rgba;       // [Fcap] premultiplied colour the tail blend uses
  uint64_t* rec;      // [Fcap] tail queue: (target << 32) | ((id + 1) << 1 | boundary), by region of 1024 ids
  uint32_t* qn;       // [tail_regions(Fcap)] tail fragments of each region (its queue entries), zero at the frame's start
  uint32_t cap;       // Fcap
  uint32_t* mask = nullptr;  // [Fcap] coverage shading: samples the ROP blend writes
  uint32_t* err = nullptr;   // device error word (kErr* bits), null: none
};

// Per-algorithm A-buffer views.
struct LLBuf {
  uint32_t* head;     // [T] 0 = empty (node 0 is the terminator); pixel-major under sample shading
  uint64_t* cd;       // [cap] (depth key << 32) | packed colour
  uint32_t* next;     // [cap]
  uint64_t cap;       // nodes including the terminator
  uint32_t* overflow; // LL fragments that found the pool full
  uint32_t* counters; // frame counters [16]; the resetting resolve moves them to counters[16..31]
  uint32_t* mask = nullptr;  // [cap] coverage shading: the node's sample mask
  float4* full = nullptr;    // [cap] ordered interlock: the shaded colour (its own tail blend uses it)
  uint32_t* err = nullptr;   // device error word (kErr* bits), null: none
};

struct KBuf {          // Simple / Spinlock
  uint32_t* count;    // [T] Simple's count (the Spinlock keeps its count in the lock word)
  uint32_t* lock;     // [T] Spinlock only: count << 1 | held (0 = empty, free); null for Simple
  uint32_t* depth;    // [T] Spinlock early-depth bound
  uint32_t* col;      // [K][T]
  uint32_t* key;      // [K][T]
  uint32_t* id;       // [K][T]
  uint32_t* mask = nullptr;  // [K][T] coverage shading
  uint32_t* err = nullptr;   // device error word (kErr* bits), null: none
  uint32_t* filled = nullptr;  // [T] Spinlock: stored places below K whose stores have completed
};

struct WBBuf {
  float4* accum;      // 4 T floats: planes of r, g, b, a (plane k at ((float*)accum) + k T)
  float* reveal;      // [T] log2 of the revealage (sum of log2(1 - alpha)); 0 = fully revealed
  // accum and reveal are indexed pixel-major under sample shading (Common::spp)
};

struct ResolveOut {
  float4* out;        // [T]
  uint32_t* fk_key;   // [T][K] or null
  uint32_t* fk_val;   // [T][K] or null
  uint32_t* counts;   // [T] or null
};

struct Common {
  uint32_t T;              // targets
  uint32_t K;              // layers
  int tail;                // tail blending on
  float4 background;
  const CodecTables* tab;
  uint32_t cov = 0;        // coverage shading: samples per pixel (0 = off)
  uint32_t spp = 1;        // sample shading: samples per pixel, targets sample-major (1 = off)
  // the frame's opaque pass (null when it has none), per sample target:
  // (depth key << 32 | id) and colour; with coverage shading sample s of
  // pixel t is entry s * T + t
  const uint64_t* opq_z = nullptr;
  const float4* opq_rgba = nullptr;
};

// One frame of a batch as the batched Linked List kernels see it.
struct BatchFrameDev {
  const uint32_t* target;  // frame-local targets
  const float* depth;
  const float4* rgba;
  float4* out;             // [T] composite over the background
  float4 background;
  uint32_t count;          // fragments
  uint32_t T;              // targets
  uint32_t target_base;    // head index of target 0 (multiple of 256)
  uint32_t frag_base;      // fragment 0 is pool node frag_base + 1 (multiple of 512)
  uint32_t mode;           // kModeList or kModeBin
  uint32_t pad[3];
};
constexpr uint32_t kModeList = 0;  // a Linked List per target (append, list walk)
constexpr uint32_t kModeBin = 1;   // a bucket per target (count, scan, scatter, ordered replay)
constexpr uint32_t kBatchMaxGroups = 16;  // (K, tail) groups: 6 layer counts x 2

struct BatchArgs {
  const BatchFrameDev* frames;  // device table, bases ascending
  uint32_t f0, f1;              // frames this launch covers
  uint32_t units;               // append: waves of 64 * chunks fragments from frames[f0].frag_base
  uint32_t tile0, tiles;        // resolve: tiles of 256 targets
  const CodecTables* tab;
  // append schedule (null: frame after frame): the frames' 512-fragment
  // slices in `rounds` rounds; round q holds slices floor(q S_f / rounds) ..
  // floor((q + 1) S_f / rounds) - 1 of every frame f (S_f its slices), frame
  // after frame. sched[0 .. rounds] are the rounds' first slices, then per
  // round the (f1 + 1) prefix offsets of the frames within it
  const uint32_t* sched = nullptr;
  uint32_t rounds = 0;
  // binned frames: per target the bucket cursor, per tile the scan base,
  // and the queue of buckets too large for one lane (with its length)
  uint32_t* cursor = nullptr;    // [cursor_cap], target t's at (t % 64) * (cursor_cap / 64) + t / 64
  uint32_t* tile_sum = nullptr;
  uint32_t* deep_q = nullptr;
  uint32_t* deep_n = nullptr;
  uint32_t cursor_cap = 0;       // a multiple of 64
};
constexpr uint32_t kBatchSlice = 512;  // fragments of a schedule slice (the frames' alignment)

// sRGB8 codec tables built on the host (exact curve, bisection in double).
void build_codec_tables(CodecTables* t);

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
// Several fills in one launch (ranges at 16-byte aligned addresses; others
// are filled by launches of their own).
struct FillRange {
  uint32_t* p;
  uint64_t n;
  uint32_t v;
};
constexpr int kMaxFillRanges = 6;
hipError_t launch_fill_ranges(const FillRange* r, int count, hipStream_t s);
hipError_t launch_batch_append(const BatchArgs& a, const LLBuf& ll, uint32_t chunks, hipStream_t s);
hipError_t launch_batch_resolve(const BatchArgs& a, const LLBuf& ll, uint32_t K, bool tail, bool pair, bool keys,
                                uint32_t grid_cap, hipStream_t s);
// Binned frames after the append launch (which counted them): the scan of
// the counts into bucket cursors over `tiles` tiles from a.tile0 and the
// frame table's frames, the scatter pass over a.units units, and the sort
// of the buckets too large for one lane to order.
hipError_t launch_bin_prepare(const BatchArgs& a, const LLBuf& ll, uint32_t tiles, uint32_t chunks, hipStream_t s);
// Composite of the binned frames among a group's tiles.
hipError_t launch_bin_resolve(const BatchArgs& a, const LLBuf& ll, uint32_t K, bool tail, uint32_t grid_cap,
                              hipStream_t s);

hipError_t launch_ll_store(const Common& c, const FragView& f, const LLBuf& ll, const TailBuf& tb, hipStream_t s);
hipError_t launch_ll_link(const Common& c, const FragView& f, const LLBuf& ll, hipStream_t s);
// Fused colour pass: store + ordered link in one kernel. Requires next[] of
// the frame's nodes to hold kUnset on entry (left so by a resetting resolve
// or by the clear).
hipError_t launch_ll_append(const Common& c, const FragView& f, const LLBuf& ll, const TailBuf& tb, int chunks,
                            hipStream_t s);
// Composite that also returns heads to 0, walked nodes' links to kUnset and
// moves the frame counters aside; pair_walk consumes node p-1 in the same
// round trip as p when the list continues there.
hipError_t launch_ll_resolve_reset(const Common& c, const LLBuf& ll, const TailBuf& tb, const ResolveOut& o,
                                   bool pair_walk, bool keys_only, hipStream_t s);

hipError_t launch_simple_store(const Common& c, const FragView& f, const KBuf& kb, const TailBuf& tb, hipStream_t s);
hipError_t launch_spin_store(const Common& c, const FragView& f, const KBuf& kb, const TailBuf& tb, hipStream_t s);
hipError_t launch_kbuf_resolve(const Common& c, const KBuf& kb, const TailBuf& tb, const ResolveOut& o, hipStream_t s);

// chunks: 64-fragment chunks per wave (1: one fragment per lane; 2 or 4:
// every chunk's loads and chain steps in flight together)
// Per-target counts of a segment's fragments (at append).
hipError_t launch_loop64_count(const Common& c, const FragView& f, uint32_t* cnt, hipStream_t s);
// The stores of a segment, once every segment of the frame is counted;
// `big` gets one mask per 64-fragment chunk of the segment (lanes whose
// pixel holds more than K fragments).
hipError_t launch_loop64_store(const Common& c, const FragView& f, uint64_t* slots, uint32_t* cnt, uint64_t* big,
                               int chunks, hipStream_t s);
hipError_t launch_loop64_classify(const Common& c, const FragView& f, const uint64_t* slots, const uint64_t* big,
                                  const TailBuf& tb, hipStream_t s);
// Also sets the slots it reads back to empty.
hipError_t launch_loop64_resolve(const Common& c, uint64_t* slots, uint32_t* cnt, const TailBuf& tb,
                                 const ResolveOut& o, hipStream_t s);

hipError_t launch_wboit_store(const Common& c, const FragView& f, const WBBuf& wb, float view_near, float view_span,
                              hipStream_t s);
hipError_t launch_wboit_resolve(const Common& c, const WBBuf& wb, const ResolveOut& o, hipStream_t s);

// Links the tail queue's records of frame-local ids [first, frags) (the ones
// that can be tails) into the per-target lists.
hipError_t launch_tail_link(const TailBuf& tb, uint32_t first, uint32_t frags, hipStream_t s);
// coverage shading (Common::cov samples per pixel, masks in the buffers)
hipError_t launch_ll_resolve_cov(const Common& c, const LLBuf& ll, const TailBuf& tb, const ResolveOut& o,
                                 hipStream_t s);
hipError_t launch_kbuf_resolve_cov(const Common& c, const KBuf& kb, const TailBuf& tb, const ResolveOut& o,
                                   hipStream_t s);
// Loop32 (oitLoop.frag.glsl): depth pass per append, colour pass per append
// at resolve time, composite
hipError_t launch_loop32_depth(const Common& c, const FragView& f, uint32_t* zs, hipStream_t s);
hipError_t launch_loop32_color(const Common& c, const FragView& f, const uint32_t* zs, uint64_t* cs,
                               const TailBuf& tb, hipStream_t s);
// Also sets the slots it reads back to empty.
hipError_t launch_loop32_resolve(const Common& c, uint32_t* zs, uint64_t* cs, const TailBuf& tb, const ResolveOut& o,
                                 hipStream_t s);
// ordered interlock: Linked List append with shaded colours, then per-pixel
// critical sections replayed in primitive order
hipError_t launch_interlock_resolve(const Common& c, const LLBuf& ll, const ResolveOut& o, hipStream_t s);
// Opaque pass: planes of n sample targets to the clear (depth 1.0, colour
// bg), then a draw: depth test LESS with depth write (u64 atomicMin of
// (key, id)), then the winners' colours
hipError_t launch_opaque_clear(uint64_t* z, float4* rgba, uint64_t n, float4 bg, hipStream_t s);
hipError_t launch_opaque_draw(const Common& c, const FragView& f, uint64_t* z, float4* rgba, hipStream_t s);
hipError_t launch_finalize_srgb8(const float4* in, uint32_t W, uint32_t H, uint32_t S, uint32_t ss,
                                 const CodecTables* tab, uint32_t* out, hipStream_t s);

hipError_t launch_codec_selftest(const CodecTables* tab, uint32_t lo, uint32_t hi, unsigned long long* mismatches,
                                 uint32_t* first_bad, hipStream_t s);

}  // namespace oit
