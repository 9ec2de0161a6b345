// oit_device.hpp - device-side building blocks of the gfx950 OIT kernels:
// depth keys, the sRGB8 colour codec, premultiplied "over", wave-level
// aggregation and the lock-free id-ordered per-target lists.
//
// Every kernel translation unit is compiled with -ffp-contract=off, so each
// float expression below rounds once per operation, exactly like the
// single-threaded restatement the parity tests compare against.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oit {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;      // empty head / slot / list end
constexpr uint32_t kUnset = 0xFFFFFFFEu;      // link word not yet written during a list build
constexpr uint64_t kEmpty64 = 0xFFFFFFFFFFFFFFFFull;

// Device error bits (oit_stats.device_errors). Every loop in which a lane
// waits for other lanes (a list insert meeting a link word that is not yet
// written, the Spinlock lock) gives up after kSpinLimit trips, every Linked
// List walk requires strictly descending values, and a walk of an (unordered)
// colour-pass tail list ends after as many hops as the arrays hold
// fragments; a lane that gives up sets a bit
// in the context's error word and leaves the loop, so a corrupted structure
// ends the kernel with an error instead of hanging the queue. A correct
// frame waits at most a few trips per lane for a link word and, for the
// lock, at most the trips of the critical sections queued before it.
constexpr uint32_t kErrListWait = 1u;  // a list insert exceeded kSpinLimit trips
constexpr uint32_t kErrLockWait = 2u;  // a Spinlock section waited kSpinLimit trips for its lock
constexpr uint32_t kErrWalk = 4u;      // a list walk met a link that does not descend (or a tail walk overran)
constexpr uint32_t kSpinLimit = 1u << 20;

__device__ __forceinline__ void report_error(uint32_t* err, uint32_t bit) {
  if (err) __hip_atomic_fetch_or(err, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Next value of a walk along a descending list: `nx` when it lies below the
// current value `cur`, else 0 (the end of the walk) with kErrWalk reported,
// so a corrupted list (a cycle, a stale link) cannot keep a walker looping.
__device__ __forceinline__ uint32_t descend(uint32_t nx, uint32_t cur, uint32_t* err) {
  if (nx < cur) return nx;
  report_error(err, kErrWalk);
  return 0u;
}

// Colour codec tables, filled on the host at context creation.
//  thr[q]  (1 <= q <= 255): smallest binary32 c whose sRGB8 code is >= q,
//          for the curve of unPremultLinearToSRGB (shaderCommon.glsl:60-69)
//          followed by packUnorm4x8 rounding; thr[0] = -inf, thr[256] = +inf.
//  dec[u]: correctly rounded unPremultSRGBToLinear(u / 255) (shaderCommon.glsl:84-94).
//  alpha[u]: u / 255 (unpackUnorm4x8).
struct CodecTables {
  float thr[257];
  float dec[256];
  float alpha[256];
};

// Order-preserving u32 key of a binary32 depth. floatBitsToUint (used by
// oitLoop64.frag.glsl:79 and oitSpinlock.frag.glsl:60,68) orders only
// non-negative depths; this key orders every float and never produces the
// empty sentinel 0xFFFFFFFF (oitRender.cpp:307).
__device__ __forceinline__ uint32_t depth_key(float d) {
  const uint32_t b = __float_as_uint(d);
  const uint32_t k = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return k == kEmpty ? 0xFFFFFFFEu : k;
}

// sRGB8 code of an unpremultiplied linear channel: the unique q with
// thr[q] <= c < thr[q+1]. A fast approximation of the curve picks the start,
// the table makes it exact (NaN and negatives give 0).
__device__ __forceinline__ uint32_t code_rgb(float c, const float* __restrict__ thr) {
  float s;
  if (c < 0.0031308f) {
    s = c * 12.92f;
  } else {
    s = 1.055f * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(c) * (1.0f / 2.4f)) - 0.055f;
  }
  int q = (int)fminf(fmaxf(s * 255.0f + 0.5f, 0.0f), 255.0f);
  while (q > 0 && !(c >= thr[q])) --q;
  while (q < 255 && c >= thr[q + 1]) ++q;
  return (uint32_t)q;
}

// shading() -> unPremultLinearToSRGB -> packUnorm4x8 (oitLinkedList.frag.glsl:
// 57,75-77). The fragment stream carries premultiplied colour, so rgb is
// divided by alpha first. Alpha: floor(clamp(a,0,1) * 255 + 0.5).
__device__ __forceinline__ uint32_t encode_color(float4 p, const CodecTables* __restrict__ tab) {
  const float a = p.w;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if (a > 0.0f) {
    r = p.x / a;
    g = p.y / a;
    b = p.z / a;
  }
  const float ac = fminf(fmaxf(a, 0.0f), 1.0f);
  const uint32_t qa = (uint32_t)floorf(ac * 255.0f + 0.5f);
  return code_rgb(r, tab->thr) | (code_rgb(g, tab->thr) << 8) | (code_rgb(b, tab->thr) << 16) | (qa << 24);
}

// Branch-free code with the threshold table in LDS: the curve evaluated with
// v_log_f32 / v_exp_f32 lands within one code of the exact one, and the two
// neighbouring thresholds settle it. oit_selftest_codec checks this against
// code_rgb over every non-negative binary32 on the device.
__device__ __forceinline__ uint32_t code_rgb_fast(float c, const float* thr) {
  float s;
  if (c < 0.0031308f) {
    s = c * 12.92f;
  } else {
    s = 1.055f * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(c) * (1.0f / 2.4f)) - 0.055f;
  }
  int q = (int)fminf(fmaxf(s * 255.0f + 0.5f, 0.0f), 255.0f);
  const float lo = thr[q];
  const float hi = thr[q + 1];
  q -= (q > 0 && c < lo) ? 1 : 0;
  q += (q < 255 && c >= hi) ? 1 : 0;
  return (uint32_t)q;
}

__device__ __forceinline__ uint32_t encode_color_lds(float4 p, const float* thr) {
  const float a = p.w;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if (a > 0.0f) {
    r = p.x / a;
    g = p.y / a;
    b = p.z / a;
  }
  const float ac = fminf(fmaxf(a, 0.0f), 1.0f);
  const uint32_t qa = (uint32_t)floorf(ac * 255.0f + 0.5f);
  return code_rgb_fast(r, thr) | (code_rgb_fast(g, thr) << 8) | (code_rgb_fast(b, thr) << 16) | (qa << 24);
}

__device__ __forceinline__ float4 decode_premult_lds(uint32_t c, const float* dec, const float* alpha) {
  const float a = alpha[c >> 24];
  return make_float4(dec[c & 255u] * a, dec[(c >> 8) & 255u] * a, dec[(c >> 16) & 255u] * a, a);
}

// unpackUnorm4x8 -> unPremultSRGBToLinear -> premultiply
// (first half of doBlendPacked, shaderCommon.glsl:117-124).
__device__ __forceinline__ float4 decode_premult(uint32_t c, const CodecTables* __restrict__ tab) {
  const float a = tab->alpha[c >> 24];
  return make_float4(tab->dec[c & 255u] * a, tab->dec[(c >> 8) & 255u] * a, tab->dec[(c >> 16) & 255u] * a, a);
}

// top over base: doBlend (shaderCommon.glsl:108-112), which is also the
// PREMULTIPLIED framebuffer blend dst = src + (1 - src.a) * dst
// (main.cpp:548-558, oit.h:55) with src = top, dst = base.
__device__ __forceinline__ float4 over(float4 top, float4 base) {
  const float t = 1.0f - top.w;
  return make_float4(top.x + t * base.x, top.y + t * base.y, top.z + t * base.z, top.w + t * base.w);
}

// --------------------------------------------------------------------------
// Wave helpers (64-wide wavefronts)
// --------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// --------------------------------------------------------------------------
// Relaxed agent-scope atomics (coherent across XCDs, no L2 writeback)
// --------------------------------------------------------------------------

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t max_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t min_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// --------------------------------------------------------------------------
// Id-ordered lock-free lists
//
// A list is a chain of link words: the head word per target, then next[x]
// for each element x. Element values are ordered descending: newest first,
// the walk order of the reference's head-exchange lists
// (oitLinkedList.frag.glsl:72); the colour-pass tail lists use the same
// order and fold newest first. Insertion of a block of values [lo, hi] (a
// run of same-target lanes):
//   old = atomic_max(link, hi)  -- publishes the block
//   if old < lo: next[lo] = old  (one store)
//   else: walk to the successor of `old` and retry there.
// next[] words of elements that are about to be inserted hold kUnset, written
// by an EARLIER kernel; a walker that reaches a published element whose link
// is still kUnset waits one loop trip for the owner's store. Only relaxed
// atomics are used: no release fences (which would write back L2) are needed
// because no other data is read during the build.
// The loops are driven by a wave-wide ballot so that one trip advances every
// active lane by one step; a lane's store of its own link happens in the same
// trip as its successful atomic, so no lane ever waits on a store that sits
// behind its own wait.
// --------------------------------------------------------------------------

// Descending list (head holds the largest value, 0 terminates). A value v
// keeps its link word at next[(v >> kShift) - kBias]: the Linked List's
// values are node indices (kShift 0, kBias 0); the colour-pass tail lists'
// values are ((id + 1) << 1) | boundary flag (kShift 1, kBias 1). `cap`
// bounds the values whose link words exist.
// Inserts block [lo, hi] (hi >= lo; its values already chained hi -> ... ->
// lo by the caller; no value of the list outside the block lies between lo
// and hi) at link word `p0` when `want`.
// Called by every lane of the wave (lanes without a block pass want=false).
// The loop condition is a wave-wide ballot, so the compiler cannot turn a
// lane's wait for a kUnset link into an inner loop that the owner of that
// link (possibly a lane of the same wave) would never leave: every trip is
// one step for every active lane, and the owner's link store happens in the
// trip of its successful atomic.
template <int kShift = 0, uint32_t kBias = 0>
__device__ __forceinline__ void list_insert_desc(bool want, uint32_t* p0, uint32_t* next, uint32_t lo,
                                                 uint32_t hi, uint32_t cap, uint32_t* err) {
  bool active = want && hi < cap;
  bool first = true;
  uint32_t* p = p0;
  uint32_t trips = 0;
  while (__ballot(active)) {
    if (active && ++trips > kSpinLimit) {
      report_error(err, kErrListWait);
      active = false;
    }
    if (active) {
      bool try_max = first;
      first = false;
      if (!try_max) {
        const uint32_t v = ld_relaxed(p);
        if (v != kUnset) {
          if (v > hi) {
            if (v < cap) p = &next[(v >> kShift) - kBias]; else active = false;  // a link outside the pool: stop
          } else {
            try_max = true;
          }
        }
      }
      if (try_max) {
        const uint32_t old = max_relaxed(p, hi);
        if (old < lo) {
          st_relaxed(&next[(lo >> kShift) - kBias], old);
          active = false;
        } else if (old < cap) {
          p = &next[(old >> kShift) - kBias];  // a larger element got there first: continue behind it
        } else {
          active = false;
        }
      }
    }
  }
}

// Fragment-stream and record reads: every input word is read once per pass,
// so these loads carry the non-temporal hint and leave the L2 and the
// last-level cache to what the next kernel reads (at 1920x1080 the Linked
// List's nodes stay cached for the composite).
__device__ __forceinline__ uint32_t ld_stream(const uint32_t* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ uint64_t ld_stream(const uint64_t* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ld_stream(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float4 ld_stream(const float4* p) {
  const float* q = reinterpret_cast<const float*>(p);
  return make_float4(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1), __builtin_nontemporal_load(q + 2),
                     __builtin_nontemporal_load(q + 3));
}

// Waits until every memory operation of the lane has completed, so the
// stores before it are visible (at the memory side, for agent-scope relaxed
// accesses) before any memory operation after it is issued.
__device__ __forceinline__ void store_fence() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt unconstrained
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Waits for every memory operation the lane has in flight, at a point that
// the whole wave reaches. After a step of conditional atomics or loads
// (`if (live) x = atomic(...)`), the compiler otherwise keeps the returns
// pending on the paths that skip their consumers and then waits for every
// operation before each conditional access of the next step, which runs a
// wave's independent chains one after another instead of together.
__device__ __forceinline__ void returns_done() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Per-target lock for critical sections whose shared data is read and
// written only with agent-scope relaxed atomics (coherent across the XCDs'
// L2 caches, so the section needs ordering but no L2 writeback or
// invalidate, which an agent-scope acquire/release would add to every lock
// operation). Bit 0 of the word is the lock; the bits above hold a count
// that lanes add to whether the lock is held or not. Acquire: a relaxed OR
// of bit 0, taken when the bit was clear; the section's accesses are issued
// only after its value has returned. Release: wait until every memory
// operation of the lane has completed, then clear bit 0 with a relaxed AND.
__device__ __forceinline__ bool lock_try(uint32_t* l) {
  return (__hip_atomic_fetch_or(l, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1u) == 0u;
}
__device__ __forceinline__ void lock_release(uint32_t* l) {
  store_fence();
  __hip_atomic_fetch_and(l, ~1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// list_insert_desc for a lane that already tried the head and learned that a
// larger element is there: it starts at that element's link word `p0` and
// reads it before its first atomic.
template <int kShift = 0, uint32_t kBias = 0>
__device__ __forceinline__ void list_insert_desc_from(bool want, uint32_t* p0, uint32_t* next, uint32_t lo,
                                                      uint32_t hi, uint32_t cap, uint32_t* err) {
  bool active = want && hi < cap;
  uint32_t* p = p0;
  uint32_t trips = 0;
  while (__ballot(active)) {
    if (active && ++trips > kSpinLimit) {
      report_error(err, kErrListWait);
      active = false;
    }
    if (active) {
      bool try_max = false;
      const uint32_t v = ld_relaxed(p);
      if (v != kUnset) {
        if (v > hi) {
          if (v < cap) p = &next[(v >> kShift) - kBias]; else active = false;
        } else {
          try_max = true;
        }
      }
      if (try_max) {
        const uint32_t old = max_relaxed(p, hi);
        if (old < lo) {
          st_relaxed(&next[(lo >> kShift) - kBias], old);
          active = false;
        } else if (old < cap) {
          p = &next[(old >> kShift) - kBias];
        } else {
          active = false;
        }
      }
    }
  }
}

}  // namespace oit
