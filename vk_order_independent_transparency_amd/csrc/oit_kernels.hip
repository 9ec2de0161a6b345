// oit_kernels.hip - gfx950 kernels of the OIT fragment-store and composite
// passes. One thread per fragment for the store passes, one thread per
// target (pixel or sample) for the composite passes; 256-thread blocks of
// four 64-wide waves.
//
// Built with: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off
//             -Xclang -target-feature -Xclang -packed-fp32-ops
#include "oit_kernels.hpp"

namespace oit {

namespace {

constexpr uint32_t kBlock = 256;

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// A run of consecutive lanes that carry the same target. Fragment ids grow
// with the lane, so a run is a block of consecutive ids of one target, and
// no fragment outside the wave can have an id inside it.
struct Run {
  uint32_t lane;
  uint32_t start_lane;
  uint32_t end_lane;
  bool start;
  bool end;
};

// Must be reached by all 64 lanes (callers keep every lane alive and give
// lanes without a fragment a target of kEmpty).
__device__ __forceinline__ Run same_target_run(uint32_t t) {
  Run r;
  r.lane = lane_id();
  const uint32_t tp = __shfl_up(t, 1u);
  const uint32_t tn = __shfl_down(t, 1u);
  r.start = (r.lane == 0) || (tp != t);
  r.end = (r.lane == 63) || (tn != t);
  const uint64_t starts = __ballot(r.start);
  const uint64_t ends = __ballot(r.end);
  const uint64_t le = (2ull << r.lane) - 1ull;  // lanes <= this lane
  r.start_lane = 63u - (uint32_t)__builtin_clzll(starts & le);
  r.end_lane = r.lane + (uint32_t)__builtin_ctzll(ends >> r.lane);
  return r;
}

// Colour-pass tail list value of frame-local fragment lid: ((lid + 1) << 1)
// | boundary flag, so that 0 can terminate the lists; its link word is
// next[lid]. Values order as ids do.
__device__ __forceinline__ uint32_t tail_value(uint32_t lid, uint32_t flag) { return ((lid + 1u) << 1) | flag; }
__device__ __forceinline__ uint32_t tail_index(uint32_t v) { return (v >> 1) - 1u; }
// values below this have link words in arrays of `cap` fragments
__device__ __forceinline__ uint32_t tail_value_cap(uint32_t cap) { return (cap + 1u) << 1; }

// Colour-pass tail lists, phase 1: a tail fragment `lid` stores its
// premultiplied colour at index lid, and its record (target << 32) | tail
// value goes to the frame's tail queue. The queue is cut into regions of
// 1024 fragment ids, each with its own counter, and a record goes to the
// region of its fragment's id: a fragment enters the tail once, so region
// r's entries fit in rec[r * 1024 ..]. The lanes of one region share one
// returning atomic, so a wave waits for one round trip however many tail
// fragments it finds. The queue holds only tail fragments, so the link pass
// reads nothing of the others. Called by every active lane of the wave that
// may hold a tail fragment.
__device__ __forceinline__ void tail_queue_push(bool tail, uint32_t t, uint32_t lid, uint32_t flag, float4 payload,
                                                const TailBuf& tb, uint32_t mask) {
  tail = tail && lid < tb.cap && tb.qn != nullptr;
  if (tail) {
    tb.rgba[lid] = payload;
    if (tb.mask) tb.mask[lid] = mask;
  }
  const uint32_t r = lid >> kTailRegionShift;
  const uint32_t lane = lane_id();
  // group the lanes by region (cross-lane work only), then one atomic per
  // group, all of them in flight together
  uint64_t rem = __ballot(tail);
  uint32_t my_leader = 0;
  uint64_t my_group = 0;
  while (rem) {
    const uint32_t leader = (uint32_t)__builtin_ctzll(rem);
    const uint32_t rl = __builtin_amdgcn_readlane(r, leader);
    const uint64_t m = __ballot(tail && r == rl);
    if ((m >> lane) & 1ull) {
      my_leader = leader;
      my_group = m;
    }
    rem &= ~m;
  }
  uint32_t q = 0;
  if (tail && lane == my_leader) q = atomicAdd(&tb.qn[r], (uint32_t)__builtin_popcountll(my_group));
  q = __shfl(q, (int)my_leader);
  if (tail)
    tb.rec[((uint64_t)r << kTailRegionShift) + q + lanes_below(my_group)] = ((uint64_t)t << 32) | tail_value(lid, flag);
}

// tail_queue_push over kU chunks of a wave (chunk u: tail[u], t[u], lid[u],
// payload[u]; no boundary flag, full coverage), with every chunk's region
// counter atomics in flight together: the wave waits one round trip for all
// of them. Must be reached by all 64 lanes.
template <int kU>
__device__ __forceinline__ void tail_queue_push_n(const bool (&tail_in)[kU], const uint32_t (&t)[kU],
                                                  const uint32_t (&lid)[kU], const float4 (&payload)[kU],
                                                  const TailBuf& tb) {
  const uint32_t lane = lane_id();
  bool tail[kU];
  uint32_t r[kU], leader[kU], q[kU];
  uint64_t group[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    tail[u] = tail_in[u] && lid[u] < tb.cap && tb.qn != nullptr;
    if (tail[u]) {
      tb.rgba[lid[u]] = payload[u];
      if (tb.mask) tb.mask[lid[u]] = 0xFFFFFFFFu;
    }
    r[u] = lid[u] >> kTailRegionShift;
    // group the lanes by region (cross-lane work only)
    uint64_t rem = __ballot(tail[u]);
    leader[u] = 0;
    group[u] = 0;
    while (rem) {
      const uint32_t l = (uint32_t)__builtin_ctzll(rem);
      const uint32_t rl = __builtin_amdgcn_readlane(r[u], l);
      const uint64_t m = __ballot(tail[u] && r[u] == rl);
      if ((m >> lane) & 1ull) {
        leader[u] = l;
        group[u] = m;
      }
      rem &= ~m;
    }
    q[u] = 0u;
    asm volatile("" : "+v"(q[u]));  // the zeros are written before any atomic goes out
  }
#pragma unroll
  for (int u = 0; u < kU; ++u)
    if (tail[u] && lane == leader[u]) q[u] = atomicAdd(&tb.qn[r[u]], (uint32_t)__builtin_popcountll(group[u]));
  returns_done();
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint32_t base = __shfl(q[u], (int)leader[u]);
    if (tail[u])
      tb.rec[((uint64_t)r[u] << kTailRegionShift) + base + lanes_below(group[u])] =
          ((uint64_t)t[u] << 32) | tail_value(lid[u], 0u);
  }
}

// A converged wave's tail fragments straight onto their targets' lists,
// for a pass that runs after every fragment of the frame is stored (no link
// pass follows), over kU chunks of the wave (chunk u: tail[u], t[u], lid[u],
// ...): neighbouring lanes of one target are chained in the wave and
// published with one exchange on the list head (k_tail_link does the same
// for queued records), every chunk's exchanges in flight together; each
// chunk's count goes to its region's counter, which the queue would
// otherwise have filled. Must be reached by all 64 lanes (lanes without a
// tail fragment pass tail = false).
template <int kU>
__device__ __forceinline__ void tail_list_push_n(const bool (&tail)[kU], const uint32_t (&t)[kU],
                                                 const uint32_t (&lid)[kU], const uint32_t (&flag)[kU],
                                                 const float4 (&payload)[kU], const TailBuf& tb) {
  const uint32_t lane = lane_id();
  bool ex[kU];
  uint32_t lo[kU], old[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const bool tl = tail[u] && lid[u] < tb.cap;
    if (tl) {
      tb.rgba[lid[u]] = payload[u];
      if (tb.mask) tb.mask[lid[u]] = 0xFFFFFFFFu;
    }
    const uint32_t v = tail_value(lid[u], flag[u]);
    const Run r = same_target_run(tl ? t[u] : kEmpty);
    const uint32_t v_prev = __shfl_up(v, 1u);
    lo[u] = __shfl(v, (int)r.start_lane);
    if (tl && !r.start) st_relaxed(&tb.next[lid[u]], v_prev);
    ex[u] = tl && r.end;
    const uint64_t m = __ballot(tl);
    if (m != 0ull && lane == (uint32_t)__builtin_ctzll(m))
      atomicAdd(&tb.qn[lid[u] >> kTailRegionShift], (uint32_t)__builtin_popcountll(m));
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    old[u] = 0u;
    asm volatile("" : "+v"(old[u]));  // the zeros are written before any exchange goes out
  }
#pragma unroll
  for (int u = 0; u < kU; ++u)
    if (ex[u]) old[u] = __hip_atomic_exchange(&tb.head[t[u]], tail_value(lid[u], flag[u]), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
  returns_done();
#pragma unroll
  for (int u = 0; u < kU; ++u)
    if (ex[u]) st_relaxed(&tb.next[tail_index(lo[u])], old[u]);
}

// Sorted register array of K (sort key, colour) pairs, ascending; inserting x
// returns the element pushed out of the back (x itself if it does not fit):
// insertionSortTail, oitCompositeDefines.glsl:116-139, unrolled into selects.
template <int K>
struct FrontK {
  uint64_t sk[K];
  uint32_t col[K];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      sk[k] = kEmpty64;
      col[k] = 0u;
    }
  }
  __device__ __forceinline__ void insert(uint64_t x, uint32_t xc, uint64_t& ev, uint32_t& evc) {
    const bool fits = x < sk[K - 1];
    ev = fits ? sk[K - 1] : x;
    evc = fits ? col[K - 1] : xc;
#pragma unroll
    for (int j = K - 1; j > 0; --j) {
      const bool before = x < sk[j - 1];
      const bool here = x < sk[j];
      sk[j] = before ? sk[j - 1] : (here ? x : sk[j]);
      col[j] = before ? col[j - 1] : (here ? xc : col[j]);
    }
    const bool first = x < sk[0];
    sk[0] = first ? x : sk[0];
    col[0] = first ? xc : col[0];
  }
  // The same insert when at most w - 1 slots can be occupied (w is wave
  // uniform): slots at and beyond w hold the +inf sentinel before and after
  // the insert, so whole blocks of four slots at or beyond w are skipped
  // with one scalar branch each (executing more slots is also exact).
  __device__ __forceinline__ void insert_bounded(uint64_t x, uint32_t xc, uint64_t& ev, uint32_t& evc, int w) {
    const bool fits = x < sk[K - 1];
    ev = fits ? sk[K - 1] : x;
    evc = fits ? col[K - 1] : xc;
#pragma unroll
    for (int b = ((K - 1) / 4) * 4; b >= 0; b -= 4) {
      if (b < w) {
#pragma unroll
        for (int j = (b + 3 < K - 1 ? b + 3 : K - 1); j >= b && j > 0; --j) {
          const bool before = x < sk[j - 1];
          const bool here = x < sk[j];
          sk[j] = before ? sk[j - 1] : (here ? x : sk[j]);
          col[j] = before ? col[j - 1] : (here ? xc : col[j]);
        }
      }
    }
    const bool first = x < sk[0];
    sk[0] = first ? x : sk[0];
    col[0] = first ? xc : col[0];
  }
};

// FrontK with a third payload moved alongside (the ordered interlock's sample
// masks under coverage shading); the same bounded insert.
template <int K>
struct FrontKM {
  uint64_t sk[K];
  uint32_t col[K];
  uint32_t msk[K];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      sk[k] = kEmpty64;
      col[k] = 0u;
      msk[k] = 0u;
    }
  }
  __device__ __forceinline__ void insert_bounded(uint64_t x, uint32_t xc, uint32_t xm, uint64_t& ev, uint32_t& evc,
                                                 int w) {
    const bool fits = x < sk[K - 1];
    ev = fits ? sk[K - 1] : x;
    evc = fits ? col[K - 1] : xc;
#pragma unroll
    for (int b = ((K - 1) / 4) * 4; b >= 0; b -= 4) {
      if (b < w) {
#pragma unroll
        for (int j = (b + 3 < K - 1 ? b + 3 : K - 1); j >= b && j > 0; --j) {
          const bool before = x < sk[j - 1];
          const bool here = x < sk[j];
          sk[j] = before ? sk[j - 1] : (here ? x : sk[j]);
          col[j] = before ? col[j - 1] : (here ? xc : col[j]);
          msk[j] = before ? msk[j - 1] : (here ? xm : msk[j]);
        }
      }
    }
    const bool first = x < sk[0];
    sk[0] = first ? x : sk[0];
    col[0] = first ? xc : col[0];
    msk[0] = first ? xm : msk[0];
  }
};

// Sorted register array of K sort keys only (depth key << 32 | node index),
// ascending: the colours are reloaded by node index after the walk, which
// keeps one third fewer registers live during the latency-bound walk.
template <int K>
struct FrontKeys {
  uint64_t sk[K];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int k = 0; k < K; ++k) sk[k] = kEmpty64;
  }
  // insertionSortTail with a wave-uniform bound w on the occupied slots + 1
  __device__ __forceinline__ uint64_t insert_bounded(uint64_t x, int w) {
    const uint64_t ev = x < sk[K - 1] ? sk[K - 1] : x;
#pragma unroll
    for (int b = ((K - 1) / 4) * 4; b >= 0; b -= 4) {
      if (b < w) {
#pragma unroll
        for (int j = (b + 3 < K - 1 ? b + 3 : K - 1); j >= b && j > 0; --j) {
          const bool before = x < sk[j - 1];
          const bool here = x < sk[j];
          sk[j] = before ? sk[j - 1] : (here ? x : sk[j]);
        }
      }
    }
    sk[0] = x < sk[0] ? x : sk[0];
    return ev;
  }
};

// Largest value of v over the active lanes, as a wave-uniform (scalar) value.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return __builtin_amdgcn_readfirstlane(v);
}

// Front-to-back blend of the first n sorted packed colours
// (oitLinkedList.frag.glsl:162-166, oitSimple.frag.glsl:161-165).
template <int K>
__device__ __forceinline__ float4 blend_front(const FrontK<K>& a, uint32_t n, const CodecTables* tab) {
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((uint32_t)k < n) c = over(c, decode_premult(a.col[k], tab));
  return c;
}

// Tail values held in registers per walk of an unordered tail list, and
// the walks a fold takes before it sorts a longer list instead. At 1080p
// with K = 8 the longest tail is 14 (1 % of the tails need a second walk);
// a window of 16 cost the K = 8 composites two to three waves per SIMD.
constexpr int kTailWindow = 8;
constexpr uint32_t kTailWalks = 4;

// One walk of an unordered tail list: the kTailWindow largest values below
// `bound`, in descending order (0 = none), by an insertion network in
// registers; returns how many values of the list lie below `bound`, and in
// `flagged` the number of boundary-flagged values of the whole list. A walk
// longer than the arrays hold (a cycle) reports kErrWalk and sets `bad`.
__device__ __forceinline__ uint32_t tail_window(const TailBuf& tb, uint32_t head, uint32_t vcap, uint32_t bound,
                                                uint32_t (&w)[kTailWindow], uint32_t& flagged, bool& bad) {
#pragma unroll
  for (int k = 0; k < kTailWindow; ++k) w[k] = 0u;
  uint32_t n = 0, fl = 0, hops = 0;
  uint32_t v = head;
  while (v != 0u && v < vcap) {  // a link outside the arrays ends the walk
    if (++hops > tb.cap) {
      report_error(tb.err, kErrWalk);
      bad = true;
      break;
    }
    const uint32_t nx = tb.next[tail_index(v)];
    fl += v & 1u;
    if (v < bound) {
      ++n;
      uint32_t x = v;
#pragma unroll
      for (int k = 0; k < kTailWindow; ++k) {
        const uint32_t hi = w[k] > x ? w[k] : x;
        x = w[k] > x ? x : w[k];
        w[k] = hi;
      }
    }
    v = nx;
  }
  flagged = fl;
  return n;
}

// A tail list longer than kTailWindow * kTailWalks, sorted into descending
// order in place: a bottom-up merge sort of the singly linked list (runs of
// 1, 2, 4, ... merged pass by pass, each pass relinking every node once),
// O(n log n) link reads instead of one walk per window. The list is the
// composite lane's own and acyclic (tail_window walked it to its end first);
// a link at or above vcap ends it. Returns the new head.
__device__ __forceinline__ uint32_t tail_sort_desc(const TailBuf& tb, uint32_t head, uint32_t vcap) {
  auto nxt = [&](uint32_t v) -> uint32_t {
    const uint32_t x = tb.next[tail_index(v)];
    return x < vcap ? x : 0u;
  };
  for (uint32_t run = 1;; run <<= 1) {
    uint32_t p = head, out = 0u, last = 0u, merges = 0u;
    while (p != 0u) {
      ++merges;
      uint32_t q = p, np = 0u;
      while (np < run && q != 0u) {
        ++np;
        q = nxt(q);
      }
      uint32_t nq = run;
      while (np > 0u || (nq > 0u && q != 0u)) {
        uint32_t e;
        if (np == 0u || (nq > 0u && q != 0u && q > p)) {  // descending: the larger value first
          e = q;
          q = nxt(q);
          --nq;
        } else {
          e = p;
          p = nxt(p);
          --np;
        }
        if (last == 0u) out = e;
        else tb.next[tail_index(last)] = e;
        last = e;
      }
      p = q;
    }
    if (last != 0u) tb.next[tail_index(last)] = 0u;
    head = out;
    if (merges <= 1u) return head;
  }
}

// Colour buffer after the colour pass: the tail fragments of target t over
// the background in id order (the ROP's primitive order, main.cpp:548-558),
// folded newest first: acc = t_n over ... over t_1, then acc over the
// background. The list is unordered, so each walk takes the next
// kTailWindow values in descending order (one walk for most targets, up to
// kTailWalks) and loads their colours four at a time; a longer tail is
// sorted in place and folded in one walk. `skip` boundary-flagged entries are in the A-buffer and not blended
// (Loop64 identical keys, see k_loop64_classify): the flagged ones with the
// lowest ids.
// tail_fold with the list's head value already loaded (a composite issues
// that load with its own first loads instead of after them).
__device__ __forceinline__ float4 tail_fold_from(const TailBuf& tb, uint32_t head, float4 bg, uint32_t skip) {
  const uint32_t vcap = tail_value_cap(tb.cap);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  if (head == 0u || head >= vcap) return over(acc, bg);
  uint32_t keep = 0xFFFFFFFFu;  // flagged entries to blend
  uint32_t bound = 0xFFFFFFFFu;
  bool bad = false;
  for (bool first = true;; first = false) {
    uint32_t w[kTailWindow];
    uint32_t flagged;
    const uint32_t n = tail_window(tb, head, vcap, bound, w, flagged, bad);
    if (first && skip > 0u) keep = flagged > skip ? flagged - skip : 0u;
    if (first && !bad && n > (uint32_t)kTailWindow * kTailWalks) {  // a long tail: sorted, then one walk
      for (uint32_t v = tail_sort_desc(tb, head, vcap); v != 0u && v < vcap;) {
        const uint32_t id = tail_index(v);
        const uint32_t nx = tb.next[id];
        if (!(v & 1u) || keep > 0u) {
          acc = over(acc, tb.rgba[id]);
          if (v & 1u) --keep;
        }
        v = nx;
      }
      break;
    }
#pragma unroll
    for (int g = 0; g < kTailWindow; g += 4) {
      if (w[g] == 0u) break;
      float4 c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = w[g + j] ? tb.rgba[tail_index(w[g + j])] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t v = w[g + j];
        if (v != 0u && (!(v & 1u) || keep > 0u)) {
          acc = over(acc, c[j]);
          if (v & 1u) --keep;
        }
      }
    }
    if (bad || n <= (uint32_t)kTailWindow) break;
    bound = w[kTailWindow - 1];
  }
  return over(acc, bg);
}

__device__ __forceinline__ float4 tail_fold(const TailBuf& tb, uint32_t t, float4 bg, uint32_t skip) {
  return tail_fold_from(tb, tb.head[t], bg, skip);
}

// Sample shading with spp samples per pixel lays targets out sample-major
// (sample * pixels + pixel). pm_index(t) is target t's place in pixel-major
// order (pixel * spp + sample) and pm_target(a) its inverse. The stream gives
// a fragment's covered samples consecutive ids, so structures indexed
// pixel-major (the WBOIT accumulators) take one wave's atomics on one run of
// entries, and composites that walk targets in pixel-major order give a
// pixel's samples neighbouring lanes, whose list nodes share cache lines.
// spp is 1, 2, 4 or 8: shifts and compares, no divisions.
__device__ __forceinline__ uint32_t pm_index(uint32_t t, uint32_t T, uint32_t spp) {
  if (spp <= 1u) return t;
  const uint32_t sh = (uint32_t)__builtin_ctz(spp);
  const uint32_t px = T >> sh;
  uint32_t smp = 0;
  for (uint32_t k = 1; k < spp; ++k) smp += t >= k * px ? 1u : 0u;
  return ((t - smp * px) << sh) | smp;
}
__device__ __forceinline__ uint32_t pm_target(uint32_t a, uint32_t T, uint32_t spp) {
  if (spp <= 1u) return a;
  const uint32_t sh = (uint32_t)__builtin_ctz(spp);
  return (a & (spp - 1u)) * (T >> sh) + (a >> sh);
}

// Colour buffer of target t before the colour-pass tail: the colour of the
// frame's opaque pass, else the clear colour.
__device__ __forceinline__ float4 dest_of(const Common& c, uint32_t t) {
  return c.opq_rgba ? c.opq_rgba[t] : c.background;
}

// Early depth test of a transparent fragment against the frame's opaque pass
// (early_fragment_tests with post_depth_coverage, oitColorDepthDefines.glsl:
// 24-37; VK_COMPARE_OP_LESS without depth write, main.cpp:526-528,541-543):
// the samples of mask m whose opaque depth key is larger than key (one test
// at target t unless coverage shading). Zero: the fragment does not exist.
__device__ __forceinline__ uint32_t opaque_visible(const uint64_t* z, uint32_t T, uint32_t cov, uint32_t t,
                                                   uint32_t key, uint32_t m) {
  if (!z) return m;
  if (cov > 1u) {
    uint32_t r = 0u;
    for (uint32_t s = 0; s < cov; ++s)
      if (((m >> s) & 1u) && key < (uint32_t)(z[(uint64_t)s * T + t] >> 32)) r |= 1u << s;
    return r;
  }
  return key < (uint32_t)(z[t] >> 32) ? m : 0u;
}

// Coverage shading: the ROP blends each tail fragment into the samples of
// its mask; per sample the newest-first fold over the sample's colour (the
// opaque pass's, else the clear colour), then the sum over samples in order
// divided by their count (the MSAA resolve). Without an opaque pass a target
// without tail fragments keeps the clear colour.
__device__ __forceinline__ float4 tail_fold_cov(const TailBuf& tb, uint32_t t, const Common& c) {
  const uint32_t S = c.cov;
  const uint32_t vcap = tail_value_cap(tb.cap);
  const uint32_t head = tb.head ? tb.head[t] : 0u;
  const bool empty = head == 0u || head >= vcap;
  if (empty && !c.opq_rgba) return c.background;
  float4 acc[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) acc[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t bound = 0xFFFFFFFFu;
  bool bad = false;
  while (!empty) {  // windows of the unordered list, newest first (tail_fold_from)
    uint32_t w[kTailWindow];
    uint32_t flagged;
    const uint32_t n = tail_window(tb, head, vcap, bound, w, flagged, bad);
    if (bound == 0xFFFFFFFFu && !bad && n > (uint32_t)kTailWindow * kTailWalks) {  // a long tail: sorted, one walk
      for (uint32_t v = tail_sort_desc(tb, head, vcap); v != 0u && v < vcap;) {
        const uint32_t id = tail_index(v);
        const uint32_t nx = tb.next[id];
        const uint32_t m = tb.mask[id];
        const float4 col = tb.rgba[id];
#pragma unroll
        for (int s = 0; s < 8; ++s)
          if ((uint32_t)s < S && ((m >> s) & 1u)) acc[s] = over(acc[s], col);
        v = nx;
      }
      break;
    }
#pragma unroll
    for (int g = 0; g < kTailWindow; g += 4) {
      if (w[g] == 0u) break;
      float4 col[4];
      uint32_t msk[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool h = w[g + j] != 0u;
        col[j] = h ? tb.rgba[tail_index(w[g + j])] : make_float4(0.f, 0.f, 0.f, 0.f);
        msk[j] = h ? tb.mask[tail_index(w[g + j])] : 0u;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int s = 0; s < 8; ++s)
          if ((uint32_t)s < S && ((msk[j] >> s) & 1u)) acc[s] = over(acc[s], col[j]);
    }
    if (bad || n <= (uint32_t)kTailWindow) break;
    bound = w[kTailWindow - 1];
  }
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if ((uint32_t)s < S) {
      const float4 cb = over(acc[s], c.opq_rgba ? c.opq_rgba[(uint64_t)s * c.T + t] : c.background);
      sum = make_float4(sum.x + cb.x, sum.y + cb.y, sum.z + cb.z, sum.w + cb.w);
    }
  }
  const float fs = (float)S;
  return make_float4(sum.x / fs, sum.y / fs, sum.z / fs, sum.w / fs);
}

// Coverage-shaded front blend (oitLinkedList.frag.glsl:147-160,
// oitSimple.frag.glsl:142-159): per sample, front to back over the sorted
// layers whose mask holds it; the sample colours summed in order and divided
// by the sample count.
template <int K>
__device__ __forceinline__ float4 blend_front_cov(const uint32_t* cols, const uint32_t* masks, uint32_t n, uint32_t S,
                                                  const float* s_dec, const float* s_alpha) {
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t s = 0; s < S; ++s) {
    float4 sc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < K; ++k)
      if ((uint32_t)k < n && ((masks[k] >> s) & 1u)) sc = over(sc, decode_premult_lds(cols[k], s_dec, s_alpha));
    sum = make_float4(sum.x + sc.x, sum.y + sc.y, sum.z + sc.z, sum.w + sc.w);
  }
  const float fs = (float)S;
  return make_float4(sum.x / fs, sum.y / fs, sum.z / fs, sum.w / fs);
}

template <int K>
__device__ __forceinline__ void write_debug(const ResolveOut& o, uint32_t t, const FrontK<K>& a, uint32_t n,
                                            uint32_t count, bool ids_from_nodes, bool vals_are_colors) {
  if (o.fk_key) {
#pragma unroll
    for (int k = 0; k < K; ++k) o.fk_key[(uint64_t)t * K + k] = (uint32_t)k < n ? (uint32_t)(a.sk[k] >> 32) : kEmpty;
  }
  if (o.fk_val) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint32_t v = kEmpty;
      if ((uint32_t)k < n) {
        v = vals_are_colors ? a.col[k] : (uint32_t)a.sk[k];
        if (ids_from_nodes) v -= 1u;  // node n holds fragment id n - 1
      }
      o.fk_val[(uint64_t)t * K + k] = v;
    }
  }
  if (o.counts) o.counts[t] = count;
}

// ---------------------------------------------------------------------------
// Fill
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t n4 = n / 4;
  uint4* p4 = reinterpret_cast<uint4*>(p);
  const uint4 v4 = make_uint4(v, v, v, v);
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += stride) p4[i] = v4;
  for (uint64_t i = n4 * 4 + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Up to kMaxFillRanges fills in one launch (a frame's clear): 16-byte stores
// over the ranges' concatenated 16-byte units, the last n mod 4 words of
// each range by the grid's first threads.
struct FillArgs {
  uint32_t* p[kMaxFillRanges];
  uint64_t n[kMaxFillRanges];
  uint64_t start[kMaxFillRanges + 1];  // prefix of the ranges' 16-byte units
  uint32_t v[kMaxFillRanges];
  int count;
};

__global__ __launch_bounds__(kBlock) void k_fill_ranges(FillArgs a) {
  const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  if (gid < 4u * (uint64_t)a.count) {
    const int j = (int)(gid >> 2);
    const uint64_t e = gid & 3u;
    if (e < (a.n[j] & 3u)) a.p[j][(a.n[j] & ~3ull) + e] = a.v[j];
  }
  const uint64_t total = a.start[a.count];
  int j = 0;
  for (uint64_t i = gid; i < total; i += stride) {
    while (i >= a.start[j + 1]) ++j;  // a thread's units increase, so its range index only grows
    const uint32_t v = a.v[j];
    reinterpret_cast<uint4*>(a.p[j])[i - a.start[j]] = make_uint4(v, v, v, v);
  }
}

// ---------------------------------------------------------------------------
// Linked List: oitLinkedList.frag.glsl
// ---------------------------------------------------------------------------

// Store half of the colour pass (oitLinkedList.frag.glsl:51-85 minus the head
// exchange): node n = id + 1 (the reference's counter, :55, advances once per
// fragment in draw order; here the allocation is the fragment's own index, so
// no global counter is contended), overflow -> tail (:59-68), node payload
// (:75-82). The link word is set to kUnset for the link kernel.
// Codec thresholds staged in LDS for the block.
__device__ __forceinline__ void stage_thresholds(float* s_thr, const CodecTables* tab) {
  for (uint32_t k = threadIdx.x; k < 257u; k += kBlock) s_thr[k] = tab->thr[k];
  __syncthreads();
}

template <bool kTail>
__global__ __launch_bounds__(kBlock) void k_ll_store(Common c, FragView f, LLBuf ll, TailBuf tb) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, c.tab);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool valid = i < f.count;
  const uint32_t lid = f.id_base + i;
  const uint64_t n = (uint64_t)lid + 1u;
  bool ovf = false;
  float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t t = kEmpty;
  if (valid && c.opq_z) {  // culled by the opaque pass: no node, no tail
    t = ld_stream(&f.target[i]);
    valid = t < c.T && opaque_visible(c.opq_z, c.T, 1u, t, depth_key(ld_stream(&f.depth[i])), 1u) != 0u;
    t = kEmpty;
  }
  if (valid) {
    col = ld_stream(&f.rgba[i]);
    if (n >= ll.cap) {
      t = ld_stream(&f.target[i]);
      ovf = t < c.T;
    } else {
      const uint32_t key = depth_key(ld_stream(&f.depth[i]));
      ll.cd[n] = ((uint64_t)key << 32) | encode_color_lds(col, s_thr);
      ll.next[n] = kUnset;
    }
  }
  const uint64_t om = __ballot(ovf);
  if (om && lane_id() == (uint32_t)__builtin_ctzll(om)) atomicAdd(ll.overflow, (uint32_t)__builtin_popcountll(om));
  if constexpr (kTail) tail_queue_push(ovf, t, lid, 0u, col, tb, 0xFFFFFFFFu);
}

// Link half: the head exchange of oitLinkedList.frag.glsl:72, done as an
// ordered insert so that every list is in descending id (newest first: the
// order a sequential draw produces). Same-target runs in a wave are chained
// locally and inserted with one atomic.
__global__ __launch_bounds__(kBlock) void k_ll_link(Common c, FragView f, LLBuf ll) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint64_t n64 = (uint64_t)f.id_base + i + 1u;
  bool valid = i < f.count && n64 < ll.cap;
  uint32_t t = valid ? ld_stream(&f.target[i]) : kEmpty;
  if (t >= c.T || (c.opq_z && !opaque_visible(c.opq_z, c.T, 1u, t, depth_key(ld_stream(&f.depth[i])), 1u))) {
    valid = false;
    t = kEmpty;
  }
  const Run r = same_target_run(t);
  const uint32_t n = (uint32_t)n64;
  // chain inside the run; the run start lane's link is written by the insert
  if (valid && !r.start) st_relaxed(&ll.next[n], n - 1u);
  const bool ins = valid && r.end;
  list_insert_desc(ins, ins ? &ll.head[pm_index(t, c.T, c.spp)] : nullptr, ll.next, n - (r.lane - r.start_lane), n,
                   (uint32_t)ll.cap, ll.err);
}

// Fused colour pass of the Linked List: store and ordered link in one
// kernel (oitLinkedList.frag.glsl:51-85). Each wave takes kU consecutive
// 64-fragment chunks and issues their loads, then their stores, then their
// head inserts back to back, so kU memory round trips overlap per wave; only
// lanes that find a larger element already at the head walk the list. The
// link words of the frame's nodes already hold kUnset (left by the previous
// frame's resetting resolve, or by the clear), so a walker never follows a
// stale link; the run chain n -> n-1 and the inserts are relaxed atomics.
// One wave's share of the fused append: fragments first .. first + 64 * kU - 1
// of a source stream whose fragment i is node node0 + i and targets head
// tbase + pm_index(target[i]) of the pool's head array.
struct AppendSrc {
  const uint32_t* target;
  const float* depth;
  const float4* rgba;
  uint64_t count;  // fragments in the source
  uint64_t first;  // the wave's first fragment
  uint32_t node0;  // node of fragment 0 (its id + 1)
  uint32_t T;      // targets of the frame
  uint32_t tbase;  // head index of the frame's target 0
  uint32_t lid0;   // frame-local id of fragment 0 (colour-pass tail records)
  const uint32_t* mask = nullptr;  // coverage shading: sample masks (a zero mask drops the fragment)
  uint32_t full = 0;               // the pixel's sample bits
  const uint64_t* opq = nullptr;   // the frame's opaque depth planes (early depth test), or null
  uint32_t cov = 0;                // coverage samples (opaque plane count)
  uint32_t spp = 1;                // sample shading: samples per pixel (heads pixel-major)
};

template <bool kTail, int kU>
__device__ __forceinline__ void ll_append_wave(const AppendSrc& f, const float* s_thr, const LLBuf& ll,
                                               const TailBuf& tb) {
  const uint32_t lane = lane_id();
  uint32_t t[kU], m[kU];
  bool node[kU], ovf[kU];
  float4 col[kU];
  float dep[kU];
  // every field of every chunk is loaded before any is looked at, so the
  // loads share one round trip (a load guarded by a test of the target
  // would wait for it)
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t i = f.first + (uint64_t)u * 64u + lane;
    const bool in = i < f.count;
    t[u] = in ? ld_stream(&f.target[i]) : kEmpty;
    dep[u] = in ? ld_stream(&f.depth[i]) : 0.f;
    col[u] = in ? ld_stream(&f.rgba[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
    m[u] = (in && f.mask) ? ld_stream(&f.mask[i]) : 0xFFFFFFFFu;
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t i = f.first + (uint64_t)u * 64u + lane;
    if (t[u] >= f.T) t[u] = kEmpty;
    if (f.mask && t[u] != kEmpty) {
      m[u] &= f.full;
      if (m[u] == 0u) t[u] = kEmpty;  // no sample covered: the shader does not run
    }
    if (f.opq && t[u] != kEmpty) {  // behind the opaque depth in every covered sample: culled
      m[u] = opaque_visible(f.opq, f.T, f.cov, t[u], depth_key(dep[u]), m[u]);
      if (m[u] == 0u) t[u] = kEmpty;
    }
    const uint64_t n64 = (uint64_t)f.node0 + i;
    node[u] = t[u] != kEmpty && n64 < ll.cap;
    ovf[u] = t[u] != kEmpty && !node[u];
    if (t[u] == kEmpty) col[u] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  uint32_t lo[kU], hi[kU];
  bool ins[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint32_t n = (uint32_t)(f.node0 + f.first + (uint64_t)u * 64u + lane);
    if (node[u]) ll.cd[n] = ((uint64_t)depth_key(dep[u]) << 32) | encode_color_lds(col[u], s_thr);
    if (node[u] && ll.mask) ll.mask[n] = m[u];  // storeMask, oitLinkedList.frag.glsl:79
    if (node[u] && ll.full) ll.full[n] = col[u];
    const Run r = same_target_run(node[u] ? t[u] : kEmpty);
    if (node[u] && !r.start) st_relaxed(&ll.next[n], n - 1u);
    ins[u] = node[u] && r.end;
    hi[u] = n;
    lo[u] = n - (r.lane - r.start_lane);
  }
  // first insert attempt of every chunk: the atomics are in flight together
  uint32_t old[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    old[u] = 0;
    if (ins[u]) old[u] = max_relaxed(&ll.head[f.tbase + pm_index(t[u], f.T, f.spp)], hi[u]);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    if (ins[u] && old[u] < lo[u]) {
      st_relaxed(&ll.next[lo[u]], old[u]);
      ins[u] = false;
    }
  }
  // a larger element was already at the head: continue behind it
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const bool walk = ins[u] && old[u] < ll.cap;
    list_insert_desc_from(walk, walk ? &ll.next[old[u]] : nullptr, ll.next, lo[u], hi[u], (uint32_t)ll.cap,
                          ll.err);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t om = __ballot(ovf[u]);
    if (om && lane == (uint32_t)__builtin_ctzll(om)) atomicAdd(ll.overflow, (uint32_t)__builtin_popcountll(om));
    if constexpr (kTail) {  // records exist only where the pool can overflow (node >= cap)
      const uint64_t i = f.first + (uint64_t)u * 64u + lane;
      const uint32_t lid = (uint32_t)(f.lid0 + i);
      if ((uint64_t)f.node0 + i >= ll.cap) tail_queue_push(ovf[u], t[u], lid, 0u, col[u], tb, m[u]);
    }
  }
}

template <bool kTail, int kU>
__global__ __launch_bounds__(kBlock) void k_ll_append(Common c, FragView f, LLBuf ll, TailBuf tb) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, c.tab);
  AppendSrc a;
  a.target = f.target;
  a.depth = f.depth;
  a.rgba = f.rgba;
  a.count = f.count;
  a.first = ((uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64u * kU);
  a.node0 = f.id_base + 1u;
  a.T = c.T;
  a.tbase = 0;
  a.lid0 = f.id_base;
  a.mask = f.mask;
  a.full = c.cov > 1u ? (1u << c.cov) - 1u : 0xFFFFFFFFu;
  a.opq = c.opq_z;
  a.cov = c.cov;
  a.spp = c.spp;
  ll_append_wave<kTail, kU>(a, s_thr, ll, tb);
}

// Walk of target t's list (t indexes the pool's head array): front K in `a`,
// list length in `cnt`, and the composite of the front K over the evicted
// tail (before the colour buffer). Resets head[t] and every walked link.
// Reached by every lane of the wave (wave_max_u32).
template <int K, bool kTail, bool kPair>
__device__ __forceinline__ float4 ll_walk_core(uint32_t t, const LLBuf& ll, const float* s_dec,
                                               const float* s_alpha, FrontK<K>& a, uint32_t& cnt) {
  a.init();
  float4 tailc = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t p = ll.head[t];
  if (p != 0u) ll.head[t] = 0u;
  cnt = 0;
  // trips of this walk so far; every active lane has the same value, so
  // readfirstlane makes it a scalar bound on the occupied slots (a trip
  // consumes at most two nodes)
  uint32_t trip = 0;
  while (p != 0u && p < ll.cap) {  // a link outside the pool ends the walk
    const int w = (int)__builtin_amdgcn_readfirstlane(kPair ? 2u * trip + 2u : trip + 1u);
    ++trip;
    const uint64_t cd0 = ll.cd[p];
    const uint32_t nx0 = ll.next[p];
    uint64_t cd1 = 0;
    uint32_t nx1 = 0;
    const uint32_t q = p - 1u;
    if (kPair) {
      cd1 = ll.cd[q];
      nx1 = ll.next[q];
    }
    ll.next[p] = kUnset;
    {
      uint64_t ev;
      uint32_t evc;
      a.insert_bounded((cd0 & 0xFFFFFFFF00000000ull) | p, (uint32_t)cd0, ev, evc, w);
      if (kTail && cnt >= (uint32_t)K) tailc = over(tailc, decode_premult_lds(evc, s_dec, s_alpha));
      ++cnt;
    }
    if (kPair && nx0 == q && q != 0u) {
      ll.next[q] = kUnset;
      uint64_t ev;
      uint32_t evc;
      a.insert_bounded((cd1 & 0xFFFFFFFF00000000ull) | q, (uint32_t)cd1, ev, evc, w);
      if (kTail && cnt >= (uint32_t)K) tailc = over(tailc, decode_premult_lds(evc, s_dec, s_alpha));
      ++cnt;
      p = descend(nx1, q, ll.err);
    } else {
      p = descend(nx0, p, ll.err);
    }
  }
  return tailc;
}

// ll_walk_core, then the front K blended over the tail colour.
template <int K, bool kTail, bool kPair>
__device__ __forceinline__ float4 ll_walk_reset(uint32_t t, const LLBuf& ll, const float* s_dec,
                                                const float* s_alpha, FrontK<K>& a, uint32_t& cnt) {
  const float4 tailc = ll_walk_core<K, kTail, kPair>(t, ll, s_dec, s_alpha, a, cnt);
  const uint32_t n = cnt < (uint32_t)K ? cnt : (uint32_t)K;
  const uint32_t nmax = wave_max_u32(n);
  float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((uint32_t)k < nmax && (uint32_t)k < n) col = over(col, decode_premult_lds(a.col[k], s_dec, s_alpha));
  return over(col, tailc);
}

// Key-only form of ll_walk_reset: the walk keeps (depth key, node) pairs and
// the front-K colours are gathered by node index at the end.
template <int K, bool kTail, bool kPair>
__device__ __forceinline__ float4 ll_walk_reset_keys(uint32_t t, const LLBuf& ll, const float* s_dec,
                                                     const float* s_alpha, FrontKeys<K>& a, uint32_t& cnt) {
  a.init();
  float4 tailc = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t p = ll.head[t];
  if (p != 0u) ll.head[t] = 0u;
  cnt = 0;
  uint32_t trip = 0;
  while (p != 0u && p < ll.cap) {  // a link outside the pool ends the walk
    const int w = (int)__builtin_amdgcn_readfirstlane(kPair ? 2u * trip + 2u : trip + 1u);
    ++trip;
    const uint32_t key0 = (uint32_t)(ll.cd[p] >> 32);
    const uint32_t nx0 = ll.next[p];
    uint32_t key1 = 0;
    uint32_t nx1 = 0;
    const uint32_t q = p - 1u;
    if (kPair) {
      key1 = (uint32_t)(ll.cd[q] >> 32);
      nx1 = ll.next[q];
    }
    ll.next[p] = kUnset;
    {
      const uint64_t ev = a.insert_bounded(((uint64_t)key0 << 32) | p, w);
      if (kTail && cnt >= (uint32_t)K)
        tailc = over(tailc, decode_premult_lds((uint32_t)ll.cd[(uint32_t)ev], s_dec, s_alpha));
      ++cnt;
    }
    if (kPair && nx0 == q && q != 0u) {
      ll.next[q] = kUnset;
      const uint64_t ev = a.insert_bounded(((uint64_t)key1 << 32) | q, w);
      if (kTail && cnt >= (uint32_t)K)
        tailc = over(tailc, decode_premult_lds((uint32_t)ll.cd[(uint32_t)ev], s_dec, s_alpha));
      ++cnt;
      p = descend(nx1, q, ll.err);
    } else {
      p = descend(nx0, p, ll.err);
    }
  }
  const uint32_t n = cnt < (uint32_t)K ? cnt : (uint32_t)K;
  const uint32_t nmax = wave_max_u32(n);
  uint32_t colr[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    colr[k] = 0u;
    if ((uint32_t)k < nmax && (uint32_t)k < n) colr[k] = (uint32_t)ll.cd[(uint32_t)a.sk[k]];
  }
  float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((uint32_t)k < nmax && (uint32_t)k < n) col = over(col, decode_premult_lds(colr[k], s_dec, s_alpha));
  return over(col, tailc);
}

__device__ __forceinline__ void stage_decode(float* s_dec, float* s_alpha, const CodecTables* tab) {
  for (uint32_t k = threadIdx.x; k < 256u; k += kBlock) {
    s_dec[k] = tab->dec[k];
    s_alpha[k] = tab->alpha[k];
  }
}

// Composite of the Linked List (oitLinkedList.frag.glsl:106-172) that leaves
// the structure clean for the next frame: head[t] = 0 (the clear of
// oitRender.cpp:206-212 done by the reader) and next[p] = kUnset for every
// walked node. With kPair, the loads for node p-1 go out with those for p:
// consecutive fragments of one target (a sphere's front and back) are
// consecutive nodes chained p -> p-1, so one round trip consumes both.
template <int K, bool kTail, bool kColorTail, bool kDebug, bool kPair>
// At K = 16 without a colour-pass tail the front K and the walk take 81
// VGPRs (five waves per SIMD); the bound holds them to 80 without spills,
// for six waves (with the tail fold it would spill).
__global__ __launch_bounds__(kBlock, (K == 16 && !kColorTail) ? 6 : 1) void k_ll_resolve_reset(Common c, LLBuf ll,
                                                                                               TailBuf tb,
                                                                                               ResolveOut o) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  if (blockIdx.x == 0 && threadIdx.x < 16u) {  // frame counters -> last-frame slot, zero for the next frame
    ll.counters[16 + threadIdx.x] = ll.counters[threadIdx.x];
    ll.counters[threadIdx.x] = 0u;
  }
  __syncthreads();
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < c.T; i += stride) {
    const uint32_t t = pm_target(i, c.T, c.spp);
    FrontK<K> a;
    uint32_t cnt;
    const float4 col = ll_walk_reset<K, kTail, kPair>(i, ll, s_dec, s_alpha, a, cnt);
    float4 d = dest_of(c, t);
    if (kColorTail) d = tail_fold(tb, t, d, 0u);
    o.out[t] = over(col, d);
    if (kDebug) write_debug<K>(o, t, a, cnt < (uint32_t)K ? cnt : (uint32_t)K, cnt, true, false);
  }
}

template <int K, bool kTail, bool kColorTail, bool kDebug, bool kPair>
__global__ __launch_bounds__(kBlock) void k_ll_resolve_reset_keys(Common c, LLBuf ll, TailBuf tb, ResolveOut o) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  if (blockIdx.x == 0 && threadIdx.x < 16u) {  // frame counters -> last-frame slot, zero for the next frame
    ll.counters[16 + threadIdx.x] = ll.counters[threadIdx.x];
    ll.counters[threadIdx.x] = 0u;
  }
  __syncthreads();
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < c.T; i += stride) {
    const uint32_t t = pm_target(i, c.T, c.spp);
    FrontKeys<K> a;
    uint32_t cnt;
    const float4 col = ll_walk_reset_keys<K, kTail, kPair>(i, ll, s_dec, s_alpha, a, cnt);
    float4 d = dest_of(c, t);
    if (kColorTail) d = tail_fold(tb, t, d, 0u);
    o.out[t] = over(col, d);
    if (kDebug) {
      const uint32_t n = cnt < (uint32_t)K ? cnt : (uint32_t)K;
      if (o.fk_key) {
#pragma unroll
        for (int k = 0; k < K; ++k)
          o.fk_key[(uint64_t)t * K + k] = (uint32_t)k < n ? (uint32_t)(a.sk[k] >> 32) : kEmpty;
      }
      if (o.fk_val) {
#pragma unroll
        for (int k = 0; k < K; ++k) o.fk_val[(uint64_t)t * K + k] = (uint32_t)k < n ? (uint32_t)a.sk[k] - 1u : kEmpty;
      }
      if (o.counts) o.counts[t] = cnt;
    }
  }
}

// Coverage-shaded composite of the Linked List (OIT_COVERAGE_SHADING,
// oitLinkedList.frag.glsl:147-160): the walk and tail as above; the front K
// blended per sample with their node masks; the colour buffer averaged over
// samples. Resets the structure like k_ll_resolve_reset.
template <int K, bool kTail>
__global__ __launch_bounds__(kBlock) void k_ll_resolve_cov(Common c, LLBuf ll, TailBuf tb, ResolveOut o) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  if (blockIdx.x == 0 && threadIdx.x < 16u) {  // frame counters -> last-frame slot, zero for the next frame
    ll.counters[16 + threadIdx.x] = ll.counters[threadIdx.x];
    ll.counters[threadIdx.x] = 0u;
  }
  __syncthreads();
  const uint32_t stride = gridDim.x * kBlock;
  // coverage shading has one target per pixel (spp = 1): head index = target
  for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < c.T; t += stride) {
    FrontK<K> a;
    uint32_t cnt;
    const float4 tailc = ll_walk_core<K, kTail, true>(t, ll, s_dec, s_alpha, a, cnt);
    const uint32_t n = cnt < (uint32_t)K ? cnt : (uint32_t)K;
    uint32_t masks[K];
#pragma unroll
    for (int k = 0; k < K; ++k) masks[k] = (uint32_t)k < n ? ll.mask[(uint32_t)a.sk[k]] : 0u;
    const float4 col = over(blend_front_cov<K>(a.col, masks, n, c.cov, s_dec, s_alpha), tailc);
    const float4 d = tail_fold_cov(tb, t, c);
    o.out[t] = over(col, d);
    write_debug<K>(o, t, a, n, cnt, true, false);
  }
}

// ---------------------------------------------------------------------------
// Batched Linked List: many frames in one pool. Frame f's fragments are pool
// nodes frag_base + 1 .. frag_base + count, its targets are head entries
// target_base .. target_base + T - 1; both bases are padded so that no append
// unit (64 * kU fragments) and no resolve tile (kBlock targets) spans two
// frames. Each frame gets exactly the single-frame result.
// ---------------------------------------------------------------------------

// Last frame in [lo, hi) whose base is <= key (bases ascend); key is wave
// uniform, so the probes are scalar loads.
template <bool kTargets>
__device__ __forceinline__ uint32_t find_frame(const BatchFrameDev* fr, uint32_t lo, uint32_t hi, uint64_t key) {
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint64_t base = kTargets ? (uint64_t)fr[mid].target_base : (uint64_t)fr[mid].frag_base;
    if (base <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Binned batch frames (BatchFrameDev::mode == kModeBin): the fragments of
// each target are gathered into one bucket of the pool by a counting sort
// (count per target, exclusive scan, scatter) instead of being linked into a
// list; buckets of more than kBinShallow entries are sorted in place by
// fragment id, newest first (the walk order of the Linked List's descending
// lists), by one wave each (a bitonic sort in registers up to 64 entries,
// in LDS up to kBinDeepMax); the composite replays each bucket newest first
// through the Linked List composite (oitLinkedList.frag.glsl:116-172), one
// lane per target, so the image is the single-frame Linked List's bit for
// bit. A sorted bucket is read with independent loads, where a list walk is
// a chain of dependent ones. Frame f's buckets fill its own node range:
// during a run head[t] holds the count of target t, the cursor of t (kept
// at bin_cursor(t): consecutive targets' cursors in different cache lines,
// so a wave's atomics on them do not queue on one line) its bucket's first
// slot before the scatter and its end after, and bucket entry s is node
// s + 1 (node 0 stays the terminator) with cd = (depth key << 32) | colour
// and next = fragment id + 1. The composite returns head[t] to 0 and the
// bucket's link words to kUnset, the pool's clean state, so a later run may
// take the frame either way.
// ---------------------------------------------------------------------------

constexpr uint32_t kBinShallow = 8;      // buckets a lane orders itself, in registers
constexpr uint32_t kBinDeepMax = 1024;   // larger ones sorted by a wave; beyond, the lane selects by id

// Position of target t's cursor: the 64 targets of a block of 64 spread over
// 64 regions of cap / 64 entries (cap a multiple of 64).
__device__ __forceinline__ uint32_t bin_cursor(uint32_t t, uint32_t cap) {
  return (t & 63u) * (cap >> 6) + (t >> 6);
}

// Count pass of one wave's share (the binned frame's form of the append):
// returnless atomics, one per run of one target.
template <int kU>
__device__ __forceinline__ void bin_count_wave(const BatchFrameDev& fr, uint64_t first, const LLBuf& ll) {
  const uint32_t lane = lane_id();
  uint32_t t[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    t[u] = i < fr.count ? ld_stream(&fr.target[i]) : kEmpty;
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint32_t tt = t[u] < fr.T ? t[u] : kEmpty;
    const Run r = same_target_run(tt);
    if (tt != kEmpty && r.end)
      __hip_atomic_fetch_add(&ll.head[fr.target_base + tt], r.lane - r.start_lane + 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Block-wide exclusive scan of one u32 per thread (kThreads a multiple of
// 64, at most 1024); returns the block total.
template <int kThreads>
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t& excl, uint32_t* s_part) {
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, (unsigned)o);
    if (lane >= (uint32_t)o) inc += y;
  }
  if (lane == 63u) s_part[w] = inc;
  __syncthreads();
  constexpr int kW = kThreads / 64;
  uint32_t before = 0, total = 0;
#pragma unroll
  for (int k = 0; k < kW; ++k) {
    const uint32_t p = s_part[k];
    if ((uint32_t)k < w) before += p;
    total += p;
  }
  __syncthreads();  // s_part may be rewritten by the next call
  excl = before + inc - v;
  return total;
}

// Scan step 1: the sum of each 256-target tile's counts (0 for a Linked
// List frame's tiles). Block 0 also empties the queue of large buckets.
__global__ __launch_bounds__(kBlock) void k_bin_tile_sums(BatchArgs b, const uint32_t* __restrict__ count) {
  __shared__ uint32_t s_part[kBlock / 64];
  const uint32_t tile = b.tile0 + blockIdx.x;
  const uint32_t f = find_frame<true>(b.frames, b.f0, b.f1, (uint64_t)tile * kBlock);
  const bool bin = b.frames[f].mode == kModeBin;
  const uint32_t v = bin ? count[(uint64_t)tile * kBlock + threadIdx.x] : 0u;
  uint32_t excl;
  const uint32_t total = block_scan_u32<kBlock>(v, excl, s_part);
  if (threadIdx.x == 0) b.tile_sum[tile] = total;
  if (blockIdx.x == 0 && threadIdx.x == 0) *b.deep_n = 0u;
}

// Scan step 2, one block per frame: each tile's first bucket slot, the
// frame's node range start plus the counts of its earlier tiles.
__global__ __launch_bounds__(1024) void k_bin_frame_scan(BatchArgs b) {
  __shared__ uint32_t s_part[16];
  const uint32_t f = b.f0 + blockIdx.x;
  if (f >= b.f1) return;
  const BatchFrameDev& fr = b.frames[f];
  if (fr.mode != kModeBin) return;
  const uint32_t t0 = fr.target_base / kBlock;
  const uint32_t nt = (fr.T + kBlock - 1u) / kBlock;
  uint32_t carry = fr.frag_base;
  for (uint32_t base = 0; base < nt; base += 1024u) {  // the same trip count on every thread
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nt ? b.tile_sum[t0 + i] : 0u;
    uint32_t excl;
    const uint32_t total = block_scan_u32<1024>(v, excl, s_part);
    if (i < nt) b.tile_sum[t0 + i] = carry + excl;
    carry += total;
  }
}

// Scan step 3: each target's cursor = the first slot of its bucket; targets
// with more than kBinShallow entries are queued for k_bin_sort.
__global__ __launch_bounds__(kBlock) void k_bin_tile_apply(BatchArgs b, const uint32_t* __restrict__ count) {
  __shared__ uint32_t s_part[kBlock / 64];
  const uint32_t tile = b.tile0 + blockIdx.x;
  const uint32_t f = find_frame<true>(b.frames, b.f0, b.f1, (uint64_t)tile * kBlock);
  if (b.frames[f].mode != kModeBin) return;  // the whole block takes the same branch
  const uint32_t t = tile * kBlock + threadIdx.x;
  const uint32_t n = count[t];
  uint32_t excl;
  (void)block_scan_u32<kBlock>(n, excl, s_part);
  b.cursor[bin_cursor(t, b.cursor_cap)] = b.tile_sum[tile] + excl;
  // queue entries of the wave from one atomic
  const bool deep = n > kBinShallow;
  const uint64_t m = __ballot(deep);
  uint32_t q = 0;
  if (lane_id() == 0u && m) q = atomicAdd(b.deep_n, (uint32_t)__builtin_popcountll(m));
  q = __shfl(q, 0);
  if (deep) b.deep_q[q + lanes_below(m)] = t;
}

// Scatter pass: each fragment of a binned frame to the next entry of its
// target's bucket (one atomic per run of one target), as a node: encoded
// (depth key, colour) and its id + 1.
template <int kU>
__global__ __launch_bounds__(kBlock) void k_bin_scatter(BatchArgs b, LLBuf ll) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, b.tab);
  const uint32_t unit = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
  if (unit >= b.units) return;
  const uint64_t g = (uint64_t)unit * (64u * kU);
  const uint32_t f = find_frame<false>(b.frames, b.f0, b.f1, g);
  const BatchFrameDev fr = b.frames[f];
  if (fr.mode != kModeBin) return;
  const uint32_t lane = lane_id();
  const uint64_t first = g - fr.frag_base;
  uint32_t t[kU];
  float dep[kU];
  float4 col[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {  // every field of every chunk in one round trip
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    const bool in = i < fr.count;
    t[u] = in ? ld_stream(&fr.target[i]) : kEmpty;
    dep[u] = in ? ld_stream(&fr.depth[i]) : 0.f;
    col[u] = in ? ld_stream(&fr.rgba[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  uint32_t base[kU];
  Run r[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {  // every chunk's cursor atomics in flight together
    const uint32_t tt = t[u] < fr.T ? t[u] : kEmpty;
    t[u] = tt;
    r[u] = same_target_run(tt);
    base[u] = 0;
    if (tt != kEmpty && r[u].end)
      base[u] = __hip_atomic_fetch_add(&b.cursor[bin_cursor(fr.target_base + tt, b.cursor_cap)],
                                       r[u].lane - r[u].start_lane + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    const uint32_t bs = __shfl(base[u], (int)r[u].end_lane);  // the run's first slot, from its last lane
    if (t[u] != kEmpty) {
      const uint32_t s = bs + (r[u].lane - r[u].start_lane) + 1u;  // node index
      ll.cd[s] = ((uint64_t)depth_key(dep[u]) << 32) | encode_color_lds(col[u], s_thr);
      ll.next[s] = (uint32_t)i + 1u;
    }
  }
}

// Orders LDS ops of one wave: every lane's earlier LDS access has completed
// and none is moved across (the wave's lanes exchange data through LDS).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Every global load the lane issued has returned, and no memory access is
// moved across (the lanes of a wave then rewrite what the others read).
__device__ __forceinline__ void loads_done() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Lane-wise 64-bit shuffle built from 32-bit ones.
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  const uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

// The queued buckets (more than kBinShallow entries), one wave each, sorted
// in place by id, newest first: up to 64 entries by a bitonic network across
// the lanes (one entry per lane, the colour word carried along), up to
// kBinDeepMax through LDS (id and entry index sorted, then the colour words
// gathered in that order); every entry is read before any is rewritten.
// Larger buckets stay as they are (the composite selects by id).
constexpr uint32_t kBinSortWaves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void k_bin_sort(BatchArgs b, LLBuf ll) {
  __shared__ uint64_t s_key[kBinSortWaves][kBinDeepMax];  // (id + 1) << 32 | entry
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  uint64_t* key = s_key[w];
  const uint32_t total = __builtin_amdgcn_readfirstlane(*b.deep_n);
  for (uint32_t e = blockIdx.x * kBinSortWaves + w; e < total; e += gridDim.x * kBinSortWaves) {
    const uint32_t t = __builtin_amdgcn_readfirstlane(b.deep_q[e]);
    const uint32_t n = __builtin_amdgcn_readfirstlane(ll.head[t]);
    const uint32_t start = __builtin_amdgcn_readfirstlane(b.cursor[bin_cursor(t, b.cursor_cap)]) - n;
    uint32_t* nx = ll.next + start + 1u;
    uint64_t* cd = ll.cd + start + 1u;
    if (n <= 64u) {
      uint64_t k = lane < n ? ((uint64_t)nx[lane] << 32) | lane : 0ull;
      uint64_t c = lane < n ? cd[lane] : 0ull;
      for (uint32_t sz = 2u; sz <= 64u; sz <<= 1) {  // bitonic across the lanes, descending
        for (uint32_t j = sz >> 1; j > 0u; j >>= 1) {
          const uint64_t ok = shfl_xor_u64(k, (int)j);
          const uint64_t oc = shfl_xor_u64(c, (int)j);
          const bool lower = (lane & j) == 0u;           // this lane holds the pair's first position
          const bool desc = (lane & sz) == 0u;           // this pair's block runs descending
          const bool take = lower == desc ? ok > k : ok < k;
          if (take) {
            k = ok;
            c = oc;
          }
        }
      }
      loads_done();
      if (lane < n) {
        nx[lane] = (uint32_t)(k >> 32);
        cd[lane] = c;
      }
    } else if (n <= kBinDeepMax) {
      uint32_t N = 128u;
      while (N < n) N <<= 1;
      for (uint32_t i = lane; i < N; i += 64u) key[i] = i < n ? ((uint64_t)nx[i] << 32) | i : 0ull;
      wave_lds_sync();
      for (uint32_t sz = 2u; sz <= N; sz <<= 1) {  // bitonic sort, descending
        for (uint32_t j = sz >> 1; j > 0u; j >>= 1) {
          for (uint32_t i = lane; i < N; i += 64u) {
            const uint32_t l = i ^ j;
            if (l > i) {
              const uint64_t x = key[i], y = key[l];
              const bool swap = (i & sz) == 0u ? x < y : x > y;
              if (swap) {
                key[i] = y;
                key[l] = x;
              }
            }
          }
          wave_lds_sync();
        }
      }
      constexpr uint32_t kPer = kBinDeepMax / 64u;  // sorted entries per lane
      uint64_t c[kPer];
#pragma unroll
      for (uint32_t m = 0; m < kPer; ++m) {
        const uint32_t j = lane + 64u * m;
        c[m] = j < n ? cd[(uint32_t)key[j]] : 0ull;
      }
      loads_done();
#pragma unroll
      for (uint32_t m = 0; m < kPer; ++m) {
        const uint32_t j = lane + 64u * m;
        if (j < n) {
          nx[j] = (uint32_t)(key[j] >> 32);
          cd[j] = c[m];
        }
      }
      wave_lds_sync();  // the next bucket rewrites key[]
    }
  }
}

// Composite of target t's bucket (n entries, nodes start + 1 .. start + n)
// by one lane: the entries newest first through the Linked List's front-K
// insertion and composite tail (ll_walk_core), then the front K blended
// over the tail; resets the bucket. Up to kBinShallow entries are ordered
// here in registers; up to kBinDeepMax k_bin_sort has ordered them; beyond,
// each step selects the newest entry not replayed yet. Lane-local only.
template <int K, bool kTail>
__device__ __forceinline__ float4 bin_replay(uint32_t t, uint32_t n, uint32_t start, const LLBuf& ll,
                                             const float* s_dec, const float* s_alpha) {
  const uint32_t* nx = ll.next + start + 1u;
  const uint64_t* cd = ll.cd + start + 1u;
  FrontK<K> a;
  a.init();
  float4 tailc = make_float4(0.f, 0.f, 0.f, 0.f);
  auto take = [&](uint32_t s, uint32_t id1, uint64_t kc) {
    uint64_t ev;
    uint32_t evc;
    a.insert((kc & 0xFFFFFFFF00000000ull) | id1, (uint32_t)kc, ev, evc);
    if (kTail && s >= (uint32_t)K) tailc = over(tailc, decode_premult_lds(evc, s_dec, s_alpha));
  };
  if (n <= kBinShallow) {
    uint32_t ids[kBinShallow];
    uint64_t kcs[kBinShallow];
#pragma unroll
    for (int j = 0; j < (int)kBinShallow; ++j) {
      ids[j] = 0u;
      kcs[j] = 0ull;
      if ((uint32_t)j < n) {
        ids[j] = nx[j];
        kcs[j] = cd[j];
      }
    }
    uint32_t prev = 0xFFFFFFFFu;
#pragma unroll
    for (int s = 0; s < (int)kBinShallow; ++s) {
      if ((uint32_t)s < n) {
        uint32_t best = 0u;
        uint64_t bkc = 0ull;
#pragma unroll
        for (int j = 0; j < (int)kBinShallow; ++j) {  // the newest entry not replayed yet
          const bool sel = ids[j] < prev && ids[j] > best;
          best = sel ? ids[j] : best;
          bkc = sel ? kcs[j] : bkc;
        }
        prev = best;
        take((uint32_t)s, best, bkc);
      }
    }
  } else if (n <= kBinDeepMax) {
    uint32_t s = 0;
    for (; s + 4u <= n; s += 4u) {  // sorted: four independent entries per trip
      uint32_t id1[4];
      uint64_t kc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        id1[q] = nx[s + q];
        kc[q] = cd[s + q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) take(s + q, id1[q], kc[q]);
    }
    for (; s < n; ++s) take(s, nx[s], cd[s]);
  } else {
    uint32_t prev = 0xFFFFFFFFu;
    for (uint32_t s = 0; s < n; ++s) {
      uint32_t best = 0u, bi = 0u;
      for (uint32_t j = 0; j < n; ++j) {
        const uint32_t id1 = nx[j];
        if (id1 < prev && id1 > best) {
          best = id1;
          bi = j;
        }
      }
      prev = best;
      take(s, best, cd[bi]);
    }
  }
  for (uint32_t j = 0; j < n; ++j) ll.next[start + 1u + j] = kUnset;
  if (n) ll.head[t] = 0u;
  const uint32_t m = n < (uint32_t)K ? n : (uint32_t)K;
  float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((uint32_t)k < m) col = over(col, decode_premult_lds(a.col[k], s_dec, s_alpha));
  return over(col, tailc);
}

template <int kU>
__global__ __launch_bounds__(kBlock) void k_batch_append(BatchArgs b, LLBuf ll) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, b.tab);
  const uint32_t unit = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
  if (unit >= b.units) return;
  uint64_t g;
  uint32_t f;
  if (b.sched) {  // interleaved: every frame advances in id order, a share of the waves in flight each
    constexpr uint32_t per = kBatchSlice / (64u * kU);  // waves per slice
    const uint32_t slice = unit / per;
    uint32_t lo = 0, hi = b.rounds;  // last round whose first slice is <= slice
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (b.sched[mid] <= slice) lo = mid;
      else hi = mid;
    }
    const uint32_t q = lo;
    const uint32_t o = slice - b.sched[q];
    const uint32_t* pre = b.sched + b.rounds + 1u + (uint64_t)q * (b.f1 + 1u);
    uint32_t a0 = 0, a1 = b.f1;  // last frame whose offset in the round is <= o
    while (a1 - a0 > 1u) {
      const uint32_t mid = (a0 + a1) >> 1;
      if (pre[mid] <= o) a0 = mid;
      else a1 = mid;
    }
    f = a0;
    const uint64_t slices = ((uint64_t)b.frames[f].count + kBatchSlice - 1u) / kBatchSlice;
    const uint64_t c = slices * q / b.rounds + (o - pre[f]);
    g = (uint64_t)b.frames[f].frag_base + c * kBatchSlice + (uint64_t)(unit % per) * (64u * kU);
  } else {
    g = (uint64_t)unit * (64u * kU);
    f = find_frame<false>(b.frames, b.f0, b.f1, g);
  }
  const BatchFrameDev fr = b.frames[f];
  if (fr.mode == kModeBin) {  // binned frame: count its targets' fragments
    bin_count_wave<kU>(fr, g - fr.frag_base, ll);
    return;
  }
  AppendSrc a;
  a.target = fr.target;
  a.depth = fr.depth;
  a.rgba = fr.rgba;
  a.count = fr.count;
  a.first = g - fr.frag_base;
  a.node0 = fr.frag_base + 1u;
  a.T = fr.T;
  a.tbase = fr.target_base;
  a.lid0 = 0;
  ll_append_wave<false, kU>(a, s_thr, ll, TailBuf{});
}

// Composite of a group's binned frames (k_batch_resolve takes its linked
// ones): one lane per target, bin_replay.
template <int K, bool kTail>
__global__ __launch_bounds__(kBlock) void k_bin_resolve(BatchArgs b, LLBuf ll) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, b.tab);
  __syncthreads();
  for (uint32_t tile = b.tile0 + blockIdx.x; tile < b.tile0 + b.tiles; tile += gridDim.x) {
    const uint32_t f = find_frame<true>(b.frames, b.f0, b.f1, (uint64_t)tile * kBlock);
    const BatchFrameDev& fr = b.frames[f];
    if (fr.mode != kModeBin) continue;  // the whole tile is one frame's
    const uint32_t t = tile * kBlock + threadIdx.x;
    const uint32_t n = ll.head[t];
    const uint32_t start = n ? b.cursor[bin_cursor(t, b.cursor_cap)] - n : 0u;
    const float4 bc = bin_replay<K, kTail>(t, n, start, ll, s_dec, s_alpha);
    const uint32_t local = t - fr.target_base;
    if (local < fr.T) fr.out[local] = over(bc, fr.background);
  }
}

template <int K, bool kTail, bool kPair, bool kKeys>
__global__ __launch_bounds__(kBlock) void k_batch_resolve(BatchArgs b, LLBuf ll) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, b.tab);
  __syncthreads();
  for (uint32_t tile = b.tile0 + blockIdx.x; tile < b.tile0 + b.tiles; tile += gridDim.x) {
    const uint32_t f = find_frame<true>(b.frames, b.f0, b.f1, (uint64_t)tile * kBlock);
    const BatchFrameDev& fr = b.frames[f];
    const uint32_t t = tile * kBlock + threadIdx.x;
    if (fr.mode == kModeBin) continue;  // k_bin_resolve's (the whole tile is one frame's)
    uint32_t cnt;
    float4 col;
    if constexpr (kKeys) {
      FrontKeys<K> a;
      col = ll_walk_reset_keys<K, kTail, kPair>(t, ll, s_dec, s_alpha, a, cnt);
    } else {
      FrontK<K> a;
      col = ll_walk_reset<K, kTail, kPair>(t, ll, s_dec, s_alpha, a, cnt);
    }
    const uint32_t local = t - fr.target_base;
    if (local < fr.T) fr.out[local] = over(col, fr.background);
  }
}

// ---------------------------------------------------------------------------
// Simple: oitSimple.frag.glsl
// ---------------------------------------------------------------------------

// Colour pass (oitSimple.frag.glsl:50-91): slot = atomicAdd(count), slot < K
// stores (colour, key, id) in layer `slot`, else the fragment tail-blends.
// A run of same-target lanes takes its slots with one atomicAdd.
template <bool kTail>
__global__ __launch_bounds__(kBlock) void k_simple_store(Common c, FragView f, KBuf kb, TailBuf tb) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, c.tab);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool valid = i < f.count;
  // the fragment's fields in one round trip, before any is tested
  uint32_t t = valid ? ld_stream(&f.target[i]) : kEmpty;
  const float dv = valid ? ld_stream(&f.depth[i]) : 0.f;
  float4 col = valid ? ld_stream(&f.rgba[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t m = (valid && f.mask) ? ld_stream(&f.mask[i]) : 0xFFFFFFFFu;
  if (t >= c.T) {
    valid = false;
    t = kEmpty;
  }
  if (valid && f.mask) {
    m &= (1u << c.cov) - 1u;
    if (m == 0u) {  // no sample covered: the shader does not run
      valid = false;
      t = kEmpty;
    }
  }
  if (valid && c.opq_z) {  // early depth test against the opaque pass
    m = opaque_visible(c.opq_z, c.T, c.cov, t, depth_key(dv), m);
    if (m == 0u) {
      valid = false;
      t = kEmpty;
    }
  }
  const uint32_t lid = f.id_base + i;
  const Run r = same_target_run(t);
  uint32_t base = 0;
  if (valid && r.start) base = atomicAdd(&kb.count[t], r.end_lane - r.lane + 1u);
  base = __shfl(base, (int)r.start_lane);
  const uint32_t slot = base + (r.lane - r.start_lane);
  bool tail = false;
  if (valid) {
    if (slot < c.K) {
      const uint64_t idx = (uint64_t)slot * c.T + t;
      kb.col[idx] = encode_color_lds(col, s_thr);
      kb.key[idx] = depth_key(dv);
      kb.id[idx] = lid;
      if (kb.mask) kb.mask[idx] = m;  // storeMask, oitSimple.frag.glsl:70
    } else {
      tail = true;
    }
  }
  if constexpr (kTail) tail_queue_push(tail, t, lid, 0u, col, tb, m);
}

// ---------------------------------------------------------------------------
// Spinlock: oitSpinlock.frag.glsl
// ---------------------------------------------------------------------------

// The word at byte offset `off` of a wave-uniform base.
__device__ __forceinline__ uint32_t* byte_at(uint32_t* base, uint64_t off) {
  return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + off);
}

// The critical section of oitSpinlock.frag.glsl:95-123, run by the lane
// that holds target t's lock on a full buffer, for the fragment a and then
// for the fragment b (a pair; either may be absent): each replaces the
// furthest element (first index wins, :96-107) when it is nearer (:109-117).
// One scan serves both sections: it finds the furthest element and the
// furthest of the others, so after a replaces the furthest, the furthest is
// a or that second one, and b's section needs no scan of its own. Each
// section sets its tail output: the evicted element, or the fragment itself.
template <int K>
__device__ __forceinline__ void spin_replace(const Common& c, const KBuf& kb, uint32_t t, bool run_a, uint32_t a_key,
                                             uint32_t a_col, uint32_t a_id, uint32_t a_mask, bool& a_tail,
                                             uint32_t& a_tail_id, uint32_t& a_tail_col, uint32_t& a_tail_mask,
                                             bool run_b, uint32_t b_key, uint32_t b_col, uint32_t b_id,
                                             uint32_t b_mask, bool& b_tail, uint32_t& b_tail_id, uint32_t& b_tail_col,
                                             uint32_t& b_tail_mask) {
  uint32_t fur = 0, sec = 0;
  uint64_t mx = 0, mx2 = 0;
  // the scan, a batch of up to eight slots at a time: a batch's loads are
  // issued together and compared after (branch-free), since a compare
  // between two loads made every slot a round trip of its own. At K = 16
  // sixteen slots in one batch held 70 VGPRs (seven waves per SIMD), two
  // batches 62 (eight), which ran faster
  constexpr int kBatch = K < 16 ? K : 8;
  // the lane's byte offset, made opaque here so that each address is formed
  // at its load (a scalar row base plus this offset) rather than hoisted out
  // of the lock loop and held, one 64-bit address per slot
  uint64_t off = (uint64_t)t * 4u;
  asm volatile("" : "+v"(off));
#pragma unroll
  for (int j0 = 0; j0 < K; j0 += kBatch) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);  // a compiler barrier between batches
    uint32_t kk[kBatch], ii[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {  // a wave-uniform row base and the lane's byte offset
      const uint64_t row = (uint64_t)(j0 + u) * c.T;
      kk[u] = ld_relaxed(byte_at(kb.key + row, off));
      ii[u] = ld_relaxed(byte_at(kb.id + row, off));
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const uint64_t v = ((uint64_t)kk[u] << 32) | ii[u];
      const uint32_t j = (uint32_t)(j0 + u);
      const bool top = v > mx, second = !top && v > mx2;
      mx2 = top ? mx : (second ? v : mx2);
      sec = top ? fur : (second ? j : sec);
      mx = top ? v : mx;
      fur = top ? j : fur;
    }
  }
  // the colours (and masks) the sections can evict, in one round trip
  const uint64_t i_fur = (uint64_t)fur * c.T + t, i_sec = (uint64_t)sec * c.T + t;
  const uint32_t col_fur = ld_relaxed(&kb.col[i_fur]), col_sec = ld_relaxed(&kb.col[i_sec]);
  uint32_t m_fur = 0u, m_sec = 0u;
  if (kb.mask) {
    m_fur = ld_relaxed(&kb.mask[i_fur]);
    m_sec = ld_relaxed(&kb.mask[i_sec]);
  }
  // the furthest element now: value, slot, colour, mask
  uint64_t cur = mx, i_cur = i_fur;
  uint32_t c_cur = col_fur, m_cur = m_fur;
  uint32_t bound = 0u;
  bool moved = false;
  if (run_a) {
    a_tail = true;
    const uint64_t me = ((uint64_t)a_key << 32) | a_id;
    if (mx > me) {
      a_tail_id = (uint32_t)mx;
      a_tail_col = col_fur;
      if (kb.mask) {
        a_tail_mask = m_fur;
        st_relaxed(&kb.mask[i_fur], a_mask);
      }
      st_relaxed(&kb.col[i_fur], a_col);
      st_relaxed(&kb.key[i_fur], a_key);
      st_relaxed(&kb.id[i_fur], a_id);
      bound = (uint32_t)(mx >> 32);
      moved = true;
      if (me > mx2) {  // a is now the furthest
        cur = me;
        c_cur = a_col;
        m_cur = a_mask;
      } else {
        cur = mx2;
        i_cur = i_sec;
        c_cur = col_sec;
        m_cur = m_sec;
      }
    }
  }
  if (run_b) {
    b_tail = true;
    const uint64_t me = ((uint64_t)b_key << 32) | b_id;
    if (cur > me) {
      b_tail_id = (uint32_t)cur;
      b_tail_col = c_cur;
      if (kb.mask) {
        b_tail_mask = m_cur;
        st_relaxed(&kb.mask[i_cur], b_mask);
      }
      st_relaxed(&kb.col[i_cur], b_col);
      st_relaxed(&kb.key[i_cur], b_key);
      st_relaxed(&kb.id[i_cur], b_id);
      bound = (uint32_t)(cur >> 32);
      moved = true;
    }
  }
  if (moved) st_relaxed(&kb.depth[t], bound);
}

// Colour pass (oitSpinlock.frag.glsl:49-129). The count of the reference
// (:86-87) lives in the lock word (count << 1 | held) and is taken first:
// one returning add per run of one target in the wave gives each fragment
// its place. A fragment whose place is below K is stored there without the
// lock (:89-93 take no decision that another fragment could change: the
// buffer is not full), and the wave then adds its stores, once they have
// completed, to the target's `filled` counter. A fragment whose place is K
// or more meets a full buffer: the early depth test (:66-69, read with the
// count) rejects it against the bound, else it takes the lock with an
// atomic OR, and runs the replacement section once `filled` shows all K
// stores, so a section never reads a slot that is still being written.
// The lock loop is driven by a wave-wide ballot: a lane that takes the lock
// runs its section (two: see the pairs below) and releases it in the same
// trip. Elements are ordered
// by (depth key, id); evicted and rejected fragments are recorded for the
// tail with their packed sRGB8 colour. A fragment the early test rejects is
// one its section would reject too (the bound is the key of an evicted
// element, and the stored elements only get nearer), so the test never
// changes the stored set or the tail; the count counts every fragment, so
// beyond K it differs from the reference's, which skips rejected ones
// (min(count, K), the composite's n, does not).
template <int K, bool kTail>
__global__ __launch_bounds__(kBlock) void k_spin_store(Common c, FragView f, KBuf kb, TailBuf tb) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, c.tab);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool valid = i < f.count;
  // the fragment's fields in one round trip, before any is tested
  const uint32_t t = valid ? ld_stream(&f.target[i]) : kEmpty;
  const float dv = valid ? ld_stream(&f.depth[i]) : 0.f;
  const float4 cv = valid ? ld_stream(&f.rgba[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t my_mask = (valid && f.mask) ? ld_stream(&f.mask[i]) : 0xFFFFFFFFu;
  if (t >= c.T) valid = false;
  if (valid && f.mask) {
    my_mask &= (1u << c.cov) - 1u;
    if (my_mask == 0u) valid = false;  // no sample covered: the shader does not run
  }
  if (valid && c.opq_z) {  // early depth test against the opaque pass
    my_mask = opaque_visible(c.opq_z, c.T, c.cov, t, depth_key(dv), my_mask);
    if (my_mask == 0u) valid = false;
  }
  const uint32_t lid = f.id_base + i;
  bool tail = false;
  uint32_t tail_id = lid;
  uint32_t tail_col = 0u;
  uint32_t tail_mask = my_mask;
  uint32_t my_key = 0u, my_col = 0u;
  if (valid) {
    my_key = depth_key(dv);
    my_col = encode_color_lds(cv, s_thr);
    tail_col = my_col;
  }
  // places: one add per run of one target, the early-test bound in the same round trip
  const Run r = same_target_run(valid ? t : kEmpty);
  uint32_t old = 0u, bound = 0u;
  if (valid && r.end)
    old = __hip_atomic_fetch_add(&kb.lock[t], 2u * (r.lane - r.start_lane + 1u), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
  if (valid) bound = ld_relaxed(&kb.depth[t]);
  old = __shfl(old, (int)r.end_lane);
  const uint32_t place = (old >> 1) + (r.lane - r.start_lane);
  const bool fill = valid && place < (uint32_t)K;
  if (fill) {
    const uint64_t idx = (uint64_t)place * c.T + t;
    st_relaxed(&kb.col[idx], my_col);
    st_relaxed(&kb.key[idx], my_key);
    st_relaxed(&kb.id[idx], lid);
    if (kb.mask) st_relaxed(&kb.mask[idx], my_mask);
  }
  if (__ballot(fill)) {  // wave-uniform: publish the stores, one add per run's last stored place
    store_fence();
    if (fill && (r.lane == r.end_lane || place + 1u == (uint32_t)K))
      __hip_atomic_fetch_add(&kb.filled[t], r.lane - r.start_lane + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // the places of a run grow with the lane, so its lanes at places K and
  // beyond are its last ones; they go in pairs: the first of a pair (a
  // sphere's front and back at one pixel) runs both sections, in id order,
  // under one acquisition, and the second gets its tail output back by
  // shuffle
  const uint32_t base = old >> 1;
  const uint32_t first_locker = r.start_lane + (base < (uint32_t)K ? (uint32_t)K - base : 0u);
  const bool locker = valid && !fill;
  const bool lead = locker && ((r.lane - first_locker) & 1u) == 0u && r.lane < r.end_lane;
  const bool follower = locker && ((r.lane - first_locker) & 1u) == 1u;
  const uint32_t p_key = __shfl_down(my_key, 1u), p_col = __shfl_down(my_col, 1u);
  const uint32_t p_mask = __shfl_down(my_mask, 1u);
  // the early test (its bound read with the places)
  const bool me_pass = locker && my_key <= bound;
  const bool p_pass = lead && p_key <= bound;
  bool p_tail = lead && !p_pass;  // the follower rejected: its own colour goes to the tail
  if (locker && !follower && !me_pass) tail = true;
  uint32_t p_tail_id = lid + 1u, p_tail_col = p_col, p_tail_mask = p_mask;
  bool pending = !follower && (me_pass || p_pass);
  uint32_t trips = 0;
  while (__ballot(pending)) {
    if (pending && ++trips > kSpinLimit) {  // a lock that is never released: give up with an error
      report_error(kb.err, kErrLockWait);
      pending = false;
    }
    if (pending) {
      const uint32_t stored = ld_relaxed(&kb.filled[t]);
      if (lock_try(&kb.lock[t])) {
        if (stored >= (uint32_t)K) {
          spin_replace<K>(c, kb, t, me_pass, my_key, my_col, lid, my_mask, tail, tail_id, tail_col, tail_mask, p_pass,
                          p_key, p_col, lid + 1u, p_mask, p_tail, p_tail_id, p_tail_col, p_tail_mask);
          pending = false;
        }
        lock_release(&kb.lock[t]);
      }
    }
  }
  // the follower's outcome, computed by its lead (lane - 1)
  const uint32_t b_tail = __shfl_up(p_tail ? 1u : 0u, 1u), b_id = __shfl_up(p_tail_id, 1u);
  const uint32_t b_col = __shfl_up(p_tail_col, 1u), b_mask = __shfl_up(p_tail_mask, 1u);
  if (follower) {
    tail = b_tail != 0u;
    tail_id = b_id;
    tail_col = b_col;
    tail_mask = b_mask;
  }
  if constexpr (kTail) {
    // every tail element blends its packed sRGB8 colour: whether a fragment
    // is rejected on arrival (the reference then blends its shaded colour,
    // oitSpinlock.frag.glsl:126-128) or stored and evicted later (its packed
    // colour, :112) depends on the arrival order, so only one colour rule
    // for both keeps the image independent of it
    float4 payload = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tail) payload = decode_premult(tail_col, c.tab);
    tail_queue_push(tail, valid ? t : kEmpty, tail_id, 0u, payload, tb, tail_mask);
  }
}

// The target's Simple / Spinlock words back to their clear values for the
// next frame (the clears of oitRender.cpp:161-169, 343-354, done by the
// composite that read them): Simple's count; the Spinlock's lock word, its
// completed-store count and its early-depth bound, which only an eviction
// sets, so only where the count passed K.
__device__ __forceinline__ void kbuf_reset(const KBuf& kb, uint32_t t, uint32_t count, uint32_t K) {
  if (count == 0u) return;
  if (kb.lock) {
    kb.lock[t] = 0u;
    if (kb.filled) kb.filled[t] = 0u;
    if (count > K) kb.depth[t] = kEmpty;
  } else {
    kb.count[t] = 0u;
  }
}

// Composite of Simple and Spinlock (oitSimple.frag.glsl:115-169 ==
// oitSpinlock.frag.glsl:153-207): n = min(count, K) elements, sorted,
// blended front to back, over the colour buffer.
template <int K, bool kColorTail, bool kDebug>
__global__ __launch_bounds__(kBlock) void k_kbuf_resolve(Common c, KBuf kb, TailBuf tb, ResolveOut o) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  __syncthreads();
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= c.T) return;
  const uint32_t count = kb.lock ? kb.lock[t] >> 1 : kb.count[t];  // Spinlock: the count in the lock word
  kbuf_reset(kb, t, count, (uint32_t)K);
  const uint32_t th = kColorTail ? tb.head[t] : 0u;                   // tail list head, with the count
  const uint32_t n = count < (uint32_t)K ? count : (uint32_t)K;
  // layers past the wave's largest count are skipped with a scalar branch,
  // and the k-th insert meets at most k occupied slots
  const uint32_t nmax = wave_max_u32(n);
  // every stored layer's planes are loaded before the first insert, so the
  // loads share one round trip instead of one each
  uint32_t lk[K], li[K], lc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    lk[k] = 0u;
    li[k] = 0u;
    lc[k] = 0u;
    if ((uint32_t)k < nmax && (uint32_t)k < n) {
      const uint64_t idx = (uint64_t)k * c.T + t;
      lk[k] = kb.key[idx];
      li[k] = kb.id[idx];
      lc[k] = kb.col[idx];
    }
  }
  FrontK<K> a;
  a.init();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if ((uint32_t)k < nmax && (uint32_t)k < n) {
      uint64_t ev;
      uint32_t evc;
      a.insert_bounded(((uint64_t)lk[k] << 32) | li[k], lc[k], ev, evc, k + 1);
    }
  }
  float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((uint32_t)k < nmax && (uint32_t)k < n) col = over(col, decode_premult_lds(a.col[k], s_dec, s_alpha));
  float4 d = dest_of(c, t);
  if (kColorTail) d = tail_fold_from(tb, th, d, 0u);
  o.out[t] = over(col, d);
  if (kDebug) write_debug<K>(o, t, a, n, count, false, false);
}

// Coverage-shaded composite of Simple and Spinlock (oitSimple.frag.glsl:
// 142-159, oitSpinlock.frag.glsl:180-197): the stored layers sorted by (depth
// key, id) with their slots, then colours and masks gathered by slot and
// blended per sample.
template <int K>
__global__ __launch_bounds__(kBlock) void k_kbuf_resolve_cov(Common c, KBuf kb, TailBuf tb, ResolveOut o) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  __syncthreads();
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= c.T) return;
  const uint32_t count = kb.lock ? kb.lock[t] >> 1 : kb.count[t];  // Spinlock: the count in the lock word
  kbuf_reset(kb, t, count, (uint32_t)K);
  const uint32_t n = count < (uint32_t)K ? count : (uint32_t)K;
  const uint32_t nmax = wave_max_u32(n);
  uint32_t lk[K], li[K];  // loaded before the first insert: one round trip
#pragma unroll
  for (int k = 0; k < K; ++k) {
    lk[k] = 0u;
    li[k] = 0u;
    if ((uint32_t)k < nmax && (uint32_t)k < n) {
      const uint64_t idx = (uint64_t)k * c.T + t;
      lk[k] = kb.key[idx];
      li[k] = kb.id[idx];
    }
  }
  FrontK<K> a;  // (key << 32 | id, slot)
  a.init();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if ((uint32_t)k < nmax && (uint32_t)k < n) {
      uint64_t ev;
      uint32_t evc;
      a.insert_bounded(((uint64_t)lk[k] << 32) | li[k], (uint32_t)k, ev, evc, k + 1);
    }
  }
  uint32_t cols[K], masks[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    cols[k] = 0u;
    masks[k] = 0u;
    if ((uint32_t)k < n) {
      const uint64_t idx = (uint64_t)a.col[k] * c.T + t;
      cols[k] = kb.col[idx];
      masks[k] = kb.mask[idx];
      a.col[k] = cols[k];
    }
  }
  const float4 col = blend_front_cov<K>(cols, masks, n, c.cov, s_dec, s_alpha);
  const float4 d = tail_fold_cov(tb, t, c);
  o.out[t] = over(col, d);
  write_debug<K>(o, t, a, n, count, false, false);
}

// ---------------------------------------------------------------------------
// Loop32: oitLoop.frag.glsl (the geometry is drawn twice)
// ---------------------------------------------------------------------------

// Depth pass (:57-100): the K nearest distinct depth keys per target, sorted,
// by a u32 atomicMin chain that stops at an empty slot or its own key; the two
// pretests use plain loads (slots only decrease within the pass). Two
// consecutive lanes of one target (a sphere's front and back) run one chain
// carrying both keys, as in k_loop64_store, with the set rule: a carried key
// that meets its equal in a slot, or equals the other carried key, is done.
// Each wave takes kU chunks of 64 consecutive fragments: the chunks' loads,
// pretests (with four probe slots) and atomic chains are in flight together,
// one dependent round trip per chain step of all of them.
template <int K, int kU>
__global__ __launch_bounds__(kBlock) void k_loop32_depth_m(Common c, FragView f, uint32_t* __restrict__ zs) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t first = ((uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64u * kU);
  uint32_t t[kU];
  float dv[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {  // every chunk's target and depth in one round trip
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    const bool in = i < f.count;
    t[u] = in ? ld_stream(&f.target[i]) : kEmpty;
    dv[u] = in ? ld_stream(&f.depth[i]) : 0.f;
  }
  uint32_t a[kU], b[kU];
  bool hasb[kU], act[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    uint32_t tt = t[u] < c.T ? t[u] : kEmpty;
    uint32_t z = 0u;
    if (tt != kEmpty) {
      z = depth_key(dv[u]);
      if (!opaque_visible(c.opq_z, c.T, 1u, tt, z, 1u)) tt = kEmpty;
    }
    const Run r = same_target_run(tt);
    const uint32_t zn = __shfl_down(z, 1u);
    const bool lead = tt != kEmpty && ((r.lane - r.start_lane) & 1u) == 0u;
    hasb[u] = lead && r.lane < r.end_lane && zn != z;
    a[u] = z;
    b[u] = 0u;
    if (hasb[u]) {
      a[u] = z < zn ? z : zn;
      b[u] = z < zn ? zn : z;
    }
    act[u] = lead;
    t[u] = tt;
  }
  // the pretests and the first slots of every chunk in one round trip
  constexpr int kProbe = K / 2 < 4 ? K / 2 : 4;
  int j[kU];
  {
    uint32_t probe[kU][kProbe > 0 ? kProbe : 1], z_last[kU], z_mid[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (act[u]) {
#pragma unroll
        for (int k = 0; k < kProbe; ++k) probe[u][k] = zs[(uint64_t)k * c.T + t[u]];
        z_last[u] = zs[(uint64_t)(K - 1) * c.T + t[u]];
        z_mid[u] = zs[(uint64_t)(K / 2) * c.T + t[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      j[u] = 0;
      if (act[u]) {
        if (a[u] > z_last[u]) {
          act[u] = false;
        } else if (a[u] > z_mid[u]) {
          j[u] = K / 2;
        } else {
#pragma unroll
          for (int k = 0; k < kProbe; ++k) j[u] += probe[u][k] < a[u] ? 1 : 0;
        }
      }
    }
  }
  // the chains, one step of every live chain per round trip
  for (;;) {
    uint32_t old[kU];
    bool any = false;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (act[u]) {
        old[u] = __hip_atomic_fetch_min(&zs[(uint64_t)j[u] * c.T + t[u]], a[u], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        any = true;
      }
    }
    returns_done();
    if (!any) break;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (!act[u]) continue;
      if (old[u] == kEmpty || old[u] == a[u]) {  // a is in slot j now (or was already): b goes on
        if (hasb[u]) {
          a[u] = b[u];
          hasb[u] = false;
        } else {
          act[u] = false;
        }
      } else if (old[u] > a[u]) {  // a took slot j and old comes out
        if (!hasb[u]) {
          a[u] = old[u];
        } else if (old[u] == b[u]) {  // the two carried keys are equal: one of them is done
          a[u] = b[u];
          hasb[u] = false;
        } else if (old[u] < b[u]) {
          a[u] = old[u];
        } else {
          a[u] = b[u];
          b[u] = old[u];
        }
      }  // else a smaller key holds slot j: a goes on unchanged
      if (++j[u] >= K) act[u] = false;
    }
  }
}

// Colour pass (:105-180): behind the K-th depth -> tail-blend the shaded
// colour; else the slot of its depth by binary search, where the colour of the
// latest fragment in primitive order is kept ((id + 1) << 32 | colour,
// atomicMax), the result of the reference's plain store in draw order.
// Each wave takes kU chunks of 64 consecutive fragments: every chunk's
// fields, then every chunk's lower-half depth slots, in one round trip each;
// the upper half only where a pixel holds more than K/2 depths below the
// fragment's.
template <int K, bool kTail, int kU>
__global__ __launch_bounds__(kBlock) void k_loop32_color_m(Common c, FragView f, const uint32_t* __restrict__ zs,
                                                           uint64_t* __restrict__ cs, TailBuf tb) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, c.tab);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t first = ((uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64u * kU);
  uint32_t t[kU], z[kU];
  float dv[kU];
  float4 cv[kU];
  bool valid[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    valid[u] = i < f.count;
    t[u] = valid[u] ? ld_stream(&f.target[i]) : kEmpty;
    dv[u] = valid[u] ? ld_stream(&f.depth[i]) : 0.f;
    cv[u] = valid[u] ? ld_stream(&f.rgba[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    if (t[u] >= c.T) valid[u] = false;
    z[u] = valid[u] ? depth_key(dv[u]) : 0u;
    if (valid[u] && !opaque_visible(c.opq_z, c.T, 1u, t[u], z[u], 1u)) valid[u] = false;
    if (!valid[u]) t[u] = kEmpty;
  }
  // the first half of every chunk's depth slots in one round trip, loaded by
  // every lane (one without a fragment reads target 0's), so no load waits
  // behind another; the second half only where the first is all nearer
  constexpr int kHalf = K / 2 > 0 ? K / 2 : K;
  uint32_t below[kU], ts[kU];
  {
    uint32_t zk[kU][kHalf];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      ts[u] = valid[u] ? t[u] : 0u;
#pragma unroll
      for (int k = 0; k < kHalf; ++k) zk[u][k] = zs[(uint64_t)k * c.T + ts[u]];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      below[u] = 0;
#pragma unroll
      for (int k = 0; k < kHalf; ++k) below[u] += valid[u] && zk[u][k] < z[u] ? 1u : 0u;
    }
  }
  if constexpr (kHalf < K) {
    bool more[kU], any = false;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      more[u] = valid[u] && below[u] == (uint32_t)kHalf;
      any = any || more[u];
    }
    if (__ballot(any)) {  // wave-uniform: then every lane loads, as above
      uint32_t zk[kU][K - kHalf];
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int k = kHalf; k < K; ++k) zk[u][k - kHalf] = zs[(uint64_t)k * c.T + ts[u]];
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int k = kHalf; k < K; ++k) below[u] += more[u] && zk[u][k - kHalf] < z[u] ? 1u : 0u;
    }
  }
  bool tail[kU];
  uint32_t lid[kU];
  float4 payload[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    lid[u] = f.id_base + (uint32_t)(first + (uint64_t)u * 64u + lane);
    tail[u] = false;
    payload[u] = valid[u] ? cv[u] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid[u]) {
      if (below[u] == (uint32_t)K) {
        tail[u] = true;
      } else {
        const uint64_t v = ((uint64_t)(lid[u] + 1u) << 32) | encode_color_lds(cv[u], s_thr);
        (void)__hip_atomic_fetch_max(&cs[(uint64_t)below[u] * c.T + t[u]], v, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  if constexpr (kTail) tail_queue_push_n<kU>(tail, t, lid, payload, tb);
}

// Composite (:186-222): the layers up to the first empty depth, front to back,
// over the colour buffer. The slots it read are set back to empty (depth
// kEmpty, colour 0) for the next frame: the stored depths are a prefix, and
// the colour pass writes only the slot of a stored depth.
template <int K, bool kColorTail, bool kDebug>
__global__ __launch_bounds__(kBlock) void k_loop32_resolve(Common c, uint32_t* __restrict__ zs,
                                                           uint64_t* __restrict__ cs, TailBuf tb, ResolveOut o) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  __syncthreads();
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= c.T) return;
  uint32_t z[K], col[K];
  uint32_t n = 0;
  bool alive = true;
  // depths are sorted with the empty slots last: the upper half is loaded
  // only where the lower half is full
  constexpr int kHalf = K / 2 > 0 ? K / 2 : K;
  const uint32_t th = kColorTail ? tb.head[t] : 0u;  // tail list head, with the depths
#pragma unroll
  for (int k = 0; k < kHalf; ++k) z[k] = zs[(uint64_t)k * c.T + t];
#pragma unroll
  for (int k = kHalf; k < K; ++k) z[k] = kEmpty;
  if (z[kHalf - 1] != kEmpty) {
#pragma unroll
    for (int k = kHalf; k < K; ++k) z[k] = zs[(uint64_t)k * c.T + t];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    alive = alive && z[k] != kEmpty;
    n += alive ? 1u : 0u;
  }
  // the stored layers' colours, loaded together once n is known
#pragma unroll
  for (int k = 0; k < K; ++k) col[k] = (uint32_t)k < n ? (uint32_t)cs[(uint64_t)k * c.T + t] : 0u;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if ((uint32_t)k < n) {
      zs[(uint64_t)k * c.T + t] = kEmpty;
      cs[(uint64_t)k * c.T + t] = 0ull;
    }
  }
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((uint32_t)k < n) acc = over(acc, decode_premult_lds(col[k], s_dec, s_alpha));
  float4 d = dest_of(c, t);
  if (kColorTail) d = tail_fold_from(tb, th, d, 0u);
  o.out[t] = over(acc, d);
  if (kDebug) {
    if (o.fk_key) {
#pragma unroll
      for (int k = 0; k < K; ++k) o.fk_key[(uint64_t)t * K + k] = (uint32_t)k < n ? z[k] : kEmpty;
    }
    if (o.fk_val) {
#pragma unroll
      for (int k = 0; k < K; ++k) o.fk_val[(uint64_t)t * K + k] = (uint32_t)k < n ? col[k] : kEmpty;
    }
    if (o.counts) o.counts[t] = n;
  }
}

// ---------------------------------------------------------------------------
// Ordered interlock: oitInterlock.frag.glsl:52-148 with pixel_interlock_ordered
// ---------------------------------------------------------------------------

// The critical sections of one pixel run one at a time in primitive order:
// the fragments are in the pixel's list (Linked List append, newest first),
// so the list is reversed and the k-buffer logic of :103-148 (early depth
// bound, fill, replace the furthest) is replayed oldest first in registers;
// each invocation's tail output (its shaded colour when rejected, the evicted
// packed colour when it replaces one) is blended by the ROP in that same
// order. The walk leaves the pool clean (heads 0, links kUnset).
//
// The k-buffer is kept sorted by (depth key, id) in registers. Ids make the
// keys unique, so "first index wins" never decides, the furthest element is
// the last one, and replacing it when the newcomer is nearer is the
// insertion that pushes the last one out (:78-101): the set, the evictions
// and so the tail outputs equal those of the slot array the shader keeps,
// each insert is bounded by the nodes seen so far (a wave-uniform trip
// count), and the composite needs no sort. With coverage shading (kCov) the
// stored layers carry their sample masks, each invocation's tail output goes
// to the samples of its own mask, and the front K are blended per sample.
//
// The first walk (two consecutive nodes per round trip where the list has
// them) keeps the newest kStack (6 or 8) nodes of each list (index, depth and
// colour, mask) in LDS, one column per thread, and reverses only the older
// ones in place; the replay takes the reversed part first, then the
// LDS entries newest last. A list of up to kStack nodes is thus read from
// memory once, the deeper part of a longer one twice.
template <int K, bool kCov>
__global__ __launch_bounds__(kBlock, (K == 16 && !kCov) ? 6 : 1) void k_interlock_resolve(Common c, LLBuf ll, ResolveOut o) {
  // K <= 8 keeps six nodes (21 KB of LDS per block: seven blocks per CU),
  // K = 16 and 32 eight (27 KB), where registers allow fewer waves anyway
  constexpr int kStack = K <= 8 ? 6 : 8;
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  __shared__ uint32_t s_id[kStack][kBlock];
  __shared__ uint64_t s_cd[kStack][kBlock];
  // masks only with coverage: without, K = 16 fits six blocks per CU
  // (26 KB each) and the launch bound holds it to the 80 VGPRs of six
  // waves per SIMD
  __shared__ uint32_t s_mk[kCov ? kStack : 1][kCov ? kBlock : 1];
  stage_decode(s_dec, s_alpha, c.tab);
  if (blockIdx.x == 0 && threadIdx.x < 16u) {  // frame counters -> last-frame slot, zero for the next frame
    ll.counters[16 + threadIdx.x] = ll.counters[threadIdx.x];
    ll.counters[threadIdx.x] = 0u;
  }
  __syncthreads();
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < c.T; i += stride) {
    const uint32_t t = pm_target(i, c.T, c.spp);  // heads are pixel-major
    const uint32_t tx = threadIdx.x;
    uint32_t p = ll.head[i];
    if (p != 0u) ll.head[i] = 0u;
    uint32_t prev = 0;
    uint32_t depth = 0;
    auto take = [&](uint32_t node, uint64_t cdv, uint32_t mv) {
      if (depth < (uint32_t)kStack) {  // newest nodes: to LDS, link released now
        ll.next[node] = kUnset;
        s_id[depth][tx] = node;
        s_cd[depth][tx] = cdv;
        if constexpr (kCov) s_mk[depth][tx] = mv;
      } else {  // older nodes: reversed in place, oldest first
        ll.next[node] = prev;
        prev = node;
      }
      ++depth;
    };
    // node p-1 is loaded with p: a sphere's front and back fragment are
    // consecutive nodes chained p -> p-1, so one round trip takes both
    while (p != 0u && p < ll.cap) {
      const uint32_t q = p - 1u;
      const uint32_t nx0 = ll.next[p];
      const uint32_t nx1 = ll.next[q];
      const uint64_t cd0 = ll.cd[p];
      const uint64_t cd1 = ll.cd[q];
      const uint32_t m0 = kCov ? ll.mask[p] : 0u;
      const uint32_t m1 = kCov ? ll.mask[q] : 0u;
      take(p, cd0, m0);
      if (nx0 == q && q != 0u) {
        take(q, cd1, m1);
        p = descend(nx1, q, ll.err);
      } else {
        p = descend(nx0, p, ll.err);
      }
    }
    uint32_t rem = depth < (uint32_t)kStack ? depth : (uint32_t)kStack;  // LDS entries left
    FrontKM<K> a;
    a.init();
    uint32_t count = 0;
    uint32_t bound = kEmpty;  // imgAuxDepth cleared to ~0, oitRender.cpp:348
    uint32_t trip = 0;        // nodes replayed so far: the same on every active lane
    float4 cb = kCov ? c.background : dest_of(c, t);
    float4 cbs[kCov ? 8 : 1];  // coverage shading: the colour buffer per sample
#pragma unroll
    for (int s = 0; s < (kCov ? 8 : 1); ++s)
      cbs[s] = (kCov && c.opq_rgba && (uint32_t)s < c.cov) ? c.opq_rgba[(uint64_t)s * c.T + t] : c.background;
    bool seen = false;
    // one invocation's critical section (:62-108) at k-buffer bound w
    auto replay = [&](uint32_t node, uint64_t cd, uint32_t mm, int w) {
      const uint32_t key = (uint32_t)(cd >> 32);
      const uint64_t me = ((uint64_t)key << 32) | (node - 1u);
      bool tail = true;
      bool own = true;
      uint32_t tcol = 0u;
      if (key <= bound) {  // :62-64 early depth
        uint64_t ev;
        uint32_t evc;
        a.insert_bounded(me, (uint32_t)cd, mm, ev, evc, w);
        if (count < (uint32_t)K) {  // :69-75 filled a free slot
          tail = false;
        } else if (ev != me) {  // :91-101 replaced the furthest
          own = false;
          tcol = evc;
          bound = (uint32_t)(ev >> 32);
        }
        ++count;
      }
      if (tail && c.tail) {  // :106-108 the invocation's output, PREMULTIPLIED blend in primitive order
        const float4 tc = own ? ll.full[node] : decode_premult_lds(tcol, s_dec, s_alpha);
        if constexpr (kCov) {
#pragma unroll
          for (int s = 0; s < 8; ++s)
            if ((uint32_t)s < c.cov && ((mm >> s) & 1u)) cbs[s] = over(tc, cbs[s]);
          seen = true;
        } else {
          cb = over(tc, cb);
        }
      }
    };
    // K <= 8: a trip of the reversed part takes two chained nodes when it
    // finds them (at K = 16 and 32 the second inlined insert costs more
    // registers, and so occupancy, than the round trips it saves)
    constexpr bool kTwo = K <= 8;
    p = prev;
    for (;;) {  // oldest first
      // a trip replays at most two nodes (one without kTwo), which bounds
      // the occupied slots on every active lane
      const int w = (int)__builtin_amdgcn_readfirstlane(kTwo ? 2u * trip + 2u : trip + 1u);
      ++trip;
      if (p != 0u && !kTwo) {  // the reversed part (only nodes below the capacity were linked)
        const uint32_t nx = ll.next[p];
        ll.next[p] = kUnset;
        const uint64_t cd0 = ll.cd[p];
        const uint32_t m0 = kCov ? ll.mask[p] : 0xFFFFFFFFu;
        replay(p, cd0, m0, w);
        p = nx;
      } else if (p != 0u) {
        // a pair chained p -> p-1 by the append is now p-1 -> p: node p+1
        // is loaded with p and taken along when it follows
        const uint32_t r = p + 1u;
        const bool rok = r < ll.cap;
        const uint32_t nx0 = ll.next[p];
        const uint32_t nx1 = rok ? ll.next[r] : 0u;
        const uint64_t cd0 = ll.cd[p];
        const uint64_t cd1 = rok ? ll.cd[r] : 0ull;
        const uint32_t m0 = kCov ? ll.mask[p] : 0xFFFFFFFFu;
        const uint32_t m1 = (kCov && rok) ? ll.mask[r] : 0xFFFFFFFFu;
        ll.next[p] = kUnset;
        replay(p, cd0, m0, w);
        if (rok && nx0 == r) {
          ll.next[r] = kUnset;
          replay(r, cd1, m1, w);
          p = nx1;
        } else {
          p = nx0;
        }
      } else if (rem > 0u) {
        --rem;
        uint32_t mm = 0xFFFFFFFFu;
        if constexpr (kCov) mm = s_mk[rem][tx];
        replay(s_id[rem][tx], s_cd[rem][tx], mm, w);
      } else {
        break;
      }
    }
    // composite (:176-230): the stored layers are already sorted
    const uint32_t n = count < (uint32_t)K ? count : (uint32_t)K;
    float4 front = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 d = cb;
    if constexpr (kCov) {
      front = blend_front_cov<K>(a.col, a.msk, n, c.cov, s_dec, s_alpha);
      if (seen || c.opq_rgba) {  // the MSAA resolve of the per-sample colour buffer
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int s = 0; s < 8; ++s)
          if ((uint32_t)s < c.cov) sum = make_float4(sum.x + cbs[s].x, sum.y + cbs[s].y, sum.z + cbs[s].z, sum.w + cbs[s].w);
        const float fs = (float)c.cov;
        d = make_float4(sum.x / fs, sum.y / fs, sum.z / fs, sum.w / fs);
      }
    } else {
      const uint32_t nmax = wave_max_u32(n);
#pragma unroll
      for (int k = 0; k < K; ++k)
        if ((uint32_t)k < nmax && (uint32_t)k < n) front = over(front, decode_premult_lds(a.col[k], s_dec, s_alpha));
    }
    o.out[t] = over(front, d);
    if (o.fk_key) {
#pragma unroll
      for (int k = 0; k < K; ++k) o.fk_key[(uint64_t)t * K + k] = (uint32_t)k < n ? (uint32_t)(a.sk[k] >> 32) : kEmpty;
    }
    if (o.fk_val) {
#pragma unroll
      for (int k = 0; k < K; ++k) o.fk_val[(uint64_t)t * K + k] = (uint32_t)k < n ? (uint32_t)a.sk[k] : kEmpty;
    }
    if (o.counts) o.counts[t] = count;
  }
}

// ---------------------------------------------------------------------------
// Loop64: oitLoop64.frag.glsl
// ---------------------------------------------------------------------------

// Ascending sort of N (a power of two) values held in registers: a bitonic
// network with compile-time indices (no spills, no dynamic indexing).
template <int N>
__device__ __forceinline__ void sort_u64(uint64_t* v) {
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const uint64_t x = v[i], y = v[j];
          const bool swap = up ? (x > y) : (x < y);
          v[i] = swap ? y : x;
          v[j] = swap ? x : y;
        }
      }
    }
  }
}

// A chunk's mask, read as one wave-uniform (scalar) value.
__device__ __forceinline__ uint64_t big_mask_load(const uint64_t* p) {
  const uint64_t m = *p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)m);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(m >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Per-target count of the frame's fragments that pass the opaque depth
// test: one returnless atomic per run of one target in a wave. Runs at
// append, so that the stores, at resolve, know every pixel's final count.
// kU chunks of 64 fragments per wave, their loads in one round trip; the
// depths are read only when an opaque pass can cull.
template <int kU>
__global__ __launch_bounds__(kBlock) void k_loop64_count(Common c, FragView f, uint32_t* __restrict__ cnt) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t first = ((uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64u * kU);
  const bool culls = c.opq_z != nullptr;
  uint32_t t[kU];
  float dv[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    const bool in = i < f.count;
    t[u] = in ? ld_stream(&f.target[i]) : kEmpty;
    dv[u] = (in && culls) ? ld_stream(&f.depth[i]) : 0.f;
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    uint32_t tt = t[u] < c.T ? t[u] : kEmpty;
    if (culls && tt != kEmpty && !opaque_visible(c.opq_z, c.T, 1u, tt, depth_key(dv[u]), 1u)) tt = kEmpty;
    const Run r = same_target_run(tt);
    if (tt != kEmpty && r.start) atomicAdd(&cnt[tt], r.end_lane - r.lane + 1u);
  }
}

// Colour pass (oitLoop64.frag.glsl:65-141): key = depth << 32 | colour (:79).
// The slots end as the K smallest keys of each pixel for any arrival order;
// this pass gets there by one of two routes per pixel, chosen by the
// pixel's count of the frame's fragments:
// - at most K fragments: every one of them is kept, so each takes a slot
//   index from its count (one returning atomic per run of one target,
//   counting down to 0) and stores its key unsorted; the composite sorts
//   the slots. No fragment of such a pixel is in the tail.
// - more than K: the two pretests of USE_EARLYDEPTH (:82-101, plain loads:
//   slots only decrease within a frame, so a stale value only lets more
//   fragments try), then the atomicMin chain (:109-121), which keeps the
//   slots sorted. Two consecutive lanes of one target (a sphere's front and
//   back fragment) run one chain that carries both keys: at each slot the
//   smaller carried key goes in, the key that comes out joins the larger
//   one, and the chain ends when both are placed or pushed past slot K-1.
//   The multiset of slots and carried keys is the same as after the two
//   single chains, so the result is identical; the pair costs one chain
//   instead of two and its lanes do not contend for the same slots.
// Each wave takes kU chunks of 64 consecutive fragments: the chunks' loads,
// counts, pretests and chain steps are in flight together.
template <int K, int kU>
__global__ __launch_bounds__(kBlock) void k_loop64_store_m(Common c, FragView f, uint64_t* __restrict__ slots,
                                                           uint32_t* __restrict__ cnt, uint64_t* __restrict__ big) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, c.tab);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t first = ((uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64u * kU);
  uint32_t t[kU];
  float dv[kU];
  float4 cv[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {  // every chunk's fields in one round trip
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    const bool in = i < f.count;
    t[u] = in ? ld_stream(&f.target[i]) : kEmpty;
    dv[u] = in ? ld_stream(&f.depth[i]) : 0.f;
    cv[u] = in ? ld_stream(&f.rgba[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  uint64_t z[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    uint32_t tt = t[u] < c.T ? t[u] : kEmpty;
    z[u] = kEmpty64;
    if (tt != kEmpty) {
      const uint32_t key = depth_key(dv[u]);
      if (opaque_visible(c.opq_z, c.T, 1u, tt, key, 1u)) z[u] = ((uint64_t)key << 32) | encode_color_lds(cv[u], s_thr);
      else tt = kEmpty;
    }
    t[u] = tt;
  }
  // every chunk's pixel counts in one round trip, loaded by every lane (one
  // without a fragment reads target 0's), so no load waits behind another
  uint32_t n[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) n[u] = cnt[t[u] != kEmpty ? t[u] : 0u];
  // pixels of at most K fragments: a slot index per fragment from the count,
  // every chunk's atomics in flight together
  uint32_t base[kU], start_lane[kU], rank[kU];
  bool small[kU], sub[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    small[u] = t[u] != kEmpty && n[u] <= (uint32_t)K;
    const Run r = same_target_run(small[u] ? t[u] : kEmpty);
    sub[u] = small[u] && r.start;
    rank[u] = r.end_lane - r.lane + 1u;  // the run's length, at its first lane
    start_lane[u] = r.start_lane;
    base[u] = 0u;
  }
#pragma unroll
  for (int u = 0; u < kU; ++u)
    if (sub[u]) base[u] = atomicSub(&cnt[t[u]], rank[u]);
  returns_done();
#pragma unroll
  for (int u = 0; u < kU; ++u)  // this lane's index + 1
    base[u] = __shfl(base[u], (int)start_lane[u]) - (lane_id() - start_lane[u]);
  // pixels of more than K: their fragments' lanes, per 64-fragment chunk,
  // for the classification (which skips the rest); pairs and chains
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t m = __ballot(t[u] != kEmpty && !small[u]);
    if (lane == 0u && first + (uint64_t)u * 64u < f.count) big[first / 64u + u] = m;
  }
  uint64_t a[kU], b[kU];
  bool hasb[kU], act[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint32_t tb = small[u] ? kEmpty : t[u];
    const Run r = same_target_run(tb);
    const uint32_t zn_lo = __shfl_down((uint32_t)z[u], 1u);
    const uint32_t zn_hi = __shfl_down((uint32_t)(z[u] >> 32), 1u);
    const uint64_t zn = ((uint64_t)zn_hi << 32) | zn_lo;
    const bool lead = tb != kEmpty && ((r.lane - r.start_lane) & 1u) == 0u;
    hasb[u] = lead && r.lane < r.end_lane;
    a[u] = z[u];
    b[u] = 0;
    if (hasb[u]) {
      a[u] = z[u] < zn ? z[u] : zn;
      b[u] = z[u] < zn ? zn : z[u];
    }
    act[u] = lead;
  }
#pragma unroll
  for (int u = 0; u < kU; ++u)  // the stores of the kept fragments (index base - 1 < K)
    if (small[u] && base[u] - 1u < (uint32_t)K) slots[(uint64_t)(base[u] - 1u) * c.T + t[u]] = z[u];
  // the pretests and the first slots of every chunk in one round trip
  constexpr int kProbe = K / 2 < 4 ? K / 2 : 4;
  int j[kU];
  {
    uint64_t probe[kU][kProbe > 0 ? kProbe : 1], s_last[kU], s_mid[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (act[u]) {
#pragma unroll
        for (int k = 0; k < kProbe; ++k) probe[u][k] = slots[(uint64_t)k * c.T + t[u]];
        s_last[u] = slots[(uint64_t)(K - 1) * c.T + t[u]];
        s_mid[u] = slots[(uint64_t)(K / 2) * c.T + t[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      j[u] = 0;
      if (act[u]) {
        if (a[u] > s_last[u]) {
          act[u] = false;
        } else if (a[u] > s_mid[u]) {
          j[u] = K / 2;
        } else {
#pragma unroll
          for (int k = 0; k < kProbe; ++k) j[u] += probe[u][k] < a[u] ? 1 : 0;
        }
      }
    }
  }
  // the chains, one step of every live chain per round trip
  for (;;) {
    uint64_t old[kU];
    bool any = false;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (act[u]) {
        old[u] = __hip_atomic_fetch_min(&slots[(uint64_t)j[u] * c.T + t[u]], a[u], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        any = true;
      }
    }
    returns_done();
    if (!any) break;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (!act[u]) continue;
      if (old[u] == kEmpty64) {  // placed in an empty slot: b (if any) goes to the next one
        if (hasb[u]) {
          a[u] = b[u];
          hasb[u] = false;
        } else {
          act[u] = false;
        }
      } else {
        const uint64_t up = old[u] > a[u] ? old[u] : a[u];  // the key that leaves slot j
        if (hasb[u] && b[u] < up) {
          a[u] = b[u];
          b[u] = up;
        } else {
          a[u] = up;
        }
      }
      if (++j[u] >= K) act[u] = false;
    }
  }
}

// Tail classification: a fragment whose key is above the final K-th key was
// evicted or rejected (the colours oitLoop64.frag.glsl:131-136 tail-blends);
// one equal to it is "boundary": identical (depth, colour) keys are the same
// bits in the A-buffer, and the resolve keeps the lowest ids of them in front.
// Only pixels of more than K fragments can have tails (the others keep every
// fragment): the store left, per 64-fragment chunk, the mask of the lanes
// whose fragment belongs to such a pixel, and a chunk with an empty mask is
// skipped. Each wave takes kU chunks and loads, one round trip each: the
// masked lanes' targets and depths, then every lane's K-th slot, then the
// colours of the lanes at or past the K-th depth; the tail fragments go
// onto their lists with every chunk's exchanges in flight together.
template <int K, int kU>
__global__ __launch_bounds__(kBlock) void k_loop64_classify(Common c, FragView f, const uint64_t* __restrict__ slots,
                                                            const uint64_t* __restrict__ bigm, TailBuf tb) {
  __shared__ float s_thr[257];
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  stage_thresholds(s_thr, c.tab);  // its barrier also publishes the decode tables
  const uint32_t lane = lane_id();
  const uint64_t first = ((uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64u * kU);
  bool big[kU];
  uint64_t any = 0;
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t m = first + (uint64_t)u * 64u < f.count ? big_mask_load(&bigm[first / 64u + u]) : 0ull;
    big[u] = (m >> lane) & 1ull;
    any |= m;
  }
  if (any == 0ull) return;  // wave-uniform: no fragment of this wave can be in the tail
  uint32_t t[kU];
  float dv[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const uint64_t i = first + (uint64_t)u * 64u + lane;
    t[u] = big[u] ? ld_stream(&f.target[i]) : 0u;
    dv[u] = big[u] ? ld_stream(&f.depth[i]) : 0.f;
  }
  // every chunk's K-th slots in one round trip, loaded by every lane (a lane
  // outside such a pixel reads target 0's: one line for all of them), so no
  // load waits behind another
  uint64_t sk[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) sk[u] = slots[(uint64_t)(K - 1) * c.T + t[u]];
  uint32_t key[kU];
  bool cand[kU];
  float4 cv[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    key[u] = depth_key(dv[u]);
    // the key's high half is the depth: a nearer fragment stays in the slots
    cand[u] = big[u] && key[u] >= (uint32_t)(sk[u] >> 32);
    cv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u)  // the candidates' colours, one round trip
    if (cand[u]) cv[u] = ld_stream(&f.rgba[first + (uint64_t)u * 64u + lane]);
  returns_done();
  bool tail[kU];
  uint32_t flag[kU], lid[kU];
  float4 payload[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    tail[u] = false;
    flag[u] = 0u;
    payload[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cand[u]) {
      const uint32_t col = encode_color_lds(cv[u], s_thr);
      const uint64_t z = ((uint64_t)key[u] << 32) | col;
      if (z >= sk[u]) {
        tail[u] = true;
        flag[u] = z == sk[u] ? 1u : 0u;
        payload[u] = decode_premult_lds(col, s_dec, s_alpha);
      }
    }
    lid[u] = f.id_base + (uint32_t)(first + (uint64_t)u * 64u + lane);
  }
  tail_list_push_n<kU>(tail, t, lid, flag, payload, tb);
}

// Composite (oitLoop64.frag.glsl:162-183): blend slots front to back until
// the first empty one. The slots it read are set back to empty for the next
// frame (the occupied ones are a prefix: slots stay sorted, and no key is
// the empty key), so the next clear need not write the whole A-buffer.
template <int K, bool kColorTail, bool kDebug>
__global__ __launch_bounds__(kBlock) void k_loop64_resolve(Common c, uint64_t* __restrict__ slots,
                                                           uint32_t* __restrict__ cnt, TailBuf tb, ResolveOut o) {
  __shared__ float s_dec[256];
  __shared__ float s_alpha[256];
  stage_decode(s_dec, s_alpha, c.tab);
  __syncthreads();
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= c.T) return;
  // slots are sorted with the empty ones last: the upper half is loaded only
  // where the lower half is full
  constexpr int kHalf = K / 2 > 0 ? K / 2 : K;
  const uint32_t th = kColorTail ? tb.head[t] : 0u;  // tail list head, with the slots
  const uint32_t n_frame = cnt[t];  // more than K: chained, sorted; else 0 and the slots unsorted
  uint64_t s[K];
#pragma unroll
  for (int k = 0; k < kHalf; ++k) s[k] = slots[(uint64_t)k * c.T + t];
  const bool upper = (uint32_t)(s[kHalf - 1] >> 32) != kEmpty;
#pragma unroll
  for (int k = kHalf; k < K; ++k) s[k] = upper ? slots[(uint64_t)k * c.T + t] : kEmpty64;
  // a pixel of at most K fragments stored them unsorted (empty slots last,
  // the empty key being the largest): sort; sorted slots stay as they are.
  // The network's size is wave-uniform: half of K where no lane holds more.
  if (__ballot(upper)) sort_u64<K>(s);
  else sort_u64<kHalf>(s);
  if (n_frame) cnt[t] = 0u;
  float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t n = 0;
  bool alive = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    alive = alive && (uint32_t)(s[k] >> 32) != kEmpty;
    if (alive) {
      col = over(col, decode_premult_lds((uint32_t)s[k], s_dec, s_alpha));
      ++n;
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((uint32_t)k < n) slots[(uint64_t)k * c.T + t] = kEmpty64;
  float4 d = dest_of(c, t);
  if (kColorTail) {
    uint32_t skip = 0;
    if (n == (uint32_t)K) {
#pragma unroll
      for (int k = 0; k < K; ++k) skip += s[k] == s[K - 1] ? 1u : 0u;
    }
    d = tail_fold_from(tb, th, d, skip);
  }
  o.out[t] = over(col, d);
  if (kDebug) {
    if (o.fk_key) {
#pragma unroll
      for (int k = 0; k < K; ++k) o.fk_key[(uint64_t)t * K + k] = (uint32_t)k < n ? (uint32_t)(s[k] >> 32) : kEmpty;
    }
    if (o.fk_val) {
#pragma unroll
      for (int k = 0; k < K; ++k) o.fk_val[(uint64_t)t * K + k] = (uint32_t)k < n ? (uint32_t)s[k] : kEmpty;
    }
    if (o.counts) o.counts[t] = n;
  }
}

// ---------------------------------------------------------------------------
// Weighted, blended OIT: oitWeighted.frag.glsl
// ---------------------------------------------------------------------------

// Colour pass (oitWeighted.frag.glsl:53-79) with the two framebuffer blends
// (main.cpp:559-577): accum += premult * w (ONE, ONE) as float atomics,
// reveal *= (1 - a) (ZERO, ONE_MINUS_SRC_COLOR) as a compare-and-swap loop.
__global__ __launch_bounds__(kBlock) void k_wboit_store(Common c, FragView f, WBBuf wb, float vnear, float vspan) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  float4 wc = make_float4(0.f, 0.f, 0.f, 0.f);
  float lr = 0.f;
  // the fragment's fields in one round trip, before any is tested
  const bool in = i < f.count;
  uint32_t t = in ? ld_stream(&f.target[i]) : kEmpty;
  const float dv = in ? ld_stream(&f.depth[i]) : 0.f;
  const float4 cv = in ? ld_stream(&f.rgba[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
  if (t >= c.T) t = kEmpty;
  if (t != kEmpty && !opaque_visible(c.opq_z, c.T, 1u, t, depth_key(dv), 1u)) t = kEmpty;
  if (t != kEmpty) {
    const float4 col = cv;
    const float view = vnear + dv * vspan;   // stands in for -viewZ (object.vert.glsl:35)
    const float depthZ = view * 10.0f;               // :65
    const float x = depthZ / 200.0f;
    const float x2 = x * x;
    const float x4 = x2 * x2;                        // pow(depthZ / 200, 4.0)
    float dw = 0.03f / (1e-5f + x4);                 // :67
    dw = fminf(fmaxf(dw, 1e-2f), 3e3f);
    float aw = fminf(1.0f, fmaxf(fmaxf(col.x, col.y), fmaxf(col.z, col.w)) * 40.0f + 0.01f);  // :69
    aw = aw * aw;                                    // :70
    const float w = aw * dw;                         // :72
    wc = make_float4(col.x * w, col.y * w, col.z * w, col.w * w);
    // revealage is a product of (1 - alpha) (the blend's ZERO, ONE_MINUS_SRC_ALPHA
    // factors, main.cpp:559-577); it is accumulated as a sum of log2 terms so
    // that every update is an add without a returned value
    lr = log2f(fmaxf(1.0f - col.w, 0.0f));
  }
  // consecutive lanes of one target (a sphere's front and back) are summed
  // in the wave first: one set of atomics per run
  const Run r = same_target_run(t);
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const float ux = __shfl_up(wc.x, o), uy = __shfl_up(wc.y, o), uz = __shfl_up(wc.z, o), uw = __shfl_up(wc.w, o);
    const float ul = __shfl_up(lr, o);
    if (r.lane >= r.start_lane + o) {
      wc.x += ux;
      wc.y += uy;
      wc.z += uz;
      wc.w += uw;
      lr += ul;
    }
  }
  if (t != kEmpty && r.end) {  // channel planes: each atomic instruction of a wave stays in few cache lines
    float* acc = reinterpret_cast<float*>(wb.accum);
    const uint32_t a = pm_index(t, c.T, c.spp);
    unsafeAtomicAdd(acc + a, wc.x);
    unsafeAtomicAdd(acc + (uint64_t)c.T + a, wc.y);
    unsafeAtomicAdd(acc + 2ull * c.T + a, wc.z);
    unsafeAtomicAdd(acc + 3ull * c.T + a, wc.w);
    unsafeAtomicAdd(&wb.reveal[a], lr);
  }
}

// Composite (oitWeighted.frag.glsl:98-109) with the WEIGHTED_COMPOSITE blend
// (main.cpp:578-588): dst = src * (1 - reveal) + dst * reveal.
// The accumulators it read are set back to the clear values (0, and 0 for
// the log2 revealage) for the next frame, where they are not already, so
// the next clear need not fill them.
// Threads run over the accumulators in their (pixel-major) order.
__global__ __launch_bounds__(kBlock) void k_wboit_resolve(Common c, WBBuf wb, ResolveOut o) {
  const uint32_t a = blockIdx.x * kBlock + threadIdx.x;
  if (a >= c.T) return;
  const uint32_t t = pm_target(a, c.T, c.spp);
  float* acc = reinterpret_cast<float*>(wb.accum);
  const float4 A = make_float4(acc[a], acc[(uint64_t)c.T + a], acc[2ull * c.T + a], acc[3ull * c.T + a]);
  const float lr = wb.reveal[a];
  if ((__float_as_uint(A.x) | __float_as_uint(A.y) | __float_as_uint(A.z) | __float_as_uint(A.w) |
       __float_as_uint(lr)) != 0u) {
    acc[a] = 0.f;
    acc[(uint64_t)c.T + a] = 0.f;
    acc[2ull * c.T + a] = 0.f;
    acc[3ull * c.T + a] = 0.f;
    wb.reveal[a] = 0.f;
  }
  const float r = exp2f(lr);  // product of (1 - alpha), kept as a log2 sum
  const float den = fmaxf(A.w, 1e-5f);
  const float4 s = make_float4(A.x / den, A.y / den, A.z / den, r);
  const float om = 1.0f - r;
  const float4 bg = dest_of(c, t);
  o.out[t] = make_float4(s.x * om + bg.x * r, s.y * om + bg.y * r, s.z * om + bg.z * r, s.w * om + bg.w * r);
  if (o.counts) o.counts[t] = 0u;
}

// ---------------------------------------------------------------------------
// Tail lists, phase 2: push the frame's queued records onto per-target
// lists, unordered (the composite restores id order, tail_fold_from).
// ---------------------------------------------------------------------------

// Work item w of the grid is chunk (w mod 2) of 512 entries of queue
// region r0 + w / 2. Neighbouring entries of one target in a wave are
// chained in the wave and published with one exchange on the list head,
// whose old value becomes the chain's last link: no insert walks a list or
// waits for another fragment's link.
template <int kU>
__global__ __launch_bounds__(kBlock) void k_tail_link(TailBuf tb, uint32_t r0, uint32_t r1) {
  constexpr uint32_t kChunk = 64u * kU;
  constexpr uint32_t kPerRegion = (1u << kTailRegionShift) / kChunk;
  const uint32_t lane = lane_id();
  const uint64_t items = (uint64_t)(r1 - r0 + 1u) * kPerRegion;
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t w = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); w < items; w += waves) {
    const uint32_t reg = r0 + (uint32_t)(w / kPerRegion);
    const uint32_t base = (uint32_t)(w % kPerRegion) * kChunk;
    const uint32_t q = __builtin_amdgcn_readfirstlane(tb.qn[reg]);
    const uint32_t n = q < (1u << kTailRegionShift) ? q : (1u << kTailRegionShift);
    if (base >= n) continue;  // wave-uniform
    const uint64_t reg0 = (uint64_t)reg << kTailRegionShift;
    uint64_t rec[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const uint32_t pos = base + (uint32_t)u * 64u + lane;
      rec[u] = pos < n ? ld_stream(&tb.rec[reg0 + pos]) : kEmpty64;
    }
    uint32_t t[kU], v[kU], lo[kU];
    bool ins[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const bool tail = rec[u] != kEmpty64;
      t[u] = tail ? (uint32_t)(rec[u] >> 32) : kEmpty;
      v[u] = (uint32_t)rec[u];
      const Run r = same_target_run(t[u]);
      const uint32_t v_prev = __shfl_up(v[u], 1u);
      lo[u] = __shfl(v[u], (int)r.start_lane);
      if (tail && !r.start) st_relaxed(&tb.next[tail_index(v[u])], v_prev);
      ins[u] = tail && r.end;
    }
    uint32_t old[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u)
      old[u] = ins[u] ? __hip_atomic_exchange(&tb.head[t[u]], v[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (ins[u]) st_relaxed(&tb.next[tail_index(lo[u])], old[u]);
  }
}

// ---------------------------------------------------------------------------
// Opaque pass: the opaque spheres drawn before the transparent ones
// (oitRender.cpp:68-122) with the BlendMode::NONE pipeline (depth test LESS,
// depth write, no blending, main.cpp:526-533)
// ---------------------------------------------------------------------------

// Planes entry of the depth clear 1.0 (oitRender.cpp:91): no fragment id.
__device__ __forceinline__ uint64_t opaque_clear_z() { return ((uint64_t)depth_key(1.0f) << 32) | 0xFFFFFFFFull; }

__global__ __launch_bounds__(kBlock) void k_opaque_clear(uint64_t* __restrict__ z, float4* __restrict__ rgba,
                                                         uint64_t n, float4 bg) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t zc = opaque_clear_z();
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
    z[j] = zc;
    rgba[j] = bg;
  }
}

// Depth phase of an opaque draw: (depth key << 32 | id) atomicMin on each
// covered sample target, so the nearest fragment wins and, among equal
// depths, the earliest in draw order (a later one fails LESS). A fragment at
// or behind the clear depth writes nothing.
template <bool kCov>
__global__ __launch_bounds__(kBlock) void k_opaque_depth(Common c, FragView f, uint64_t* __restrict__ z) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= f.count) return;
  const uint32_t t = ld_stream(&f.target[i]);
  if (t >= c.T) return;
  const uint32_t key = depth_key(ld_stream(&f.depth[i]));
  if (key >= depth_key(1.0f)) return;
  const uint64_t v = ((uint64_t)key << 32) | (f.id_base + i);
  if constexpr (kCov) {
    const uint32_t m = ld_stream(&f.mask[i]) & ((1u << c.cov) - 1u);
    for (uint32_t s = 0; s < c.cov; ++s)
      if ((m >> s) & 1u)
        (void)__hip_atomic_fetch_min(&z[(uint64_t)s * c.T + t], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    (void)__hip_atomic_fetch_min(&z[t], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Colour phase: the fragment that won a sample target writes its colour
// there (blending off). Runs after its draw's depth phase; a later draw that
// takes the target over rewrites the colour in its own colour phase.
template <bool kCov>
__global__ __launch_bounds__(kBlock) void k_opaque_color(Common c, FragView f, const uint64_t* __restrict__ z,
                                                         float4* __restrict__ rgba) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= f.count) return;
  const uint32_t t = ld_stream(&f.target[i]);
  if (t >= c.T) return;
  const uint32_t key = depth_key(ld_stream(&f.depth[i]));
  if (key >= depth_key(1.0f)) return;
  const uint64_t v = ((uint64_t)key << 32) | (f.id_base + i);
  if constexpr (kCov) {
    const uint32_t m = ld_stream(&f.mask[i]) & ((1u << c.cov) - 1u);
    for (uint32_t s = 0; s < c.cov; ++s) {
      const uint64_t j = (uint64_t)s * c.T + t;
      if (((m >> s) & 1u) && z[j] == v) rgba[j] = ld_stream(&f.rgba[i]);
    }
  } else {
    if (z[t] == v) rgba[t] = ld_stream(&f.rgba[i]);
  }
}

// Display image: average of `S` sample planes and of each ss x ss block of a
// (W*ss) x (H*ss) premultiplied image (MSAA resolve, main.cpp:700-714;
// supersample downsample, :715-731), summed in sample, row, column order;
// then sRGB8 rgb (exact thresholds) and linear 8-bit alpha, RGBA bytes.
__global__ __launch_bounds__(kBlock) void k_finalize_srgb8(const float4* __restrict__ in, uint32_t W, uint32_t H,
                                                           uint32_t S, uint32_t ss, const CodecTables* tab,
                                                           uint32_t* __restrict__ out) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, tab);
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= W * H) return;
  const uint32_t x = p % W, y = p / W;
  const uint64_t Wf = (uint64_t)W * ss, Hf = (uint64_t)H * ss;
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t s = 0; s < S; ++s)
    for (uint32_t dy = 0; dy < ss; ++dy)
      for (uint32_t dx = 0; dx < ss; ++dx) {
        const float4 v = in[(s * Hf + (uint64_t)y * ss + dy) * Wf + (uint64_t)x * ss + dx];
        sum = make_float4(sum.x + v.x, sum.y + v.y, sum.z + v.z, sum.w + v.w);
      }
  const float n = (float)(S * ss * ss);
  const float a = fminf(fmaxf(sum.w / n, 0.0f), 1.0f);
  out[p] = code_rgb_fast(sum.x / n, s_thr) | (code_rgb_fast(sum.y / n, s_thr) << 8) |
           (code_rgb_fast(sum.z / n, s_thr) << 16) | ((uint32_t)floorf(a * 255.0f + 0.5f) << 24);
}

// Device self-check of the branch-free encode against the exact table search
// over a range of binary32 bit patterns.
__global__ __launch_bounds__(kBlock) void k_codec_selftest(const CodecTables* tab, uint32_t lo, uint32_t hi,
                                                           unsigned long long* mismatches, uint32_t* first_bad) {
  __shared__ float s_thr[257];
  stage_thresholds(s_thr, tab);
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  unsigned long long bad = 0;
  uint32_t first = 0xFFFFFFFFu;
  for (uint64_t b = (uint64_t)lo + (uint64_t)blockIdx.x * kBlock + threadIdx.x; b < hi; b += stride) {
    const float c = __uint_as_float((uint32_t)b);
    if (code_rgb_fast(c, s_thr) != code_rgb(c, tab->thr)) {
      ++bad;
      if ((uint32_t)b < first) first = (uint32_t)b;
    }
  }
  if (bad) {
    atomicAdd(mismatches, bad);
    atomicMin(first_bad, first);
  }
}

template <template <int> class F, typename... A>
hipError_t dispatch_k(uint32_t K, A&&... args) {
  switch (K) {
    case 1: return F<1>::run(args...);
    case 2: return F<2>::run(args...);
    case 4: return F<4>::run(args...);
    case 8: return F<8>::run(args...);
    case 16: return F<16>::run(args...);
    case 32: return F<32>::run(args...);
    default: return hipErrorInvalidValue;
  }
}

template <int K>
struct KBufResolve {
  static hipError_t run(const Common& c, const KBuf& kb, const TailBuf& tb, const ResolveOut& o, hipStream_t s) {
    const bool dbg = o.fk_key || o.fk_val || o.counts;
    const bool ctail = c.tail && tb.head != nullptr;
    const dim3 g(blocks_for(c.T)), b(kBlock);
#define OIT_KB(CT, DBG) hipLaunchKernelGGL((k_kbuf_resolve<K, CT, DBG>), g, b, 0, s, c, kb, tb, o)
    if (ctail) { if (dbg) OIT_KB(true, true); else OIT_KB(true, false); }
    else { if (dbg) OIT_KB(false, true); else OIT_KB(false, false); }
#undef OIT_KB
    return hipGetLastError();
  }
};

template <int K>
struct SpinStore {
  static hipError_t run(const Common& c, const FragView& f, const KBuf& kb, const TailBuf& tb, hipStream_t s) {
    const dim3 g(blocks_for(f.count)), b(kBlock);
    if (c.tail && tb.head) hipLaunchKernelGGL((k_spin_store<K, true>), g, b, 0, s, c, f, kb, tb);
    else hipLaunchKernelGGL((k_spin_store<K, false>), g, b, 0, s, c, f, kb, tb);
    return hipGetLastError();
  }
};

template <int K>
struct Loop64Store {
  static hipError_t run(const Common& c, const FragView& f, uint64_t* slots, uint32_t* cnt, uint64_t* big,
                        int chunks, hipStream_t s) {
    if (f.count == 0) return hipSuccess;
    const int u = chunks >= 4 ? 4 : (chunks >= 2 ? 2 : 1);
    const uint64_t per_block = (uint64_t)kBlock * u;
    const dim3 g((uint32_t)((f.count + per_block - 1) / per_block)), b(kBlock);
    if (u == 4) hipLaunchKernelGGL((k_loop64_store_m<K, 4>), g, b, 0, s, c, f, slots, cnt, big);
    else if (u == 2) hipLaunchKernelGGL((k_loop64_store_m<K, 2>), g, b, 0, s, c, f, slots, cnt, big);
    else hipLaunchKernelGGL((k_loop64_store_m<K, 1>), g, b, 0, s, c, f, slots, cnt, big);
    return hipGetLastError();
  }
};

template <int K>
struct Loop64Classify {
  static hipError_t run(const Common& c, const FragView& f, const uint64_t* slots, const uint64_t* big,
                        const TailBuf& tb, hipStream_t s) {
    if (f.count == 0) return hipSuccess;
    constexpr int kU = 4;  // 64-fragment chunks per wave
    const uint32_t blocks = (uint32_t)((f.count + kBlock * kU - 1) / (kBlock * kU));
    hipLaunchKernelGGL((k_loop64_classify<K, kU>), dim3(blocks), dim3(kBlock), 0, s, c, f, slots, big, tb);
    return hipGetLastError();
  }
};

template <int K>
struct Loop64Resolve {
  static hipError_t run(const Common& c, uint64_t* slots, uint32_t* cnt, const TailBuf& tb, const ResolveOut& o,
                        hipStream_t s) {
    const bool dbg = o.fk_key || o.fk_val || o.counts;
    const bool ctail = c.tail && tb.head != nullptr;
    const dim3 g(blocks_for(c.T)), b(kBlock);
#define OIT_L64(CT, DBG) hipLaunchKernelGGL((k_loop64_resolve<K, CT, DBG>), g, b, 0, s, c, slots, cnt, tb, o)
    if (ctail) { if (dbg) OIT_L64(true, true); else OIT_L64(true, false); }
    else { if (dbg) OIT_L64(false, true); else OIT_L64(false, false); }
#undef OIT_L64
    return hipGetLastError();
  }
};

template <int K>
struct LLResolveReset {
  static hipError_t run(const Common& c, const LLBuf& ll, const TailBuf& tb, const ResolveOut& o, bool pair,
                        bool keys_only, hipStream_t s) {
    if (keys_only) return run_keys(c, ll, tb, o, pair, s);
    const bool dbg = o.fk_key || o.fk_val || o.counts;
    const bool ctail = tb.head != nullptr;
    uint32_t blocks = blocks_for(c.T);
    if (blocks > 8192u) blocks = 8192u;
    if (blocks == 0) blocks = 1;
    const dim3 g(blocks), b(kBlock);
#define OIT_LLR(TAIL, CT, DBG, PAIR) \
  hipLaunchKernelGGL((k_ll_resolve_reset<K, TAIL, CT, DBG, PAIR>), g, b, 0, s, c, ll, tb, o)
#define OIT_LLR_P(TAIL, CT, DBG) \
  do { if (pair) OIT_LLR(TAIL, CT, DBG, true); else OIT_LLR(TAIL, CT, DBG, false); } while (0)
    if (c.tail) {
      if (ctail) { if (dbg) OIT_LLR_P(true, true, true); else OIT_LLR_P(true, true, false); }
      else { if (dbg) OIT_LLR_P(true, false, true); else OIT_LLR_P(true, false, false); }
    } else {
      if (dbg) OIT_LLR_P(false, false, true); else OIT_LLR_P(false, false, false);
    }
#undef OIT_LLR_P
#undef OIT_LLR
    return hipGetLastError();
  }
  static hipError_t run_keys(const Common& c, const LLBuf& ll, const TailBuf& tb, const ResolveOut& o, bool pair,
                             hipStream_t s) {
    const bool dbg = o.fk_key || o.fk_val || o.counts;
    const bool ctail = tb.head != nullptr;
    uint32_t blocks = blocks_for(c.T);
    if (blocks > 8192u) blocks = 8192u;
    if (blocks == 0) blocks = 1;
    const dim3 g(blocks), b(kBlock);
#define OIT_LLK(TAIL, CT, DBG, PAIR) \
  hipLaunchKernelGGL((k_ll_resolve_reset_keys<K, TAIL, CT, DBG, PAIR>), g, b, 0, s, c, ll, tb, o)
#define OIT_LLK_P(TAIL, CT, DBG) \
  do { if (pair) OIT_LLK(TAIL, CT, DBG, true); else OIT_LLK(TAIL, CT, DBG, false); } while (0)
    if (c.tail) {
      if (ctail) { if (dbg) OIT_LLK_P(true, true, true); else OIT_LLK_P(true, true, false); }
      else { if (dbg) OIT_LLK_P(true, false, true); else OIT_LLK_P(true, false, false); }
    } else {
      if (dbg) OIT_LLK_P(false, false, true); else OIT_LLK_P(false, false, false);
    }
#undef OIT_LLK_P
#undef OIT_LLK
    return hipGetLastError();
  }
};

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t want = (n / 4 + kBlock - 1) / kBlock;
  const uint32_t g = (uint32_t)(want < 4096 ? (want ? want : 1) : 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_ranges(const FillRange* r, int count, hipStream_t s) {
  FillArgs a{};
  int m = 0;
  a.start[0] = 0;
  for (int i = 0; i < count; ++i) {
    if (r[i].n == 0) continue;
    if (m == kMaxFillRanges || (reinterpret_cast<uintptr_t>(r[i].p) & 15u)) {  // not packable: a launch of its own
      const hipError_t e = launch_fill_u32(r[i].p, r[i].n, r[i].v, s);
      if (e != hipSuccess) return e;
      continue;
    }
    a.p[m] = r[i].p;
    a.n[m] = r[i].n;
    a.v[m] = r[i].v;
    a.start[m + 1] = a.start[m] + r[i].n / 4;
    ++m;
  }
  if (m == 0) return hipSuccess;
  a.count = m;
  for (int i = m + 1; i <= kMaxFillRanges; ++i) a.start[i] = a.start[m];
  uint64_t want = (a.start[m] + kBlock - 1) / kBlock;
  if (want < 1) want = 1;  // the remainders
  const uint32_t g = (uint32_t)(want < 4096 ? want : 4096);
  hipLaunchKernelGGL(k_fill_ranges, dim3(g), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ll_store(const Common& c, const FragView& f, const LLBuf& ll, const TailBuf& tb, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  const dim3 g(blocks_for(f.count)), b(kBlock);
  if (c.tail && tb.head) hipLaunchKernelGGL((k_ll_store<true>), g, b, 0, s, c, f, ll, tb);
  else hipLaunchKernelGGL((k_ll_store<false>), g, b, 0, s, c, f, ll, tb);
  return hipGetLastError();
}

hipError_t launch_ll_link(const Common& c, const FragView& f, const LLBuf& ll, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ll_link, dim3(blocks_for(f.count)), dim3(kBlock), 0, s, c, f, ll);
  return hipGetLastError();
}

hipError_t launch_ll_append(const Common& c, const FragView& f, const LLBuf& ll, const TailBuf& tb, int chunks,
                            hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  const int u = chunks >= 8 ? 8 : 4;
  const uint64_t per_block = (uint64_t)kBlock * u;
  const dim3 g((uint32_t)((f.count + per_block - 1) / per_block)), b(kBlock);
  const bool tail = c.tail && tb.head;
  if (u == 8) {
    if (tail) hipLaunchKernelGGL((k_ll_append<true, 8>), g, b, 0, s, c, f, ll, tb);
    else hipLaunchKernelGGL((k_ll_append<false, 8>), g, b, 0, s, c, f, ll, tb);
  } else {
    if (tail) hipLaunchKernelGGL((k_ll_append<true, 4>), g, b, 0, s, c, f, ll, tb);
    else hipLaunchKernelGGL((k_ll_append<false, 4>), g, b, 0, s, c, f, ll, tb);
  }
  return hipGetLastError();
}

hipError_t launch_ll_resolve_reset(const Common& c, const LLBuf& ll, const TailBuf& tb, const ResolveOut& o,
                                   bool pair_walk, bool keys_only, hipStream_t s) {
  return dispatch_k<LLResolveReset>(c.K, c, ll, tb, o, pair_walk, keys_only, s);
}

hipError_t launch_simple_store(const Common& c, const FragView& f, const KBuf& kb, const TailBuf& tb, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  const dim3 g(blocks_for(f.count)), b(kBlock);
  if (c.tail && tb.head) hipLaunchKernelGGL((k_simple_store<true>), g, b, 0, s, c, f, kb, tb);
  else hipLaunchKernelGGL((k_simple_store<false>), g, b, 0, s, c, f, kb, tb);
  return hipGetLastError();
}

hipError_t launch_spin_store(const Common& c, const FragView& f, const KBuf& kb, const TailBuf& tb, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  return dispatch_k<SpinStore>(c.K, c, f, kb, tb, s);
}

hipError_t launch_kbuf_resolve(const Common& c, const KBuf& kb, const TailBuf& tb, const ResolveOut& o, hipStream_t s) {
  return dispatch_k<KBufResolve>(c.K, c, kb, tb, o, s);
}

hipError_t launch_loop64_count(const Common& c, const FragView& f, uint32_t* cnt, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  constexpr int kU = 4;
  const uint32_t blocks = (uint32_t)((f.count + kBlock * kU - 1) / (kBlock * kU));
  hipLaunchKernelGGL((k_loop64_count<kU>), dim3(blocks), dim3(kBlock), 0, s, c, f, cnt);
  return hipGetLastError();
}

hipError_t launch_loop64_store(const Common& c, const FragView& f, uint64_t* slots, uint32_t* cnt, uint64_t* big,
                               int chunks, hipStream_t s) {
  return dispatch_k<Loop64Store>(c.K, c, f, slots, cnt, big, chunks, s);
}

hipError_t launch_loop64_classify(const Common& c, const FragView& f, const uint64_t* slots, const uint64_t* big,
                                  const TailBuf& tb, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  return dispatch_k<Loop64Classify>(c.K, c, f, slots, big, tb, s);
}

hipError_t launch_loop64_resolve(const Common& c, uint64_t* slots, uint32_t* cnt, const TailBuf& tb,
                                 const ResolveOut& o, hipStream_t s) {
  return dispatch_k<Loop64Resolve>(c.K, c, slots, cnt, tb, o, s);
}

hipError_t launch_wboit_store(const Common& c, const FragView& f, const WBBuf& wb, float view_near, float view_span,
                              hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_wboit_store, dim3(blocks_for(f.count)), dim3(kBlock), 0, s, c, f, wb, view_near, view_span);
  return hipGetLastError();
}

hipError_t launch_wboit_resolve(const Common& c, const WBBuf& wb, const ResolveOut& o, hipStream_t s) {
  hipLaunchKernelGGL(k_wboit_resolve, dim3(blocks_for(c.T)), dim3(kBlock), 0, s, c, wb, o);
  return hipGetLastError();
}

hipError_t launch_codec_selftest(const CodecTables* tab, uint32_t lo, uint32_t hi, unsigned long long* mismatches,
                                 uint32_t* first_bad, hipStream_t s) {
  hipLaunchKernelGGL(k_codec_selftest, dim3(4096), dim3(kBlock), 0, s, tab, lo, hi, mismatches, first_bad);
  return hipGetLastError();
}

hipError_t launch_tail_link(const TailBuf& tb, uint32_t first, uint32_t frags, hipStream_t s) {
  const uint32_t n = frags < tb.cap ? frags : tb.cap;
  if (first >= n || !tb.qn) return hipSuccess;
  // a wave per 512 entries of a region of 1024 ids
  const uint32_t r0 = first >> kTailRegionShift, r1 = (n - 1u) >> kTailRegionShift;
  constexpr uint32_t kPerRegion = (1u << kTailRegionShift) / 512u;
  uint64_t blocks = ((uint64_t)(r1 - r0 + 1u) * kPerRegion + kBlock / 64 - 1) / (kBlock / 64);
  if (blocks > 8192u) blocks = 8192u;
  hipLaunchKernelGGL(k_tail_link<8>, dim3((uint32_t)blocks), dim3(kBlock), 0, s, tb, r0, r1);
  return hipGetLastError();
}

namespace {

template <int K>
struct BatchResolve {
  static hipError_t run(const BatchArgs& a, const LLBuf& ll, bool tail, bool pair, bool keys, uint32_t grid_cap,
                        hipStream_t s) {
    if (a.tiles == 0) return hipSuccess;
    uint32_t blocks = a.tiles;
    if (grid_cap && blocks > grid_cap) blocks = grid_cap;
    const dim3 g(blocks), b(kBlock);
#define OIT_BR(TAIL, PAIR, KEYS) hipLaunchKernelGGL((k_batch_resolve<K, TAIL, PAIR, KEYS>), g, b, 0, s, a, ll)
#define OIT_BR_K(TAIL, PAIR) do { if (keys) OIT_BR(TAIL, PAIR, true); else OIT_BR(TAIL, PAIR, false); } while (0)
#define OIT_BR_P(TAIL) do { if (pair) OIT_BR_K(TAIL, true); else OIT_BR_K(TAIL, false); } while (0)
    if (tail) OIT_BR_P(true);
    else OIT_BR_P(false);
#undef OIT_BR_P
#undef OIT_BR_K
#undef OIT_BR
    return hipGetLastError();
  }
};

}  // namespace

hipError_t launch_batch_append(const BatchArgs& a, const LLBuf& ll, uint32_t chunks, hipStream_t s) {
  if (a.units == 0) return hipSuccess;
  const dim3 g((a.units + kBlock / 64 - 1) / (kBlock / 64)), b(kBlock);
  if (chunks == 8) hipLaunchKernelGGL((k_batch_append<8>), g, b, 0, s, a, ll);
  else if (chunks == 4) hipLaunchKernelGGL((k_batch_append<4>), g, b, 0, s, a, ll);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_bin_prepare(const BatchArgs& a, const LLBuf& ll, uint32_t tiles, uint32_t chunks, hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin_tile_sums, dim3(tiles), dim3(kBlock), 0, s, a, ll.head);
  hipLaunchKernelGGL(k_bin_frame_scan, dim3(a.f1 - a.f0), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(k_bin_tile_apply, dim3(tiles), dim3(kBlock), 0, s, a, ll.head);
  if (a.units) {
    const dim3 g((a.units + kBlock / 64 - 1) / (kBlock / 64)), b(kBlock);
    if (chunks == 8) hipLaunchKernelGGL((k_bin_scatter<8>), g, b, 0, s, a, ll);
    else hipLaunchKernelGGL((k_bin_scatter<4>), g, b, 0, s, a, ll);
  }
  hipLaunchKernelGGL(k_bin_sort, dim3(2048), dim3(kBlock), 0, s, a, ll);
  return hipGetLastError();
}

template <int K>
struct BinResolve {
  static hipError_t run(const BatchArgs& a, const LLBuf& ll, bool tail, uint32_t grid_cap, hipStream_t s) {
    if (a.tiles == 0) return hipSuccess;
    uint32_t blocks = a.tiles;
    if (grid_cap && blocks > grid_cap) blocks = grid_cap;
    if (tail) hipLaunchKernelGGL((k_bin_resolve<K, true>), dim3(blocks), dim3(kBlock), 0, s, a, ll);
    else hipLaunchKernelGGL((k_bin_resolve<K, false>), dim3(blocks), dim3(kBlock), 0, s, a, ll);
    return hipGetLastError();
  }
};

hipError_t launch_bin_resolve(const BatchArgs& a, const LLBuf& ll, uint32_t K, bool tail, uint32_t grid_cap,
                              hipStream_t s) {
  return dispatch_k<BinResolve>(K, a, ll, tail, grid_cap, s);
}

hipError_t launch_batch_resolve(const BatchArgs& a, const LLBuf& ll, uint32_t K, bool tail, bool pair, bool keys,
                                uint32_t grid_cap, hipStream_t s) {
  return dispatch_k<BatchResolve>(K, a, ll, tail, pair, keys, grid_cap, s);
}

namespace {

template <int K>
struct LLResolveCov {
  static hipError_t run(const Common& c, const LLBuf& ll, const TailBuf& tb, const ResolveOut& o, hipStream_t s) {
    uint32_t blocks = blocks_for(c.T);
    if (blocks > 8192u) blocks = 8192u;
    if (blocks == 0) blocks = 1;
    if (c.tail) hipLaunchKernelGGL((k_ll_resolve_cov<K, true>), dim3(blocks), dim3(kBlock), 0, s, c, ll, tb, o);
    else hipLaunchKernelGGL((k_ll_resolve_cov<K, false>), dim3(blocks), dim3(kBlock), 0, s, c, ll, tb, o);
    return hipGetLastError();
  }
};

template <int K>
struct KBufResolveCov {
  static hipError_t run(const Common& c, const KBuf& kb, const TailBuf& tb, const ResolveOut& o, hipStream_t s) {
    hipLaunchKernelGGL((k_kbuf_resolve_cov<K>), dim3(blocks_for(c.T)), dim3(kBlock), 0, s, c, kb, tb, o);
    return hipGetLastError();
  }
};

}  // namespace

hipError_t launch_ll_resolve_cov(const Common& c, const LLBuf& ll, const TailBuf& tb, const ResolveOut& o,
                                 hipStream_t s) {
  if (c.cov < 2u || c.cov > 8u || !ll.mask) return hipErrorInvalidValue;
  return dispatch_k<LLResolveCov>(c.K, c, ll, tb, o, s);
}

hipError_t launch_kbuf_resolve_cov(const Common& c, const KBuf& kb, const TailBuf& tb, const ResolveOut& o,
                                   hipStream_t s) {
  if (c.cov < 2u || c.cov > 8u || !kb.mask) return hipErrorInvalidValue;
  return dispatch_k<KBufResolveCov>(c.K, c, kb, tb, o, s);
}

namespace {

template <int K>
struct Loop32Depth {
  static hipError_t run(const Common& c, const FragView& f, uint32_t* zs, hipStream_t s) {
    if (f.count == 0) return hipSuccess;
    constexpr int kU = 2;  // 64-fragment chunks per wave
    const uint32_t blocks = (uint32_t)((f.count + kBlock * kU - 1) / (kBlock * kU));
    hipLaunchKernelGGL((k_loop32_depth_m<K, kU>), dim3(blocks), dim3(kBlock), 0, s, c, f, zs);
    return hipGetLastError();
  }
};

template <int K>
struct Loop32Color {
  static hipError_t run(const Common& c, const FragView& f, const uint32_t* zs, uint64_t* cs, const TailBuf& tb,
                        hipStream_t s) {
    if (f.count == 0) return hipSuccess;
    constexpr int kU = 2;  // 64-fragment chunks per wave
    const dim3 g((uint32_t)((f.count + kBlock * kU - 1) / (kBlock * kU))), b(kBlock);
    if (c.tail && tb.head) hipLaunchKernelGGL((k_loop32_color_m<K, true, kU>), g, b, 0, s, c, f, zs, cs, tb);
    else hipLaunchKernelGGL((k_loop32_color_m<K, false, kU>), g, b, 0, s, c, f, zs, cs, tb);
    return hipGetLastError();
  }
};

template <int K>
struct Loop32Resolve {
  static hipError_t run(const Common& c, uint32_t* zs, uint64_t* cs, const TailBuf& tb, const ResolveOut& o,
                        hipStream_t s) {
    const bool dbg = o.fk_key || o.fk_val || o.counts;
    const bool ctail = c.tail && tb.head != nullptr;
    const dim3 g(blocks_for(c.T)), b(kBlock);
#define OIT_L32(CT, DBG) hipLaunchKernelGGL((k_loop32_resolve<K, CT, DBG>), g, b, 0, s, c, zs, cs, tb, o)
    if (ctail) { if (dbg) OIT_L32(true, true); else OIT_L32(true, false); }
    else { if (dbg) OIT_L32(false, true); else OIT_L32(false, false); }
#undef OIT_L32
    return hipGetLastError();
  }
};

template <int K>
struct InterlockResolve {
  static hipError_t run(const Common& c, const LLBuf& ll, const ResolveOut& o, hipStream_t s) {
    uint32_t blocks = blocks_for(c.T);
    if (blocks > 8192u) blocks = 8192u;
    if (blocks == 0) blocks = 1;
    if (c.cov > 1u)
      hipLaunchKernelGGL((k_interlock_resolve<K, true>), dim3(blocks), dim3(kBlock), 0, s, c, ll, o);
    else
      hipLaunchKernelGGL((k_interlock_resolve<K, false>), dim3(blocks), dim3(kBlock), 0, s, c, ll, o);
    return hipGetLastError();
  }
};

}  // namespace

hipError_t launch_loop32_depth(const Common& c, const FragView& f, uint32_t* zs, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  return dispatch_k<Loop32Depth>(c.K, c, f, zs, s);
}

hipError_t launch_loop32_color(const Common& c, const FragView& f, const uint32_t* zs, uint64_t* cs,
                               const TailBuf& tb, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  return dispatch_k<Loop32Color>(c.K, c, f, zs, cs, tb, s);
}

hipError_t launch_loop32_resolve(const Common& c, uint32_t* zs, uint64_t* cs, const TailBuf& tb, const ResolveOut& o,
                                 hipStream_t s) {
  return dispatch_k<Loop32Resolve>(c.K, c, zs, cs, tb, o, s);
}

hipError_t launch_interlock_resolve(const Common& c, const LLBuf& ll, const ResolveOut& o, hipStream_t s) {
  if (!ll.full || (c.cov > 1u && !ll.mask)) return hipErrorInvalidValue;
  return dispatch_k<InterlockResolve>(c.K, c, ll, o, s);
}

hipError_t launch_opaque_clear(uint64_t* z, float4* rgba, uint64_t n, float4 bg, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 8192u) blocks = 8192u;
  hipLaunchKernelGGL(k_opaque_clear, dim3((uint32_t)blocks), dim3(kBlock), 0, s, z, rgba, n, bg);
  return hipGetLastError();
}

hipError_t launch_opaque_draw(const Common& c, const FragView& f, uint64_t* z, float4* rgba, hipStream_t s) {
  if (f.count == 0) return hipSuccess;
  const dim3 g(blocks_for(f.count)), b(kBlock);
  if (c.cov > 1u) {
    if (c.cov > 8u || !f.mask) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_opaque_depth<true>, g, b, 0, s, c, f, z);
    hipLaunchKernelGGL(k_opaque_color<true>, g, b, 0, s, c, f, z, rgba);
  } else {
    hipLaunchKernelGGL(k_opaque_depth<false>, g, b, 0, s, c, f, z);
    hipLaunchKernelGGL(k_opaque_color<false>, g, b, 0, s, c, f, z, rgba);
  }
  return hipGetLastError();
}

hipError_t launch_finalize_srgb8(const float4* in, uint32_t W, uint32_t H, uint32_t S, uint32_t ss,
                                 const CodecTables* tab, uint32_t* out, hipStream_t s) {
  const uint64_t P = (uint64_t)W * H;
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize_srgb8, dim3(blocks_for(P)), dim3(kBlock), 0, s, in, W, H, S, ss, tab, out);
  return hipGetLastError();
}

}  // namespace oit
