// oit_batch.hip - host runtime of the batched Linked List (oit_batch_*).
//
// Many frames share one resident node pool and one head array. The plan
// orders frames by (K, tail_blend) so that each group's targets form one
// contiguous head range, pads every frame's fragments to 512 (an append unit
// never spans two frames) and its targets to 256 (a resolve tile never spans
// two frames), and lays the frame table out for the kernels' scalar binary
// search. A run is: one append launch over every frame, then one resolve
// launch per group (optionally on forked streams, optionally replayed from a
// captured hipGraph). The resolve leaves heads at 0 and walked links at
// kUnset, so the pool is clean for the next run without a clear pass.
//
// The per-frame semantics are those of the single-frame Linked List
// (oitLinkedList.frag.glsl:51-85 append, :106-172 composite); a pool sized
// from the exact fragment counts never overflows, so no colour-pass tail.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "oit.h"
#include "oit_kernels.hpp"

using oit::BatchArgs;
using oit::BatchFrameDev;
using oit::LLBuf;

namespace {

constexpr uint32_t kTile = 256;        // resolve tile (one block of targets)
constexpr uint32_t kFragAlign = 512;   // append unit at 8 chunks per wave (one schedule slice)
constexpr uint32_t kRounds = 1024;     // rounds of the interleaved append schedule
constexpr uint64_t kMaxPoolSlots = 0xFFFFF000ull;  // node indices stay below kUnset
constexpr uint64_t kMaxHeads = 0xFFFF0000ull;  // the head array, rounded up to 2^16, stays below 2^32
constexpr uint64_t kMaxFrameFragments = 0x7FFFFFFEull;
constexpr int kSide = 3;               // forked streams for concurrent groups
constexpr uint32_t kDefaultBlocks = 8192;
// fragments per target from which a frame is binned: none by default (on
// the 64-frame mix every binned frame ran slower than its lists; the tuner
// sweeps the threshold)
constexpr float kDefaultBinDensity = 3.0e38f;

int st(hipError_t e) {
  if (e == hipSuccess) return OIT_OK;
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory) return OIT_ERR_OOM;
  if (e == hipErrorInvalidValue) return OIT_ERR_INVALID_ARG;
  return OIT_ERR_DEVICE;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

int k_index(uint32_t K) {
  switch (K) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    case 16: return 4;
    case 32: return 5;
    default: return -1;
  }
}

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

struct Group {
  uint32_t K, tail, f0, f1, tile0, tiles;
  bool bin;  // holds binned frames
};

struct Plan {
  std::vector<BatchFrameDev> table;  // group order
  std::vector<uint32_t> sched;       // interleaved append schedule (BatchArgs::sched)
  uint32_t rounds = 0;
  std::vector<Group> groups;
  uint64_t pool_slots = 0;  // padded fragments (nodes without the terminator)
  uint64_t heads = 0;       // padded targets
  bool bin = false;         // some frame is binned
};

bool same_table(const std::vector<BatchFrameDev>& a, const std::vector<BatchFrameDev>& b) {
  return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(BatchFrameDev)) == 0);
}

// Frames with at least `bin_density` fragments per target are binned
// (bucket per target) instead of linked (list per target).
int make_plan(const oit_batch_frame* fr, uint32_t n, Plan& p, float bin_density) {
  if (n > 0 && !fr) return OIT_ERR_INVALID_ARG;
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) {
    const oit_batch_frame& f = fr[i];
    if (f.width == 0 || f.height == 0) return OIT_ERR_INVALID_ARG;
    if (!(f.samples == 1 || f.samples == 2 || f.samples == 4 || f.samples == 8)) return OIT_ERR_INVALID_ARG;
    if (k_index(f.layers) < 0 || f.tail_blend > 1u || f.reserved != 0u) return OIT_ERR_INVALID_ARG;
    if ((uint64_t)f.width * f.height * f.samples > 0xFFFFFFFFull) return OIT_ERR_INVALID_ARG;
    if (!f.out || ((uintptr_t)f.out & 15u)) return OIT_ERR_INVALID_ARG;
    if (f.frags.count > kMaxFrameFragments) return OIT_ERR_CAPACITY;
    if (f.frags.coverage) return OIT_ERR_UNSUPPORTED;  // batches hold sample-shaded or single-sample frames
    if (f.frags.count > 0) {
      if (!f.frags.target || !f.frags.depth || !f.frags.rgba) return OIT_ERR_INVALID_ARG;
      if ((uintptr_t)f.frags.rgba & 15u) return OIT_ERR_INVALID_ARG;
    }
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (fr[a].layers != fr[b].layers) return fr[a].layers < fr[b].layers;
    return fr[a].tail_blend < fr[b].tail_blend;
  });
  p.table.clear();
  p.groups.clear();
  uint64_t frag = 0, tgt = 0;
  for (uint32_t j = 0; j < n; ++j) {
    const oit_batch_frame& f = fr[order[j]];
    BatchFrameDev d;
    std::memset(&d, 0, sizeof d);
    d.target = f.frags.target;
    d.depth = f.frags.depth;
    d.rgba = reinterpret_cast<const float4*>(f.frags.rgba);
    d.out = reinterpret_cast<float4*>(f.out);
    d.background = make_float4(f.background[0], f.background[1], f.background[2], f.background[3]);
    d.count = (uint32_t)f.frags.count;
    d.T = f.width * f.height * f.samples;
    d.target_base = (uint32_t)tgt;
    d.frag_base = (uint32_t)frag;
    d.mode = (f.frags.count > 0 && (double)f.frags.count >= (double)bin_density * (double)d.T) ? oit::kModeBin
                                                                                                : oit::kModeList;
    frag += round_up(f.frags.count, kFragAlign);
    tgt += round_up(d.T, kTile);
    if (frag > kMaxPoolSlots || tgt > kMaxHeads) return OIT_ERR_CAPACITY;
    if (p.groups.empty() || p.groups.back().K != f.layers || p.groups.back().tail != f.tail_blend) {
      if (!p.groups.empty()) p.groups.back().f1 = j;
      Group g;
      g.K = f.layers;
      g.tail = f.tail_blend;
      g.f0 = j;
      g.f1 = j + 1;
      g.tile0 = d.target_base / kTile;
      g.tiles = 0;
      g.bin = false;
      p.groups.push_back(g);
    }
    if (d.mode == oit::kModeBin) {
      p.groups.back().bin = true;
      p.bin = true;
    }
    p.table.push_back(d);
  }
  if (!p.groups.empty()) p.groups.back().f1 = n;
  for (size_t g = 0; g < p.groups.size(); ++g) {
    const uint64_t end = g + 1 < p.groups.size() ? (uint64_t)p.groups[g + 1].tile0 : tgt / kTile;
    p.groups[g].tiles = (uint32_t)(end - p.groups[g].tile0);
  }
  p.pool_slots = frag;
  p.heads = tgt;
  // interleaved append: up to kRounds rounds; round q takes slices
  // floor(q S / R) .. floor((q + 1) S / R) - 1 of every frame
  uint64_t max_slices = 0;
  std::vector<uint64_t> sl(n);
  for (uint32_t j = 0; j < n; ++j) {
    sl[j] = (p.table[j].count + oit::kBatchSlice - 1) / oit::kBatchSlice;
    max_slices = std::max(max_slices, sl[j]);
  }
  const uint32_t R = (uint32_t)std::min<uint64_t>(kRounds, std::max<uint64_t>(max_slices, 1));
  p.rounds = R;
  p.sched.assign((size_t)R + 1 + (size_t)R * (n + 1), 0u);
  for (uint32_t q = 0; q <= R; ++q) {
    uint64_t start = 0;
    for (uint32_t j = 0; j < n; ++j) start += sl[j] * q / R;
    p.sched[q] = (uint32_t)start;
  }
  for (uint32_t q = 0; q < R; ++q) {
    uint32_t* pre = p.sched.data() + R + 1 + (size_t)q * (n + 1);
    uint64_t acc = 0;
    for (uint32_t j = 0; j < n; ++j) {
      pre[j] = (uint32_t)acc;
      acc += sl[j] * (q + 1) / R - sl[j] * q / R;
    }
    pre[n] = (uint32_t)acc;
  }
  return OIT_OK;
}

oit_batch_tuning default_tuning() {
  oit_batch_tuning t;
  std::memset(&t, 0, sizeof t);
  t.append_chunks = 4;
  for (int k = 0; k < 6; ++k)
    for (int tl = 0; tl < 2; ++tl) t.resolve_blocks[k][tl] = kDefaultBlocks;
  t.bin_density = kDefaultBinDensity;
  return t;
}

bool tuning_valid(const oit_batch_tuning& t) {
  if (t.append_chunks != 4 && t.append_chunks != 8) return false;
  if (!(t.bin_density >= 0.f)) return false;  // NaN too
  for (int k = 0; k < 6; ++k)
    for (int tl = 0; tl < 2; ++tl)
      if (t.resolve_variant[k][tl] & ~(OIT_VARIANT_LL_SINGLE_WALK | OIT_VARIANT_LL_KEYS_ONLY)) return false;
  return true;
}

}  // namespace

struct oit_batch {
  int device = 0;
  uint32_t flags = 0;
  oit::CodecTables* d_tab = nullptr;
  // resident pool
  uint32_t* head = nullptr;
  uint64_t* cd = nullptr;
  uint32_t* next = nullptr;
  uint32_t* counters = nullptr;  // [0] overflow (stays 0: pools are exact)
  uint64_t cap_nodes = 0;        // allocated nodes, terminator included
  uint64_t cap_heads = 0;
  // binned frames: bucket cursor and deep-bucket queue per head, scan base
  // per tile, queue lengths per group
  uint32_t* cursor = nullptr;
  uint32_t* deep_q = nullptr;
  uint32_t* tile_sum = nullptr;
  uint32_t* deep_n = nullptr;
  bool dirty = false;            // a run did not complete: re-initialise before the next
  // frame table
  BatchFrameDev* d_table = nullptr;
  BatchFrameDev* h_table = nullptr;  // pinned staging
  uint32_t table_cap = 0;
  uint32_t* d_sched = nullptr;  // interleaved append schedule, uploaded with the table
  uint32_t* h_sched = nullptr;
  size_t sched_cap = 0;
  hipEvent_t table_ev = nullptr;
  bool table_ev_pending = false;
  std::vector<BatchFrameDev> uploaded;
  bool uploaded_valid = false;
  // concurrency and graphs
  hipStream_t side[kSide] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kSide] = {};
  hipStream_t capture = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<BatchFrameDev> graph_table;
  uint32_t graph_version = 0;
  // kernel choices
  oit_batch_tuning tuning = default_tuning();
  uint32_t tuning_version = 1;
};

namespace {

void release(oit_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  (void)hipFree(b->d_tab);
  (void)hipFree(b->head);
  (void)hipFree(b->cd);
  (void)hipFree(b->next);
  (void)hipFree(b->counters);
  (void)hipFree(b->cursor);
  (void)hipFree(b->deep_q);
  (void)hipFree(b->tile_sum);
  (void)hipFree(b->deep_n);
  (void)hipFree(b->d_table);
  if (b->h_table) (void)hipHostFree(b->h_table);
  (void)hipFree(b->d_sched);
  if (b->h_sched) (void)hipHostFree(b->h_sched);
  if (b->table_ev) (void)hipEventDestroy(b->table_ev);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  for (int i = 0; i < kSide; ++i) {
    if (b->join_ev[i]) (void)hipEventDestroy(b->join_ev[i]);
    if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
  }
  if (b->capture) (void)hipStreamDestroy(b->capture);
  delete b;
}

LLBuf pool_view(const oit_batch* b) {
  LLBuf ll;
  ll.head = b->head;
  ll.cd = b->cd;
  ll.next = b->next;
  ll.cap = b->cap_nodes;
  ll.overflow = b->counters;
  ll.counters = b->counters;
  return ll;
}

// Grows the pool (never shrinks) and brings it to the clean state.
int ensure_pool(oit_batch* b, uint64_t nodes, uint64_t heads, hipStream_t s) {
  int rc;
  if (nodes > b->cap_nodes || heads > b->cap_heads) {
    if ((rc = st(hipDeviceSynchronize()))) return rc;
    if (b->gexec) {  // captured launches hold the old pointers
      (void)hipGraphExecDestroy(b->gexec);
      b->gexec = nullptr;
    }
  }
  if (nodes > b->cap_nodes) {
    (void)hipFree(b->cd);
    (void)hipFree(b->next);
    b->cd = nullptr;
    b->next = nullptr;
    b->cap_nodes = 0;
    const uint64_t want = round_up(nodes, 1ull << 20);
    const uint64_t cap = want < kMaxPoolSlots + 1 ? want : nodes;
    if ((rc = st(hipMalloc((void**)&b->cd, cap * sizeof(uint64_t))))) return rc;
    if ((rc = st(hipMalloc((void**)&b->next, cap * sizeof(uint32_t))))) return rc;
    b->cap_nodes = cap;
    if ((rc = st(oit::launch_fill_u32(b->next, cap, oit::kUnset, s)))) return rc;
  }
  if (heads > b->cap_heads) {
    (void)hipFree(b->head);
    (void)hipFree(b->cursor);
    (void)hipFree(b->deep_q);
    (void)hipFree(b->tile_sum);
    b->head = b->cursor = b->deep_q = b->tile_sum = nullptr;
    b->cap_heads = 0;
    const uint64_t cap = round_up(heads, 1ull << 16);
    if ((rc = st(hipMalloc((void**)&b->head, cap * sizeof(uint32_t))))) return rc;
    if ((rc = st(hipMalloc((void**)&b->cursor, cap * sizeof(uint32_t))))) return rc;
    if ((rc = st(hipMalloc((void**)&b->deep_q, cap * sizeof(uint32_t))))) return rc;
    if ((rc = st(hipMalloc((void**)&b->tile_sum, cap / kTile * sizeof(uint32_t))))) return rc;
    b->cap_heads = cap;
    if ((rc = st(oit::launch_fill_u32(b->head, cap, 0u, s)))) return rc;
  }
  if (b->dirty) {
    if (b->cap_nodes && (rc = st(oit::launch_fill_u32(b->next, b->cap_nodes, oit::kUnset, s)))) return rc;
    if (b->cap_heads && (rc = st(oit::launch_fill_u32(b->head, b->cap_heads, 0u, s)))) return rc;
    b->dirty = false;
  }
  return OIT_OK;
}

int ensure_table(oit_batch* b, uint32_t n) {
  if (n <= b->table_cap) return OIT_OK;
  int rc;
  if ((rc = st(hipDeviceSynchronize()))) return rc;
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  (void)hipFree(b->d_table);
  if (b->h_table) (void)hipHostFree(b->h_table);
  b->d_table = nullptr;
  b->h_table = nullptr;
  b->table_cap = 0;
  b->uploaded_valid = false;
  b->table_ev_pending = false;
  const uint32_t cap = std::max<uint32_t>(n, 64);
  if ((rc = st(hipMalloc((void**)&b->d_table, cap * sizeof(BatchFrameDev))))) return rc;
  if ((rc = st(hipHostMalloc((void**)&b->h_table, cap * sizeof(BatchFrameDev), hipHostMallocDefault)))) return rc;
  b->table_cap = cap;
  return OIT_OK;
}

// Stream-ordered upload of the frame table, skipped when the device copy
// already holds it.
int upload_table(oit_batch* b, const Plan& p, hipStream_t s) {
  int rc;
  if ((rc = ensure_table(b, (uint32_t)p.table.size()))) return rc;
  if (b->uploaded_valid && same_table(b->uploaded, p.table)) return OIT_OK;
  if (p.table.empty()) return OIT_OK;
  if (b->table_ev_pending) {  // the previous copy has read the staging buffer
    if ((rc = st(hipEventSynchronize(b->table_ev)))) return rc;
    b->table_ev_pending = false;
  }
  if (p.sched.size() > b->sched_cap) {
    if ((rc = st(hipDeviceSynchronize()))) return rc;
    if (b->gexec) {  // captured launches hold the old pointer
      (void)hipGraphExecDestroy(b->gexec);
      b->gexec = nullptr;
    }
    (void)hipFree(b->d_sched);
    if (b->h_sched) (void)hipHostFree(b->h_sched);
    b->d_sched = nullptr;
    b->h_sched = nullptr;
    b->sched_cap = 0;
    const size_t cap = std::max<size_t>(p.sched.size(), 4096);
    if ((rc = st(hipMalloc((void**)&b->d_sched, cap * sizeof(uint32_t))))) return rc;
    if ((rc = st(hipHostMalloc((void**)&b->h_sched, cap * sizeof(uint32_t), hipHostMallocDefault)))) return rc;
    b->sched_cap = cap;
  }
  std::memcpy(b->h_table, p.table.data(), p.table.size() * sizeof(BatchFrameDev));
  std::memcpy(b->h_sched, p.sched.data(), p.sched.size() * sizeof(uint32_t));
  if ((rc = st(hipMemcpyAsync(b->d_table, b->h_table, p.table.size() * sizeof(BatchFrameDev),
                              hipMemcpyHostToDevice, s))))
    return rc;
  if ((rc = st(hipMemcpyAsync(b->d_sched, b->h_sched, p.sched.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              s))))
    return rc;
  if ((rc = st(hipEventRecord(b->table_ev, s)))) return rc;
  b->table_ev_pending = true;
  b->uploaded = p.table;
  b->uploaded_valid = true;
  return OIT_OK;
}

BatchArgs args_for(const oit_batch* b, const Plan& p) {
  BatchArgs a;
  a.frames = b->d_table;
  a.f0 = 0;
  a.f1 = (uint32_t)p.table.size();
  a.units = 0;
  a.tile0 = 0;
  a.tiles = 0;
  a.tab = b->d_tab;
  a.cursor = b->cursor;
  a.tile_sum = b->tile_sum;
  a.deep_q = b->deep_q;
  a.deep_n = b->deep_n;
  a.cursor_cap = (uint32_t)b->cap_heads;
  return a;
}

int launch_append(const oit_batch* b, const Plan& p, uint32_t chunks, hipStream_t s) {
  BatchArgs a = args_for(b, p);
  a.units = (uint32_t)(p.pool_slots / (64u * chunks));
  if (!(b->flags & OIT_BATCH_FRAME_ORDER) && !p.table.empty()) {
    a.sched = b->d_sched;
    a.rounds = p.rounds;
  }
  return st(oit::launch_batch_append(a, pool_view(b), chunks, s));
}

// Binned frames' scan and scatter, after the append launch counted them.
int launch_bin(const oit_batch* b, const Plan& p, uint32_t chunks, hipStream_t s) {
  if (!p.bin) return OIT_OK;
  BatchArgs a = args_for(b, p);
  a.units = (uint32_t)(p.pool_slots / (64u * chunks));
  return st(oit::launch_bin_prepare(a, pool_view(b), (uint32_t)(p.heads / kTile), chunks, s));
}

int launch_group(const oit_batch* b, const Plan& p, const Group& g, uint32_t gi, uint32_t variant, uint32_t blocks,
                 hipStream_t s) {
  BatchArgs a = args_for(b, p);
  a.f0 = g.f0;
  a.f1 = g.f1;
  a.tile0 = g.tile0;
  a.tiles = g.tiles;
  (void)gi;
  int rc = st(oit::launch_batch_resolve(a, pool_view(b), g.K, g.tail != 0, !(variant & OIT_VARIANT_LL_SINGLE_WALK),
                                        (variant & OIT_VARIANT_LL_KEYS_ONLY) != 0, blocks, s));
  if (!rc && g.bin) rc = st(oit::launch_bin_resolve(a, pool_view(b), g.K, g.tail != 0, blocks, s));
  return rc;
}

// Append + every group's resolve on s (resolve groups forked when asked).
int enqueue(oit_batch* b, const Plan& p, hipStream_t s) {
  int rc;
  if ((rc = launch_append(b, p, b->tuning.append_chunks, s))) return rc;
  if ((rc = launch_bin(b, p, b->tuning.append_chunks, s))) return rc;
  const bool fork = (b->flags & OIT_BATCH_CONCURRENT) && p.groups.size() > 1;
  if (fork && (rc = st(hipEventRecord(b->fork_ev, s)))) return rc;
  bool used[kSide] = {};
  for (size_t g = 0; g < p.groups.size(); ++g) {
    const Group& gr = p.groups[g];
    const int ki = k_index(gr.K);
    hipStream_t gs = s;
    if (fork && g > 0) {
      const int j = (int)((g - 1) % kSide);
      gs = b->side[j];
      if (!used[j]) {
        if ((rc = st(hipStreamWaitEvent(gs, b->fork_ev, 0)))) return rc;
        used[j] = true;
      }
    }
    if ((rc = launch_group(b, p, gr, (uint32_t)g, b->tuning.resolve_variant[ki][gr.tail],
                           b->tuning.resolve_blocks[ki][gr.tail], gs)))
      return rc;
  }
  for (int j = 0; j < kSide; ++j) {
    if (!used[j]) continue;
    if ((rc = st(hipEventRecord(b->join_ev[j], b->side[j])))) return rc;
    if ((rc = st(hipStreamWaitEvent(s, b->join_ev[j], 0)))) return rc;
  }
  return OIT_OK;
}

int run_graph(oit_batch* b, const Plan& p, hipStream_t s) {
  int rc;
  if (b->gexec && b->graph_version == b->tuning_version && same_table(b->graph_table, p.table))
    return st(hipGraphLaunch(b->gexec, s));
  // the table a captured graph reads is uploaded before capture and stays
  if ((rc = st(hipStreamSynchronize(s)))) return rc;
  if ((rc = upload_table(b, p, s))) return rc;
  if ((rc = st(hipStreamSynchronize(s)))) return rc;
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  if ((rc = st(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal)))) return rc;
  const int erc = enqueue(b, p, b->capture);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->capture, &graph);
  if (erc) {
    if (graph) (void)hipGraphDestroy(graph);
    return erc;
  }
  if ((rc = st(ec))) return rc;
  rc = st(hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  if (rc) {
    b->gexec = nullptr;
    return rc;
  }
  b->graph_table = p.table;
  b->graph_version = b->tuning_version;
  return st(hipGraphLaunch(b->gexec, s));
}

int prepare(oit_batch* b, const oit_batch_frame* frames, uint32_t n, Plan& p, hipStream_t s) {
  int rc;
  if ((rc = make_plan(frames, n, p, b->tuning.bin_density))) return rc;
  return ensure_pool(b, p.pool_slots + 1, p.heads, s);
}

}  // namespace

extern "C" {

int oit_batch_create(oit_batch** out, int device, uint32_t flags) {
  if (!out) return OIT_ERR_INVALID_ARG;
  *out = nullptr;
  if (flags & ~(uint32_t)(OIT_BATCH_CONCURRENT | OIT_BATCH_GRAPH | OIT_BATCH_FRAME_ORDER)) return OIT_ERR_INVALID_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return OIT_ERR_DEVICE;
  }
  oit_batch* b = new (std::nothrow) oit_batch();
  if (!b) return OIT_ERR_OOM;
  b->device = device;
  b->flags = flags;
  DeviceGuard g(device);
  oit::CodecTables host;
  oit::build_codec_tables(&host);
  int rc = st(hipMalloc((void**)&b->d_tab, sizeof host));
  if (!rc) rc = st(hipMemcpy(b->d_tab, &host, sizeof host, hipMemcpyHostToDevice));
  if (!rc) rc = st(hipMalloc((void**)&b->counters, 32 * sizeof(uint32_t)));
  if (!rc) rc = st(hipMemset(b->counters, 0, 32 * sizeof(uint32_t)));
  if (!rc) rc = st(hipMalloc((void**)&b->deep_n, oit::kBatchMaxGroups * sizeof(uint32_t)));
  if (!rc) rc = st(hipMemset(b->deep_n, 0, oit::kBatchMaxGroups * sizeof(uint32_t)));
  if (!rc) rc = st(hipEventCreateWithFlags(&b->table_ev, hipEventDisableTiming));
  if (!rc) rc = st(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
  for (int i = 0; !rc && i < kSide; ++i) {
    rc = st(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
    if (!rc) rc = st(hipEventCreateWithFlags(&b->join_ev[i], hipEventDisableTiming));
  }
  if (!rc) rc = st(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
  if (rc) {
    release(b);
    return rc;
  }
  *out = b;
  return OIT_OK;
}

int oit_batch_destroy(oit_batch* b) {
  if (!b) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  (void)hipDeviceSynchronize();
  release(b);
  return OIT_OK;
}

int oit_batch_query(const oit_batch* b, const oit_batch_frame* frames, uint32_t n, oit_batch_sizes* out) {
  if (!b || !out) return OIT_ERR_INVALID_ARG;
  Plan p;
  const int rc = make_plan(frames, n, p, b->tuning.bin_density);
  if (rc) return rc;
  out->pool_nodes = p.pool_slots + 1;
  out->head_entries = p.heads;
  // nodes 12 B; per head: the head, a bucket cursor and a deep-queue entry; a scan base per tile
  out->bytes = out->pool_nodes * 12u + out->head_entries * 12u + out->head_entries / kTile * 4u;
  out->groups = (uint32_t)p.groups.size();
  out->frames = n;
  return OIT_OK;
}

int oit_batch_reserve(oit_batch* b, uint64_t pool_nodes, uint64_t head_entries) {
  if (!b) return OIT_ERR_INVALID_ARG;
  if (pool_nodes > kMaxPoolSlots + 1 || head_entries > kMaxHeads) return OIT_ERR_CAPACITY;
  DeviceGuard g(b->device);
  const int rc = ensure_pool(b, pool_nodes, head_entries, nullptr);
  if (rc) return rc;
  return st(hipStreamSynchronize(nullptr));
}

int oit_batch_run(oit_batch* b, const oit_batch_frame* frames, uint32_t n, void* stream) {
  if (!b) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  Plan p;
  int rc = prepare(b, frames, n, p, s);
  if (rc) return rc;
  if (p.table.empty()) return OIT_OK;
  if (b->flags & OIT_BATCH_GRAPH) {
    rc = run_graph(b, p, s);
  } else {
    rc = upload_table(b, p, s);
    if (!rc) rc = enqueue(b, p, s);
  }
  if (rc) b->dirty = true;
  return rc;
}

int oit_batch_tune(oit_batch* b, const oit_batch_frame* frames, uint32_t n, void* stream) {
  if (!b) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  Plan p;
  int rc = prepare(b, frames, n, p, s);
  if (rc) return rc;
  if (p.table.empty()) return OIT_OK;
  if ((rc = upload_table(b, p, s))) return rc;
  const size_t G = p.groups.size();
  std::vector<hipEvent_t> ev(2 * G + 2, nullptr);
  for (auto& e : ev)
    if ((rc = st(hipEventCreate(&e)))) break;
  oit_batch_tuning best = b->tuning;
  auto resolve_all = [&](const uint32_t* variant, const uint32_t* blocks, bool timed) -> int {
    for (size_t gi = 0; gi < G; ++gi) {
      int r;
      if (timed && (r = st(hipEventRecord(ev[2 + 2 * gi], s)))) return r;
      if ((r = launch_group(b, p, p.groups[gi], (uint32_t)gi, variant[gi], blocks[gi], s))) return r;
      if (timed && (r = st(hipEventRecord(ev[3 + 2 * gi], s)))) return r;
    }
    return OIT_OK;
  };
  std::vector<uint32_t> cur_v(G), cur_b(G);
  for (size_t gi = 0; gi < G; ++gi) {
    const int ki = k_index(p.groups[gi].K);
    cur_v[gi] = b->tuning.resolve_variant[ki][p.groups[gi].tail];
    cur_b[gi] = b->tuning.resolve_blocks[ki][p.groups[gi].tail];
  }
  // frames binned from which density (fragments per target): the whole run
  // timed for each candidate, from never to always
  if (!rc) {
    const float dens[7] = {3.0e38f, 32.f, 16.f, 8.f, 4.f, 2.f, 0.f};
    float best_ms = 0.f;
    for (int di = 0; di < 7 && !rc; ++di) {
      Plan pd;
      if ((rc = make_plan(frames, n, pd, dens[di]))) break;
      if ((rc = upload_table(b, pd, s))) break;
      float ms_min = 0.f;
      for (int rep = 0; rep < 2 && !rc; ++rep) {
        if ((rc = st(hipEventRecord(ev[0], s)))) break;
        if ((rc = launch_append(b, pd, b->tuning.append_chunks, s))) break;
        if ((rc = launch_bin(b, pd, b->tuning.append_chunks, s))) break;
        for (size_t gi = 0; gi < G && !rc; ++gi)
          rc = launch_group(b, pd, pd.groups[gi], (uint32_t)gi, cur_v[gi], cur_b[gi], s);
        if (rc) break;
        if ((rc = st(hipEventRecord(ev[1], s)))) break;
        if ((rc = st(hipEventSynchronize(ev[1])))) break;
        float ms = 0.f;
        if ((rc = st(hipEventElapsedTime(&ms, ev[0], ev[1])))) break;
        if (rep == 0 || ms < ms_min) ms_min = ms;
      }
      if (rc) break;
      if (di == 0 || ms_min < best_ms) {
        best_ms = ms_min;
        best.bin_density = dens[di];
      }
    }
    if (!rc) rc = make_plan(frames, n, p, best.bin_density);
    if (!rc) rc = upload_table(b, p, s);
  }
  // append: chunks per wave
  if (!rc) {
    float best_ms = 0.f;
    for (uint32_t chunks : {4u, 8u}) {
      float ms_min = 0.f;
      for (int rep = 0; rep < 3 && !rc; ++rep) {
        if ((rc = st(hipEventRecord(ev[0], s)))) break;
        if ((rc = launch_append(b, p, chunks, s))) break;
        if ((rc = launch_bin(b, p, chunks, s))) break;
        if ((rc = st(hipEventRecord(ev[1], s)))) break;
        if ((rc = resolve_all(cur_v.data(), cur_b.data(), false))) break;
        if ((rc = st(hipEventSynchronize(ev[1])))) break;
        float ms = 0.f;
        if ((rc = st(hipEventElapsedTime(&ms, ev[0], ev[1])))) break;
        if (rep == 0 || ms < ms_min) ms_min = ms;
      }
      if (rc) break;
      if (chunks == 4u || ms_min < best_ms) {
        best_ms = ms_min;
        best.append_chunks = chunks;
      }
    }
  }
  // resolve: variant and block cap per group
  const uint32_t variants[4] = {0u, OIT_VARIANT_LL_SINGLE_WALK, OIT_VARIANT_LL_KEYS_ONLY,
                                OIT_VARIANT_LL_SINGLE_WALK | OIT_VARIANT_LL_KEYS_ONLY};
  const uint32_t caps[3] = {kDefaultBlocks, 2048u, 0u};
  std::vector<float> gbest(G, 0.f);
  std::vector<bool> have(G, false);
  for (int vi = 0; vi < 4 && !rc; ++vi) {
    for (int ci = 0; ci < 3 && !rc; ++ci) {
      std::vector<uint32_t> v(G, variants[vi]), c(G, caps[ci]);
      std::vector<float> tmin(G, 0.f);
      for (int rep = 0; rep < 2 && !rc; ++rep) {
        if ((rc = launch_append(b, p, best.append_chunks, s))) break;
        if ((rc = launch_bin(b, p, best.append_chunks, s))) break;
        if ((rc = resolve_all(v.data(), c.data(), true))) break;
        if ((rc = st(hipStreamSynchronize(s)))) break;
        for (size_t gi = 0; gi < G && !rc; ++gi) {
          float ms = 0.f;
          rc = st(hipEventElapsedTime(&ms, ev[2 + 2 * gi], ev[3 + 2 * gi]));
          if (rep == 0 || ms < tmin[gi]) tmin[gi] = ms;
        }
      }
      for (size_t gi = 0; gi < G && !rc; ++gi) {
        if (!have[gi] || tmin[gi] < gbest[gi]) {
          have[gi] = true;
          gbest[gi] = tmin[gi];
          const int ki = k_index(p.groups[gi].K);
          best.resolve_variant[ki][p.groups[gi].tail] = variants[vi];
          best.resolve_blocks[ki][p.groups[gi].tail] = caps[ci];
        }
      }
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc) {
    b->dirty = true;
    return rc;
  }
  b->tuning = best;
  ++b->tuning_version;
  return OIT_OK;
}

int oit_batch_get_tuning(const oit_batch* b, oit_batch_tuning* out) {
  if (!b || !out) return OIT_ERR_INVALID_ARG;
  *out = b->tuning;
  return OIT_OK;
}

int oit_batch_set_tuning(oit_batch* b, const oit_batch_tuning* t) {
  if (!b || !t || !tuning_valid(*t)) return OIT_ERR_INVALID_ARG;
  b->tuning = *t;
  ++b->tuning_version;
  return OIT_OK;
}

}  // extern "C"
