// oit_api.hip - host runtime behind include/oit.h: configuration, A-buffer
// sizing and ownership, per-frame pass order, and kernel dispatch.
//
// Reference counterparts: State and its rebuild logic (oit.h:64-116,
// main.cpp:130-250), createFrameImages sizing (oit.cpp:84-163), the
// clear/draw method pairs dispatched by Sample::onRender (oitRender.cpp:
// 28-154, 156-425).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "oit.h"
#include "oit_kernels.hpp"

using oit::kEmpty;

namespace {

// --- sRGB8 codec tables (host) ---------------------------------------------
// Encode curve of unPremultLinearToSRGB (shaderCommon.glsl:60-69) and the
// decode curve of unPremultSRGBToLinear (shaderCommon.glsl:84-94), evaluated
// in double; thresholds are found by bisection over binary32 bit patterns.
double srgb_curve(double c) { return c < (double)0.0031308f ? c * 12.92 : std::pow(c, 1.0 / 2.4) * 1.055 - 0.055; }

int srgb_code(float c) {
  double s = srgb_curve((double)c);
  if (!(s > 0.0)) s = 0.0;
  if (s > 1.0) s = 1.0;
  return (int)std::floor(s * 255.0 + 0.5);
}

float from_bits(uint32_t b) {
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

}  // namespace

namespace oit {

void build_codec_tables(CodecTables* t) {
  t->thr[0] = -INFINITY;
  t->thr[256] = INFINITY;
  for (int q = 1; q <= 255; ++q) {
    uint32_t lo = 0u, hi = 0x3F800000u;
    while (hi - lo > 1u) {
      const uint32_t mid = lo + (hi - lo) / 2u;
      if (srgb_code(from_bits(mid)) >= q) hi = mid; else lo = mid;
    }
    t->thr[q] = from_bits(hi);
  }
  for (int u = 0; u < 256; ++u) {
    const double f = (double)u / 255.0;
    const double lin = f < (double)0.04045f ? f / 12.92 : std::pow((f + 0.055) / 1.055, 2.4);
    t->dec[u] = (float)lin;
    t->alpha[u] = (float)u / 255.0f;
  }
}

}  // namespace oit

namespace {

constexpr uint32_t kMaxFrameFragments = 0x7FFFFFFEu;  // ids and tail links fit in 31 bits
constexpr uint32_t kCounterWords = 32;  // frame counters [0..15], last frame's [16..31]; then the error word
constexpr uint64_t kMaxPoolNodes = 0xFFFFFFFEull;      // node indices stay below the kUnset sentinel

enum Stage { kStClear = 0, kStStore, kStLink, kStClassify, kStTailLink, kStResolve, kStCount };

struct EventPair {
  hipEvent_t a, b;
  int stage;
};

bool valid_layers(uint32_t k) { return k == 1 || k == 2 || k == 4 || k == 8 || k == 16 || k == 32; }

}  // namespace

struct oit_ctx {
  int device = 0;
  bool configured = false;
  oit_config cfg{};
  uint32_t T = 0;
  oit::CodecTables* d_tab = nullptr;
  uint32_t* d_counters = nullptr;  // [1] LL overflow; [16..31] last frame (LL)
  // Linked List
  uint32_t* ll_head = nullptr;
  uint64_t* ll_cd = nullptr;
  uint32_t* ll_next = nullptr;
  uint64_t ll_cap = 0;
  // Linked List pool state: "clean" = heads 0 and every link word kUnset, as
  // the resetting resolve leaves them; dirty_upto bounds the nodes a frame
  // that was not resolved may have touched.
  bool ll_clean = false;
  uint64_t ll_dirty_upto = 0;
  bool counters_clean = false;
  bool stats_moved = false;  // the last resolve moved the counters to [16..31]
  // Simple / Spinlock
  uint32_t* kb_count = nullptr;
  uint32_t* kb_lock = nullptr;
  uint32_t* kb_depth = nullptr;
  uint32_t* kb_filled = nullptr;  // Spinlock: completed stores below K per target
  uint32_t* kb_col = nullptr;
  uint32_t* kb_key = nullptr;
  uint32_t* kb_id = nullptr;
  // Loop64
  uint64_t* l64 = nullptr;
  uint32_t* l64_cnt = nullptr;  // [T] the frame's fragments per target (0 = clean)
  uint64_t* l64_big = nullptr;  // per 64-fragment chunk of the frame: lanes whose pixel holds more than K
  bool l64_clean = false;  // every slot empty and every count 0 (a fill, or the last frame's composite)
  bool kb_clean = false;   // Simple / Spinlock: counts, lock words and bounds at their clear values, likewise
  bool l32_clean = false;  // Loop32: likewise for the depth and colour slots
  bool wb_clean = false;   // WBOIT: accumulators at the clear values (a fill, or the last composite)
  // WBOIT
  float4* wb_accum = nullptr;
  float* wb_reveal = nullptr;
  // colour-pass tail lists
  uint32_t* tl_head = nullptr;
  uint32_t* tl_next = nullptr;
  float4* tl_rgba = nullptr;
  uint64_t* tl_rec = nullptr;
  uint32_t* tl_qn = nullptr;  // tail fragments (queue entries) per region [oit::tail_regions(tl_cap)]
  uint32_t tl_cap = 0;
  bool tl_head_valid = false;  // cleared for the current frame
  bool tl_clean = false;       // tail heads and region counters all 0: filled, and no tail pushed since
  // frame state: oit_clear opens a frame; the first resolve links the tail
  // lists, after which the frame takes no more appends.
  bool frame_open = false;
  bool resolved = false;
  uint64_t frame_frags = 0;
  uint64_t appends = 0;
  std::vector<oit::FragView> segments;
  // sizes
  uint64_t bytes_abuf = 0, bytes_aux = 0;
  // allocated elements of each frame buffer: a reconfiguration reuses a
  // buffer while the new configuration fits, so a context that serves many
  // frame sizes stops allocating once it has seen the largest
  uint64_t cap_ll_head = 0, ll_alloc = 0;
  uint64_t cap_kb_count = 0, cap_kb_lock = 0, cap_kb_depth = 0, cap_kb_filled = 0, cap_kb_elems = 0;
  uint64_t cap_l64 = 0, cap_l64_cnt = 0, cap_l64_big = 0, cap_wb = 0, cap_tl_head = 0;
  // coverage shading: sample masks of nodes, k-buffer slots and tail entries
  uint32_t* ll_mask = nullptr;
  float4* il_full = nullptr;   // ordered interlock: shaded colour per node [ll_alloc]
  uint32_t* l32_z = nullptr;   // Loop32 depth slots [K][T]
  uint64_t* l32_c = nullptr;   // Loop32 colour slots [K][T]: (id + 1) << 32 | colour
  uint64_t cap_il_full = 0, cap_l32 = 0;
  uint32_t* kb_mask = nullptr;
  uint32_t* tl_mask = nullptr;
  uint64_t cap_ll_mask = 0, cap_kb_mask = 0, cap_tl_mask = 0;
  // opaque pass: (depth key << 32 | id) and colour per sample target
  // [T * max(coverage samples, 1)], valid for the frame once opq_active
  uint64_t* opq_z = nullptr;
  float4* opq_rgba = nullptr;
  uint64_t cap_opq = 0, cap_opq_rgba = 0;
  bool opq_active = false;
  uint64_t opq_frags = 0;  // opaque fragments drawn this frame (their ids)
  // the last stream-ordered call's completion: buffers are freed only after it
  hipEvent_t work_ev = nullptr;
  bool work_pending = false;
  // profiling
  bool profiling = false;
  std::vector<EventPair> events;
  std::vector<hipEvent_t> event_pool;
  double prof_ms[kStCount] = {0};
  uint64_t prof_launches[kStCount] = {0};
};

namespace {

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

template <typename T>
int dev_alloc(T** p, uint64_t count) {
  *p = nullptr;
  if (count == 0) return OIT_OK;
  void* q = nullptr;
  if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return OIT_ERR_OOM;
  }
  *p = static_cast<T*>(q);
  return OIT_OK;
}

template <typename T>
void dev_free(T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

// Waits for the context's last stream-ordered work (before buffers it may
// use are freed).
void wait_idle(oit_ctx* c) {
  if (c->work_pending) {
    (void)hipEventSynchronize(c->work_ev);
    c->work_pending = false;
  }
}

void mark_work(oit_ctx* c, hipStream_t s) {
  if (c->work_ev && hipEventRecord(c->work_ev, s) == hipSuccess) c->work_pending = true;
}

// Makes *p hold at least `need` elements, reusing the allocation when it
// fits; *grown reports a new allocation (its contents are undefined).
template <typename T>
int reserve(oit_ctx* c, T** p, uint64_t* cap, uint64_t need, bool* grown = nullptr) {
  if (need == 0 || (*p && need <= *cap)) return OIT_OK;
  wait_idle(c);
  dev_free(p);
  *cap = 0;
  const int rc = dev_alloc(p, need);
  if (rc) return rc;
  *cap = need;
  if (grown) *grown = true;
  return OIT_OK;
}

template <typename T>
void release(oit_ctx* c, T** p, uint64_t* cap) {
  if (*p) {
    wait_idle(c);
    dev_free(p);
  }
  *cap = 0;
}

void free_frame_buffers(oit_ctx* c) {
  wait_idle(c);
  c->cap_ll_head = c->ll_alloc = 0;
  c->cap_kb_count = c->cap_kb_lock = c->cap_kb_depth = c->cap_kb_filled = c->cap_kb_elems = 0;
  c->cap_l64 = c->cap_l64_cnt = c->cap_l64_big = c->cap_wb = c->cap_tl_head = 0;
  c->cap_ll_mask = c->cap_kb_mask = c->cap_tl_mask = 0;
  dev_free(&c->ll_mask);
  dev_free(&c->il_full);
  dev_free(&c->l32_z);
  dev_free(&c->l32_c);
  c->l32_clean = false;
  c->cap_il_full = c->cap_l32 = 0;
  dev_free(&c->kb_mask);
  dev_free(&c->tl_mask);
  dev_free(&c->opq_z);
  dev_free(&c->opq_rgba);
  c->cap_opq = c->cap_opq_rgba = 0;
  c->opq_active = false;
  dev_free(&c->ll_head);
  dev_free(&c->ll_cd);
  dev_free(&c->ll_next);
  dev_free(&c->kb_count);
  dev_free(&c->kb_lock);
  dev_free(&c->kb_depth);
  dev_free(&c->kb_filled);
  dev_free(&c->kb_col);
  dev_free(&c->kb_key);
  dev_free(&c->kb_id);
  dev_free(&c->l64);
  dev_free(&c->l64_cnt);
  dev_free(&c->l64_big);
  c->l64_clean = false;
  c->kb_clean = false;
  dev_free(&c->wb_accum);
  dev_free(&c->wb_reveal);
  c->wb_clean = false;
  dev_free(&c->tl_head);
  dev_free(&c->tl_next);
  dev_free(&c->tl_rgba);
  dev_free(&c->tl_rec);
  dev_free(&c->tl_qn);
  c->tl_cap = 0;
  c->tl_head_valid = false;
  c->tl_clean = false;
  c->ll_cap = 0;
  c->bytes_abuf = c->bytes_aux = 0;
  c->configured = false;
}

int status_of(hipError_t e) {
  if (e == hipSuccess) return OIT_OK;
  if (e == hipErrorOutOfMemory) return OIT_ERR_OOM;
  if (e == hipErrorInvalidValue) return OIT_ERR_INVALID_ARG;
  return OIT_ERR_DEVICE;
}

hipEvent_t take_event(oit_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Runs `launch` bracketed by events when profiling is on.
template <typename F>
hipError_t staged(oit_ctx* c, int stage, hipStream_t s, F&& launch) {
  if (!c->profiling) return launch();
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return launch();
  (void)hipEventRecord(a, s);
  const hipError_t e = launch();
  (void)hipEventRecord(b, s);
  c->events.push_back({a, b, stage});
  return e;
}

oit::Common common_of(const oit_ctx* c) {
  oit::Common k;
  k.T = c->T;
  k.K = c->cfg.layers;
  k.tail = c->cfg.tail_blend ? 1 : 0;
  k.background = make_float4(c->cfg.background[0], c->cfg.background[1], c->cfg.background[2], c->cfg.background[3]);
  k.tab = c->d_tab;
  k.cov = c->cfg.coverage_samples > 1 ? c->cfg.coverage_samples : 0;
  k.spp = c->cfg.samples > 1 ? c->cfg.samples : 1;
  if (c->opq_active) {
    k.opq_z = c->opq_z;
    k.opq_rgba = c->opq_rgba;
  }
  return k;
}

bool coverage(const oit_ctx* c) { return c->cfg.coverage_samples > 1; }

// The context's device error word (oit::kErr* bits, sticky until
// oit_get_stats reads it), after the 32 frame counters.
uint32_t* dev_err(const oit_ctx* c) { return c->d_counters + kCounterWords; }

oit::TailBuf tail_of(const oit_ctx* c) {
  oit::TailBuf t;
  const bool on = c->tl_head != nullptr && c->tl_head_valid;
  t.head = on ? c->tl_head : nullptr;
  t.next = c->tl_next;
  t.rgba = c->tl_rgba;
  t.rec = c->tl_rec;
  t.qn = c->tl_qn;
  t.cap = c->tl_cap;
  t.mask = c->cfg.coverage_samples > 1 ? c->tl_mask : nullptr;
  t.err = dev_err(c);
  return t;
}

// The Linked List and the ordered interlock keep fragments in the node pool.
bool pooled(uint32_t alg) { return alg == OIT_LINKEDLIST || alg == OIT_INTERLOCK; }

// The ordered interlock blends its tail in the resolve, in primitive order.
bool uses_tail_lists(const oit_ctx* c) {
  return c->cfg.tail_blend && c->cfg.algorithm != OIT_WEIGHTED && c->cfg.algorithm != OIT_INTERLOCK;
}

// Makes the per-fragment tail arrays hold `need` fragments, keeping this
// frame's records (stream-ordered copies; the old buffers are freed after a
// stream synchronisation).
int ensure_tail_capacity(oit_ctx* c, uint64_t need, hipStream_t s) {
  if (!c->tl_head || c->cap_tl_head < c->T) {
    int rc = reserve(c, &c->tl_head, &c->cap_tl_head, c->T);
    if (rc) return rc;
    c->tl_head_valid = false;
  }
  if (!c->tl_head_valid) {
    hipError_t e = oit::launch_fill_u32(c->tl_head, c->T, 0u, s);  // empty tail lists
    if (e == hipSuccess && c->tl_qn) e = oit::launch_fill_u32(c->tl_qn, oit::tail_regions(c->tl_cap), 0u, s);
    if (e != hipSuccess) return status_of(e);
    c->tl_head_valid = true;
  }
  if (need <= c->tl_cap) return OIT_OK;
  uint64_t cap = c->tl_cap ? c->tl_cap : 1024;
  while (cap < need) cap *= 2;
  if (c->cfg.max_fragments && cap > c->cfg.max_fragments) cap = c->cfg.max_fragments;
  if (cap < need) return OIT_ERR_CAPACITY;
  if (cap > kMaxFrameFragments) cap = kMaxFrameFragments;
  uint32_t* nn = nullptr;
  float4* nr = nullptr;
  uint64_t* nrec = nullptr;
  uint32_t* nm = nullptr;
  uint32_t* nq = nullptr;
  const uint64_t regions = oit::tail_regions(cap);
  int rc = dev_alloc(&nn, cap);
  if (!rc) rc = dev_alloc(&nr, cap);
  if (!rc) rc = dev_alloc(&nrec, cap);
  if (!rc) rc = dev_alloc(&nq, regions);
  if (!rc && coverage(c)) rc = dev_alloc(&nm, cap);
  if (rc) {
    dev_free(&nn);
    dev_free(&nr);
    dev_free(&nrec);
    dev_free(&nq);
    dev_free(&nm);
    return rc;
  }
  // region counters: this frame's so far, zero beyond
  hipError_t fe = oit::launch_fill_u32(nq, regions, 0u, s);
  if (fe == hipSuccess && c->tl_cap && c->tl_qn)
    fe = hipMemcpyAsync(nq, c->tl_qn, oit::tail_regions(c->tl_cap) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
  if (c->tl_cap) {
    (void)hipMemcpyAsync(nn, c->tl_next, (uint64_t)c->tl_cap * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(nr, c->tl_rgba, (uint64_t)c->tl_cap * sizeof(float4), hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(nrec, c->tl_rec, (uint64_t)c->tl_cap * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
    if (nm && c->tl_mask)
      (void)hipMemcpyAsync(nm, c->tl_mask, (uint64_t)c->tl_cap * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    (void)hipStreamSynchronize(s);
  }
  dev_free(&c->tl_next);
  dev_free(&c->tl_rgba);
  dev_free(&c->tl_rec);
  dev_free(&c->tl_qn);
  dev_free(&c->tl_mask);
  c->tl_next = nn;
  c->tl_rgba = nr;
  c->tl_rec = nrec;
  c->tl_qn = nq;
  c->tl_mask = nm;
  c->cap_tl_mask = nm ? cap : 0;
  c->tl_cap = (uint32_t)cap;
  return fe == hipSuccess ? OIT_OK : status_of(fe);
}

// Grows the ordered interlock's node pool (it holds every fragment of the
// frame) to at least `need` nodes, keeping this frame's nodes: stream-ordered
// copies, new links kUnset, the old buffers freed after the stream drains.
int ensure_pool_nodes(oit_ctx* c, uint64_t need, hipStream_t s) {
  if (need <= c->ll_alloc) return OIT_OK;
  if (need > kMaxPoolNodes) return OIT_ERR_CAPACITY;
  uint64_t cap = c->ll_alloc ? c->ll_alloc : 1024;
  while (cap < need) cap *= 2;
  if (cap > kMaxPoolNodes) cap = kMaxPoolNodes;
  const bool cov = coverage(c);
  uint64_t* ncd = nullptr;
  uint32_t* nnext = nullptr;
  float4* nfull = nullptr;
  uint32_t* nmask = nullptr;
  int rc = dev_alloc(&ncd, cap);
  if (!rc) rc = dev_alloc(&nnext, cap);
  if (!rc) rc = dev_alloc(&nfull, cap);
  if (!rc && cov) rc = dev_alloc(&nmask, cap);
  if (rc) {
    dev_free(&ncd);
    dev_free(&nnext);
    dev_free(&nfull);
    dev_free(&nmask);
    return rc;
  }
  const uint64_t old = c->ll_alloc;
  hipError_t e = hipSuccess;
  if (old) {
    e = hipMemcpyAsync(ncd, c->ll_cd, old * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(nnext, c->ll_next, old * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && c->il_full)
      e = hipMemcpyAsync(nfull, c->il_full, old * sizeof(float4), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && nmask && c->ll_mask)
      e = hipMemcpyAsync(nmask, c->ll_mask, old * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
  }
  if (e == hipSuccess) e = oit::launch_fill_u32(nnext + old, cap - old, oit::kUnset, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    dev_free(&ncd);
    dev_free(&nnext);
    dev_free(&nfull);
    dev_free(&nmask);
    return status_of(e);
  }
  wait_idle(c);
  dev_free(&c->ll_cd);
  dev_free(&c->ll_next);
  dev_free(&c->il_full);
  dev_free(&c->ll_mask);
  c->ll_cd = ncd;
  c->ll_next = nnext;
  c->il_full = nfull;
  c->ll_mask = nmask;
  c->ll_alloc = c->ll_cap = cap;
  c->cap_il_full = cap;
  c->cap_ll_mask = nmask ? cap : 0;
  c->bytes_abuf = cap * ((cov ? 16u : 12u) + 16u);
  return OIT_OK;
}

}  // namespace

extern "C" {

const char* oit_status_string(int status) {
  switch (status) {
    case OIT_OK: return "ok";
    case OIT_ERR_INVALID_ARG: return "invalid argument";
    case OIT_ERR_OOM: return "out of device memory";
    case OIT_ERR_OVER_BUDGET: return "over memory budget";
    case OIT_ERR_DEVICE: return "device error";
    case OIT_ERR_UNSUPPORTED: return "unsupported configuration";
    case OIT_ERR_CAPACITY: return "capacity exceeded";
    case OIT_ERR_STATE: return "invalid call order";
    case OIT_ERR_NOT_READY: return "not ready";
    default: return "unknown status";
  }
}

int oit_create(oit_ctx** out, int device) {
  if (!out) return OIT_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return OIT_ERR_DEVICE;
  }
  oit_ctx* c = new (std::nothrow) oit_ctx();
  if (!c) return OIT_ERR_OOM;
  c->device = device;
  DeviceGuard g(device);
  oit::CodecTables host;
  oit::build_codec_tables(&host);
  int rc = dev_alloc(&c->d_tab, 1);
  if (!rc) rc = dev_alloc(&c->d_counters, kCounterWords + 1);
  if (!rc && hipMemcpy(c->d_tab, &host, sizeof host, hipMemcpyHostToDevice) != hipSuccess) rc = OIT_ERR_DEVICE;
  if (!rc && hipMemset(c->d_counters, 0, (kCounterWords + 1) * sizeof(uint32_t)) != hipSuccess) rc = OIT_ERR_DEVICE;
  if (!rc && hipEventCreateWithFlags(&c->work_ev, hipEventDisableTiming) != hipSuccess) rc = OIT_ERR_DEVICE;
  if (rc) {
    dev_free(&c->d_tab);
    dev_free(&c->d_counters);
    delete c;
    return rc;
  }
  *out = c;
  return OIT_OK;
}

int oit_destroy(oit_ctx* c) {
  if (!c) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(c->device);
  (void)hipDeviceSynchronize();
  free_frame_buffers(c);
  dev_free(&c->d_tab);
  dev_free(&c->d_counters);
  if (c->work_ev) (void)hipEventDestroy(c->work_ev);
  for (auto& p : c->events) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  delete c;
  return OIT_OK;
}

int oit_configure(oit_ctx* c, const oit_config* cfg) {
  if (!c || !cfg) return OIT_ERR_INVALID_ARG;
  const uint32_t alg = cfg->algorithm;
  if (alg > OIT_WEIGHTED) return OIT_ERR_INVALID_ARG;
  if (alg != OIT_WEIGHTED && !valid_layers(cfg->layers)) return OIT_ERR_INVALID_ARG;
  if (cfg->width == 0 || cfg->height == 0) return OIT_ERR_INVALID_ARG;
  const uint32_t S = cfg->samples ? cfg->samples : 1;
  if (S != 1 && S != 2 && S != 4 && S != 8) return OIT_ERR_INVALID_ARG;
  const uint64_t T = (uint64_t)cfg->width * cfg->height * S;
  if (T >= 0xFFFFFFFFull) return OIT_ERR_INVALID_ARG;
  for (int k = 0; k < 4; ++k)
    if (!std::isfinite(cfg->background[k])) return OIT_ERR_INVALID_ARG;
  const uint32_t cov = cfg->coverage_samples;
  if (cov > 1) {  // OIT_COVERAGE_SHADING
    if (cov != 2 && cov != 4 && cov != 8) return OIT_ERR_INVALID_ARG;
    if (S != 1) return OIT_ERR_INVALID_ARG;  // per-pixel fragments
    if (alg == OIT_LOOP || alg == OIT_LOOP64 || alg == OIT_WEIGHTED) return OIT_ERR_UNSUPPORTED;  // README.md:33-34
    if (alg == OIT_LINKEDLIST && (cfg->variant & OIT_VARIANT_LL_SPLIT_APPEND)) return OIT_ERR_UNSUPPORTED;
  }
  uint64_t ll_cap = 0;
  if (pooled(alg)) {  // oit.cpp:119-126 (the interlock's pool also grows to the frame's fragments)
    ll_cap = cfg->ll_pool_nodes;
    if (ll_cap == 0) ll_cap = (uint64_t)(cfg->ll_nodes_per_pixel ? cfg->ll_nodes_per_pixel : 10) * T;
    if (ll_cap < 1) ll_cap = 1;
    if (ll_cap > kMaxPoolNodes) return OIT_ERR_CAPACITY;
  }
  DeviceGuard g(c->device);
  c->configured = false;
  c->cfg = *cfg;
  c->cfg.samples = S;
  if (alg == OIT_WEIGHTED && !valid_layers(c->cfg.layers)) c->cfg.layers = 1;
  c->T = (uint32_t)T;
  const uint32_t K = c->cfg.layers;
  // buffers of the other algorithms go; this algorithm's are reused if they fit
  if (!pooled(alg)) {
    release(c, &c->ll_head, &c->cap_ll_head);
    uint64_t dummy = c->ll_alloc;
    release(c, &c->ll_cd, &dummy);
    release(c, &c->ll_next, &c->ll_alloc);
    c->ll_cap = 0;
  }
  if (alg != OIT_SIMPLE && alg != OIT_SPINLOCK) {
    release(c, &c->kb_count, &c->cap_kb_count);
    uint64_t d1 = c->cap_kb_elems, d2 = c->cap_kb_elems;
    release(c, &c->kb_col, &d1);
    release(c, &c->kb_key, &d2);
    release(c, &c->kb_id, &c->cap_kb_elems);
  }
  if (alg != OIT_SPINLOCK) {
    release(c, &c->kb_lock, &c->cap_kb_lock);
    release(c, &c->kb_depth, &c->cap_kb_depth);
    release(c, &c->kb_filled, &c->cap_kb_filled);
  }
  if (!pooled(alg)) release(c, &c->ll_mask, &c->cap_ll_mask);
  if (alg != OIT_INTERLOCK) release(c, &c->il_full, &c->cap_il_full);
  if (alg != OIT_LOOP) {
    uint64_t d = c->cap_l32;
    release(c, &c->l32_z, &d);
    release(c, &c->l32_c, &c->cap_l32);
  }
  if (alg != OIT_SIMPLE && alg != OIT_SPINLOCK) release(c, &c->kb_mask, &c->cap_kb_mask);
  if (alg != OIT_LOOP64) {
    release(c, &c->l64, &c->cap_l64);
    release(c, &c->l64_cnt, &c->cap_l64_cnt);
    release(c, &c->l64_big, &c->cap_l64_big);
  }
  if (alg != OIT_WEIGHTED) {
    uint64_t d = c->cap_wb;
    release(c, &c->wb_accum, &d);
    release(c, &c->wb_reveal, &c->cap_wb);
  }
  int rc = OIT_OK;
  bool ll_grown = false;
  switch (alg) {
    case OIT_LINKEDLIST:
    case OIT_INTERLOCK: {
      rc = reserve(c, &c->ll_head, &c->cap_ll_head, T, &ll_grown);
      if (!rc && ll_cap > c->ll_alloc) {  // node payloads and links share one capacity
        uint64_t d = c->ll_alloc;
        rc = reserve(c, &c->ll_cd, &d, ll_cap, &ll_grown);
        if (!rc) rc = reserve(c, &c->ll_next, &c->ll_alloc, ll_cap, &ll_grown);
      }
      if (!rc && cov > 1) rc = reserve(c, &c->ll_mask, &c->cap_ll_mask, c->ll_alloc);
      if (!rc && cov <= 1) release(c, &c->ll_mask, &c->cap_ll_mask);
      if (!rc && alg == OIT_INTERLOCK) rc = reserve(c, &c->il_full, &c->cap_il_full, c->ll_alloc);
      c->ll_cap = alg == OIT_INTERLOCK ? c->ll_alloc : ll_cap;
      c->bytes_abuf = c->ll_cap * ((cov > 1 ? 16u : 12u) + (alg == OIT_INTERLOCK ? 16u : 0u));
      c->bytes_aux = T * 4u;
      break;
    }
    case OIT_LOOP: {  // oit.cpp:127-133: depth and colour planes
      if (T * K > c->cap_l32) {
        uint64_t d = c->cap_l32;
        rc = reserve(c, &c->l32_z, &d, T * K);
        if (!rc) rc = reserve(c, &c->l32_c, &c->cap_l32, T * K);
      }
      c->bytes_abuf = T * K * 12u;
      c->bytes_aux = 0;
      break;
    }
    case OIT_SIMPLE:
    case OIT_SPINLOCK: {  // oit.cpp:102-118 (the Spinlock's count lives in its lock word)
      if (alg == OIT_SIMPLE) rc = reserve(c, &c->kb_count, &c->cap_kb_count, T);
      else release(c, &c->kb_count, &c->cap_kb_count);
      if (!rc && T * K > c->cap_kb_elems) {
        uint64_t d1 = c->cap_kb_elems, d2 = c->cap_kb_elems;
        rc = reserve(c, &c->kb_col, &d1, T * K);
        if (!rc) rc = reserve(c, &c->kb_key, &d2, T * K);
        if (!rc) rc = reserve(c, &c->kb_id, &c->cap_kb_elems, T * K);
      }
      if (!rc && alg == OIT_SPINLOCK) rc = reserve(c, &c->kb_lock, &c->cap_kb_lock, T);
      if (!rc && alg == OIT_SPINLOCK) rc = reserve(c, &c->kb_depth, &c->cap_kb_depth, T);
      if (!rc && alg == OIT_SPINLOCK) rc = reserve(c, &c->kb_filled, &c->cap_kb_filled, T);
      if (!rc && cov > 1) rc = reserve(c, &c->kb_mask, &c->cap_kb_mask, T * K);
      if (!rc && cov <= 1) release(c, &c->kb_mask, &c->cap_kb_mask);
      c->bytes_abuf = T * K * (cov > 1 ? 16u : 12u);
      c->bytes_aux = T * (alg == OIT_SPINLOCK ? 12u : 4u);
      break;
    }
    case OIT_LOOP64: {  // oit.cpp:134-140
      rc = reserve(c, &c->l64, &c->cap_l64, T * K);
      if (!rc) rc = reserve(c, &c->l64_cnt, &c->cap_l64_cnt, T);
      c->bytes_abuf = T * K * 8u;
      c->bytes_aux = T * 4u;
      break;
    }
    case OIT_WEIGHTED: {  // oit.h:234-235 (binary32 here)
      if (T > c->cap_wb) {
        uint64_t d = c->cap_wb;
        rc = reserve(c, &c->wb_accum, &d, T);
        if (!rc) rc = reserve(c, &c->wb_reveal, &c->cap_wb, T);
      }
      c->bytes_abuf = T * 20u;
      c->bytes_aux = 0;
      break;
    }
  }
  if (!rc && uses_tail_lists(c)) {
    rc = reserve(c, &c->tl_head, &c->cap_tl_head, T);
    if (!rc && cfg->max_fragments && cfg->max_fragments > c->tl_cap) {
      const uint64_t cap = cfg->max_fragments < kMaxFrameFragments ? cfg->max_fragments : kMaxFrameFragments;
      uint64_t d1 = c->tl_cap, d2 = c->tl_cap, d3 = c->tl_cap, d4 = oit::tail_regions(c->tl_cap);
      rc = reserve(c, &c->tl_next, &d1, cap);
      if (!rc) rc = reserve(c, &c->tl_rgba, &d2, cap);
      if (!rc) rc = reserve(c, &c->tl_rec, &d3, cap);
      if (!rc) rc = reserve(c, &c->tl_qn, &d4, oit::tail_regions(cap));
      if (!rc) c->tl_cap = (uint32_t)cap;
    }
    if (!rc && cov > 1 && c->tl_cap) rc = reserve(c, &c->tl_mask, &c->cap_tl_mask, c->tl_cap);
    if (!rc && cov <= 1) release(c, &c->tl_mask, &c->cap_tl_mask);
  } else if (!rc) {
    release(c, &c->tl_head, &c->cap_tl_head);
    uint64_t d1 = c->tl_cap, d2 = c->tl_cap, d3 = c->tl_cap, d4 = 0;
    release(c, &c->tl_next, &d1);
    release(c, &c->tl_rgba, &d2);
    release(c, &c->tl_rec, &d3);
    release(c, &c->tl_qn, &d4);
    release(c, &c->tl_mask, &c->cap_tl_mask);
    c->tl_cap = 0;
  }
  if (rc) {
    free_frame_buffers(c);
    return rc;
  }
  c->tl_head_valid = false;
  c->tl_clean = false;
  c->l64_clean = false;  // slots beyond the last frame's T x K may hold an older frame's keys
  c->l32_clean = false;
  c->wb_clean = false;
  c->kb_clean = false;
  if (pooled(alg) && ll_grown) {
    // a new allocation is filled clean (every head 0, every link kUnset) by
    // the next clear, on its stream
    c->ll_clean = false;
    c->ll_dirty_upto = c->ll_alloc;
  }
  if (!pooled(alg)) {
    c->ll_clean = false;
    c->ll_dirty_upto = 0;
  }
  c->counters_clean = false;
  c->stats_moved = false;
  c->configured = true;
  c->frame_open = false;
  c->resolved = false;
  c->frame_frags = 0;
  c->appends = 0;
  c->segments.clear();
  return OIT_OK;
}

int oit_trim(oit_ctx* c) {
  if (!c) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(c->device);
  free_frame_buffers(c);
  c->ll_clean = false;
  c->ll_dirty_upto = 0;
  c->frame_open = false;
  c->resolved = false;
  return OIT_OK;
}

int oit_query_sizes(const oit_ctx* c, oit_sizes* out) {
  if (!c || !out) return OIT_ERR_INVALID_ARG;
  if (!c->configured) return OIT_ERR_STATE;
  std::memset(out, 0, sizeof *out);
  out->abuffer_bytes = c->bytes_abuf;
  out->aux_bytes = c->bytes_aux + (c->tl_head ? (uint64_t)c->T * 4u : 0u);
  out->tail_bytes = (uint64_t)c->tl_cap * (4u + 16u + 8u) + oit::tail_regions(c->tl_cap) * 4u;
  out->total_bytes = out->abuffer_bytes + out->aux_bytes + out->tail_bytes;
  out->ll_capacity = c->ll_cap;
  switch (c->cfg.algorithm) {
    case OIT_LINKEDLIST: out->bytes_per_element = 12; break;
    case OIT_INTERLOCK: out->bytes_per_element = 28; break;
    case OIT_LOOP: out->bytes_per_element = 12; break;
    case OIT_SIMPLE:
    case OIT_SPINLOCK: out->bytes_per_element = 12; break;
    case OIT_LOOP64: out->bytes_per_element = 8; break;
    default: out->bytes_per_element = 20; break;
  }
  out->targets = c->T;
  out->resident_bytes = c->cap_ll_head * 4u + c->ll_alloc * 12u + c->cap_kb_count * 4u + c->cap_kb_lock * 4u +
                        c->cap_kb_depth * 4u + c->cap_kb_filled * 4u + c->cap_kb_elems * 12u + c->cap_l64 * 8u + c->cap_l64_cnt * 4u + c->cap_l64_big * 8u + c->cap_wb * 20u +
                        c->cap_tl_head * 4u + (uint64_t)c->tl_cap * (4u + 16u + 8u) + oit::tail_regions(c->tl_cap) * 4u +
                        (c->cap_ll_mask + c->cap_kb_mask + c->cap_tl_mask) * 4u + c->cap_il_full * 16u +
                        c->cap_l32 * 12u + c->cap_opq * 8u + c->cap_opq_rgba * 16u;
  return OIT_OK;
}

int oit_clear(oit_ctx* c, void* stream) {
  if (!c) return OIT_ERR_INVALID_ARG;
  if (!c->configured) return OIT_ERR_STATE;
  DeviceGuard g(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t T = c->T, K = c->cfg.layers;
  const bool ll = pooled(c->cfg.algorithm);
  const bool need_ll_fill = ll && !c->ll_clean;
  // only the Linked List's kernels write the frame counters (its resolve
  // leaves them at 0): the other algorithms' stay as the last fill left them
  const bool need_counters = !c->counters_clean;
  // tail lists: filled unless no tail was pushed since the last fill (a
  // Linked List frame whose pool could not overflow)
  const bool need_tail = c->tl_head && (!ll || c->tl_next) && !c->tl_clean;
  // every fill of the clear in one launch, none where nothing needs one: the
  // steady state of the Linked List (its resolve left heads at 0, links at
  // kUnset and the counters at 0: the clear of oitRender.cpp:206-212 is done
  // by the reader) and of WBOIT (its composite left the accumulators at 0).
  // Tail-list heads not filled here are filled on first use
  // (ensure_tail_capacity).
  oit::FillRange fr[8];
  int nf = 0;
  auto fill = [&](void* p, uint64_t n, uint32_t v) { fr[nf++] = oit::FillRange{static_cast<uint32_t*>(p), n, v}; };
  if (need_counters) fill(c->d_counters, 16, 0u);
  bool l64_fill = false, l32_fill = false, wb_fill = false, kb_fill = false;
  switch (c->cfg.algorithm) {
    case OIT_LOOP:  // depths to all ones (oitRender.cpp:242-265), colour slots to 0; the composite leaves them so
      l32_fill = !c->l32_clean;
      if (l32_fill) {
        fill(c->l32_z, T * K, kEmpty);
        fill(c->l32_c, 2 * T * K, 0u);
      }
      break;
    case OIT_INTERLOCK:
    case OIT_LINKEDLIST:  // head and counter to 0, oitRender.cpp:206-212; links of a dirty range to kUnset
      // every allocated head: a larger earlier frame may have left some
      if (need_ll_fill) {
        fill(c->ll_head, c->cap_ll_head, 0u);
        fill(c->ll_next, c->ll_dirty_upto ? c->ll_dirty_upto : c->ll_alloc, oit::kUnset);
      }
      break;
    case OIT_SIMPLE:  // count to 0, oitRender.cpp:161-169; the composite leaves it so
      kb_fill = !c->kb_clean;
      if (kb_fill) fill(c->kb_count, T, 0u);
      break;
    case OIT_SPINLOCK:  // count, spin to 0, depth to ~0, oitRender.cpp:343-354 (the count is in the lock word); likewise
      kb_fill = !c->kb_clean;
      if (kb_fill) {
        fill(c->kb_lock, T, 0u);
        fill(c->kb_depth, T, kEmpty);
        fill(c->kb_filled, T, 0u);
      }
      break;
    case OIT_LOOP64:  // every byte to 0xFF, oitRender.cpp:307; the composite leaves its slots empty
      l64_fill = !c->l64_clean;
      if (l64_fill) {
        fill(c->l64, 2 * T * K, kEmpty);
        fill(c->l64_cnt, T, 0u);
      }
      break;
    case OIT_WEIGHTED:  // accum 0, reveal 1 (log2 0), oitRender.cpp:394-397; the composite leaves them so
      wb_fill = !c->wb_clean;
      if (wb_fill) {
        fill(c->wb_accum, 4 * T, 0u);
        fill(c->wb_reveal, T, 0u);
      }
      break;
  }
  if (need_tail) {
    fill(c->tl_head, T, 0u);  // empty tail lists
    if (c->tl_qn) fill(c->tl_qn, oit::tail_regions(c->tl_cap), 0u);  // and an empty tail queue
  }
  hipError_t e = hipSuccess;
  if (nf > 0) e = staged(c, kStClear, s, [&]() -> hipError_t { return oit::launch_fill_ranges(fr, nf, s); });
  if (e == hipSuccess && l64_fill) c->l64_clean = true;
  if (e == hipSuccess && l32_fill) c->l32_clean = true;
  if (e == hipSuccess && wb_fill) c->wb_clean = true;
  if (e == hipSuccess && kb_fill) c->kb_clean = true;
  if (e == hipSuccess) c->counters_clean = true;
  if (e == hipSuccess && ll) {
    c->ll_clean = true;
    c->ll_dirty_upto = 0;
  }
  c->stats_moved = false;
  if (e == hipSuccess && need_tail) c->tl_clean = true;
  c->tl_head_valid = e == hipSuccess && c->tl_head != nullptr && c->tl_clean;
  c->frame_frags = 0;
  c->appends = 0;
  c->opq_active = false;
  c->opq_frags = 0;
  c->segments.clear();
  c->frame_open = e == hipSuccess;
  c->resolved = false;
  if (nf > 0) mark_work(c, s);
  return status_of(e);
}

int oit_draw_opaque(oit_ctx* c, const oit_fragments* f, void* stream) {
  if (!c || !f) return OIT_ERR_INVALID_ARG;
  if (!c->configured) return OIT_ERR_STATE;
  if (!c->frame_open || c->resolved || c->appends > 0) return OIT_ERR_STATE;  // opaque first, oitRender.cpp:113-125
  if (f->count > 0 && (!f->target || !f->depth || !f->rgba)) return OIT_ERR_INVALID_ARG;
  if (f->count > 0 && coverage(c) && !f->coverage) return OIT_ERR_INVALID_ARG;
  if (f->count > kMaxFrameFragments || c->opq_frags + f->count > kMaxFrameFragments) return OIT_ERR_CAPACITY;
  DeviceGuard g(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSuccess;
  if (!c->opq_active) {  // the frame's first opaque draw: depth 1.0, clear colour (oitRender.cpp:89-91)
    const uint64_t n = (uint64_t)c->T * (coverage(c) ? c->cfg.coverage_samples : 1u);
    int rc = reserve(c, &c->opq_z, &c->cap_opq, n);
    if (!rc) rc = reserve(c, &c->opq_rgba, &c->cap_opq_rgba, n);
    if (rc) return rc;
    const float4 bg = make_float4(c->cfg.background[0], c->cfg.background[1], c->cfg.background[2],
                                  c->cfg.background[3]);
    e = staged(c, kStClear, s, [&] { return oit::launch_opaque_clear(c->opq_z, c->opq_rgba, n, bg, s); });
    if (e != hipSuccess) return status_of(e);
    c->opq_active = true;
  }
  if (f->count > 0) {
    oit::FragView v;
    v.target = f->target;
    v.depth = f->depth;
    v.rgba = reinterpret_cast<const float4*>(f->rgba);
    v.count = (uint32_t)f->count;
    v.id_base = (uint32_t)c->opq_frags;
    v.mask = coverage(c) ? f->coverage : nullptr;
    const oit::Common k = common_of(c);
    e = staged(c, kStStore, s, [&] { return oit::launch_opaque_draw(k, v, c->opq_z, c->opq_rgba, s); });
    if (e != hipSuccess) return status_of(e);
    c->opq_frags += f->count;
  }
  mark_work(c, s);
  return OIT_OK;
}

int oit_append(oit_ctx* c, const oit_fragments* f, void* stream) {
  if (!c || !f) return OIT_ERR_INVALID_ARG;
  if (!c->configured) return OIT_ERR_STATE;
  if (!c->frame_open || c->resolved) return OIT_ERR_STATE;
  if (f->count == 0) return OIT_OK;
  if (!f->target || !f->depth || !f->rgba) return OIT_ERR_INVALID_ARG;
  if (coverage(c) && !f->coverage) return OIT_ERR_INVALID_ARG;  // coverage shading needs the sample masks
  if (f->count > kMaxFrameFragments || c->frame_frags + f->count > kMaxFrameFragments) return OIT_ERR_CAPACITY;
  DeviceGuard g(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t end = c->frame_frags + f->count;
  const uint32_t alg = c->cfg.algorithm;
  if (uses_tail_lists(c)) {
    const bool need = alg != OIT_LINKEDLIST || end + 1 > c->ll_cap;  // LL: only when the pool can overflow
    if (need) {
      int rc = ensure_tail_capacity(c, end, s);
      if (rc) return rc;
      c->tl_clean = false;  // this append may push tail fragments
    }
  }
  oit::FragView v;
  v.target = f->target;
  v.depth = f->depth;
  v.rgba = reinterpret_cast<const float4*>(f->rgba);
  v.count = (uint32_t)f->count;
  v.id_base = (uint32_t)c->frame_frags;
  v.mask = coverage(c) ? f->coverage : nullptr;
  const oit::Common k = common_of(c);
  const oit::TailBuf tb = tail_of(c);
  hipError_t e = hipSuccess;
  switch (alg) {
    case OIT_LINKEDLIST: {
      oit::LLBuf ll{c->ll_head, c->ll_cd, c->ll_next, c->ll_cap, c->d_counters + 1, c->d_counters};
      ll.err = dev_err(c);
      ll.mask = coverage(c) ? c->ll_mask : nullptr;
      c->ll_clean = false;
      c->counters_clean = false;
      const uint64_t touched = end + 1 < c->ll_cap ? end + 1 : c->ll_cap;
      if (touched > c->ll_dirty_upto) c->ll_dirty_upto = touched;
      if (c->cfg.variant & OIT_VARIANT_LL_SPLIT_APPEND) {
        e = staged(c, kStStore, s, [&] { return oit::launch_ll_store(k, v, ll, tb, s); });
        if (e == hipSuccess) e = staged(c, kStLink, s, [&] { return oit::launch_ll_link(k, v, ll, s); });
      } else {
        e = staged(c, kStStore, s, [&] {
          return oit::launch_ll_append(k, v, ll, tb, (c->cfg.variant & OIT_VARIANT_LL_APPEND_8) ? 8 : 4, s);
        });
      }
      break;
    }
    case OIT_INTERLOCK: {  // nodes of every fragment; the critical sections run in the resolve
      int rc = ensure_pool_nodes(c, end + 1, s);
      if (rc) return rc;
      oit::LLBuf ll{c->ll_head, c->ll_cd, c->ll_next, c->ll_cap, c->d_counters + 1, c->d_counters};
      ll.err = dev_err(c);
      ll.mask = coverage(c) ? c->ll_mask : nullptr;
      ll.full = c->il_full;
      c->ll_clean = false;
      c->counters_clean = false;
      if (end + 1 > c->ll_dirty_upto) c->ll_dirty_upto = end + 1;
      const oit::TailBuf none{};
      e = staged(c, kStStore, s, [&] { return oit::launch_ll_append(k, v, ll, none, 4, s); });
      break;
    }
    case OIT_LOOP: {  // depth pass now; the colour pass runs over the recorded batches at resolve
      c->l32_clean = false;
      e = staged(c, kStStore, s, [&] { return oit::launch_loop32_depth(k, v, c->l32_z, s); });
      c->segments.push_back(v);
      break;
    }
    case OIT_SIMPLE: {
      c->kb_clean = false;
      oit::KBuf kb{c->kb_count, nullptr, nullptr, c->kb_col, c->kb_key, c->kb_id};
      kb.err = dev_err(c);
      kb.mask = coverage(c) ? c->kb_mask : nullptr;
      e = staged(c, kStStore, s, [&] { return oit::launch_simple_store(k, v, kb, tb, s); });
      break;
    }
    case OIT_SPINLOCK: {
      c->kb_clean = false;
      oit::KBuf kb{nullptr, c->kb_lock, c->kb_depth, c->kb_col, c->kb_key, c->kb_id};
      kb.err = dev_err(c);
      kb.filled = c->kb_filled;
      kb.mask = coverage(c) ? c->kb_mask : nullptr;
      e = staged(c, kStStore, s, [&] { return oit::launch_spin_store(k, v, kb, tb, s); });
      break;
    }
    case OIT_LOOP64: {  // counted now; stored at resolve, once every pixel's count is known
      c->l64_clean = false;
      e = staged(c, kStStore, s, [&] { return oit::launch_loop64_count(k, v, c->l64_cnt, s); });
      c->segments.push_back(v);
      break;
    }
    case OIT_WEIGHTED: {
      oit::WBBuf wb{c->wb_accum, c->wb_reveal};
      c->wb_clean = false;
      e = staged(c, kStStore, s, [&] {
        return oit::launch_wboit_store(k, v, wb, c->cfg.wboit_view_near, c->cfg.wboit_view_span, s);
      });
      break;
    }
  }
  mark_work(c, s);
  if (e != hipSuccess) return status_of(e);
  c->frame_frags = end;
  c->appends++;
  return OIT_OK;
}

int oit_resolve_debug(oit_ctx* c, float* out_rgba, uint32_t* fk_key, uint32_t* fk_val, uint32_t* counts,
                      void* stream) {
  if (!c || !out_rgba) return OIT_ERR_INVALID_ARG;
  if (!c->configured || !c->frame_open) return OIT_ERR_STATE;
  if (c->resolved) return OIT_ERR_STATE;  // one composite per frame (the Linked List recycles its heads)
  DeviceGuard g(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t alg = c->cfg.algorithm;
  const bool link_tails = !c->resolved;
  if (link_tails && alg == OIT_LOOP64 && c->cfg.tail_blend && !c->segments.empty()) {
    int rc = ensure_tail_capacity(c, c->frame_frags, s);
    if (rc) return rc;
    c->tl_clean = false;  // the classification pushes tail fragments
  }
  const oit::Common k = common_of(c);
  const oit::TailBuf tb = tail_of(c);
  oit::ResolveOut o{reinterpret_cast<float4*>(out_rgba), fk_key, fk_val, counts};
  hipError_t e = hipSuccess;
  if (link_tails && alg == OIT_LOOP64 && !c->segments.empty()) {  // the stores of every segment (counted at append)
    const int chunks = (c->cfg.variant & OIT_VARIANT_STORE_CHUNKS_1) ? 1
                       : (c->cfg.variant & OIT_VARIANT_STORE_CHUNKS_4) ? 4 : 2;
    uint64_t masks = 0;  // one per 64-fragment chunk of each segment
    for (const auto& v : c->segments) masks += (v.count + 63u) / 64u;
    int rc = reserve(c, &c->l64_big, &c->cap_l64_big, masks ? masks : 1);
    if (rc) return rc;
    uint64_t m0 = 0;
    for (const auto& v : c->segments) {
      uint64_t* big = c->l64_big + m0;
      e = staged(c, kStStore, s, [&] { return oit::launch_loop64_store(k, v, c->l64, c->l64_cnt, big, chunks, s); });
      if (e != hipSuccess) return status_of(e);
      m0 += (v.count + 63u) / 64u;
    }
    if (c->cfg.tail_blend && tb.head) {
      m0 = 0;
      for (const auto& v : c->segments) {
        const uint64_t* big = c->l64_big + m0;
        e = staged(c, kStClassify, s, [&] { return oit::launch_loop64_classify(k, v, c->l64, big, tb, s); });
        if (e != hipSuccess) return status_of(e);
        m0 += (v.count + 63u) / 64u;
      }
    }
  }
  if (link_tails && alg == OIT_LOOP) {  // the second geometry pass (oitLoop.frag.glsl:105-180)
    for (const auto& v : c->segments) {
      e = staged(c, kStClassify, s, [&] { return oit::launch_loop32_color(k, v, c->l32_z, c->l32_c, tb, s); });
      if (e != hipSuccess) return status_of(e);
    }
  }
  // Linked List: only fragments whose node would be at or past the pool's
  // end (id >= cap - 1) can be tails
  const uint64_t first_tail = alg == OIT_LINKEDLIST ? (c->ll_cap ? c->ll_cap - 1 : 0) : 0;
  const bool no_tails = !tb.next || first_tail >= c->frame_frags;
  // Loop64's classification, which runs after every store, links its tail
  // fragments itself; the other algorithms queue them for this pass
  if (link_tails && tb.head && !no_tails && alg != OIT_LOOP64) {
    e = staged(c, kStTailLink, s, [&] {
      return oit::launch_tail_link(tb, (uint32_t)first_tail, (uint32_t)c->frame_frags, s);
    });
    if (e != hipSuccess) return status_of(e);
  }
  c->resolved = true;
  oit::TailBuf tbr = tb;
  if (no_tails) tbr.head = nullptr;  // no tail records can exist
  switch (alg) {
    case OIT_LINKEDLIST: {
      oit::LLBuf ll{c->ll_head, c->ll_cd, c->ll_next, c->ll_cap, c->d_counters + 1, c->d_counters};
      ll.err = dev_err(c);
      const bool pair = !(c->cfg.variant & OIT_VARIANT_LL_SINGLE_WALK);
      // at K = 32 the keys-only composite holds 95 VGPRs against 129 (five
      // waves per SIMD against three): 3840x2160x4 frame 4.70 -> 4.37 ms
      const bool keys = (c->cfg.variant & OIT_VARIANT_LL_KEYS_ONLY) != 0 || c->cfg.layers >= 32;
      if (coverage(c)) {
        ll.mask = c->ll_mask;
        e = staged(c, kStResolve, s, [&] { return oit::launch_ll_resolve_cov(k, ll, tbr, o, s); });
      } else {
        e = staged(c, kStResolve, s, [&] { return oit::launch_ll_resolve_reset(k, ll, tbr, o, pair, keys, s); });
      }
      if (e == hipSuccess) {
        c->ll_clean = true;
        c->ll_dirty_upto = 0;
        c->counters_clean = true;
        c->stats_moved = true;
      }
      break;
    }
    case OIT_SIMPLE:
    case OIT_SPINLOCK: {
      oit::KBuf kb{alg == OIT_SPINLOCK ? nullptr : c->kb_count, alg == OIT_SPINLOCK ? c->kb_lock : nullptr,
                   c->kb_depth, c->kb_col, c->kb_key, c->kb_id};
      kb.filled = alg == OIT_SPINLOCK ? c->kb_filled : nullptr;
      if (coverage(c)) {
        kb.mask = c->kb_mask;
        e = staged(c, kStResolve, s, [&] { return oit::launch_kbuf_resolve_cov(k, kb, tbr, o, s); });
      } else {
        e = staged(c, kStResolve, s, [&] { return oit::launch_kbuf_resolve(k, kb, tbr, o, s); });
      }
      c->kb_clean = e == hipSuccess;  // the composite set every word it read back to its clear value
      break;
    }
    case OIT_LOOP64:
      e = staged(c, kStResolve, s, [&] { return oit::launch_loop64_resolve(k, c->l64, c->l64_cnt, tbr, o, s); });
      c->l64_clean = e == hipSuccess;  // the composite emptied every slot it read
      break;
    case OIT_LOOP:
      e = staged(c, kStResolve, s, [&] { return oit::launch_loop32_resolve(k, c->l32_z, c->l32_c, tbr, o, s); });
      c->l32_clean = e == hipSuccess;  // the composite emptied every slot it read
      break;
    case OIT_INTERLOCK: {
      oit::LLBuf ll{c->ll_head, c->ll_cd, c->ll_next, c->ll_cap, c->d_counters + 1, c->d_counters};
      ll.err = dev_err(c);
      ll.mask = coverage(c) ? c->ll_mask : nullptr;
      ll.full = c->il_full;
      e = staged(c, kStResolve, s, [&] { return oit::launch_interlock_resolve(k, ll, o, s); });
      if (e == hipSuccess) {
        c->ll_clean = true;
        c->ll_dirty_upto = 0;
        c->counters_clean = true;
        c->stats_moved = true;
      }
      break;
    }
    case OIT_WEIGHTED: {
      oit::WBBuf wb{c->wb_accum, c->wb_reveal};
      e = staged(c, kStResolve, s, [&] { return oit::launch_wboit_resolve(k, wb, o, s); });
      c->wb_clean = e == hipSuccess;  // the composite set every accumulator back to the clear values
      break;
    }
  }
  mark_work(c, s);
  return status_of(e);
}

int oit_finalize_srgb8(oit_ctx* c, const float* in_rgba, uint32_t width, uint32_t height, uint32_t samples,
                       uint32_t supersample, uint8_t* out_rgba8, void* stream) {
  if (!c || !in_rgba || !out_rgba8 || width == 0 || height == 0) return OIT_ERR_INVALID_ARG;
  if (samples != 1 && samples != 2 && samples != 4 && samples != 8) return OIT_ERR_INVALID_ARG;
  if (supersample < 1 || supersample > 8) return OIT_ERR_INVALID_ARG;
  if ((uint64_t)width * height * samples * supersample * supersample >= 0xFFFFFFFFull) return OIT_ERR_INVALID_ARG;
  if (((uintptr_t)in_rgba & 15u) || ((uintptr_t)out_rgba8 & 3u)) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(c->device);
  return status_of(oit::launch_finalize_srgb8(reinterpret_cast<const float4*>(in_rgba), width, height, samples,
                                              supersample, c->d_tab, reinterpret_cast<uint32_t*>(out_rgba8),
                                              static_cast<hipStream_t>(stream)));
}

int oit_resolve(oit_ctx* c, float* out_rgba, void* stream) {
  return oit_resolve_debug(c, out_rgba, nullptr, nullptr, nullptr, stream);
}

int oit_get_stats(oit_ctx* c, oit_stats* out, void* stream) {
  if (!c || !out) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t h[16] = {0};
  uint32_t err = 0;
  const uint32_t* src = c->d_counters + (c->stats_moved ? 16 : 0);
  if (hipMemcpyAsync(h, src, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess) return OIT_ERR_DEVICE;
  if (hipMemcpyAsync(&err, dev_err(c), sizeof err, hipMemcpyDeviceToHost, s) != hipSuccess) return OIT_ERR_DEVICE;
  if (hipStreamSynchronize(s) != hipSuccess) return OIT_ERR_DEVICE;
  // tail fragments: counted per region of 1024 ids where they were found
  uint64_t tails = 0;
  if (c->tl_qn && c->tl_head_valid) {
    std::vector<uint32_t> rg(oit::tail_regions(c->tl_cap));
    if (hipMemcpyAsync(rg.data(), c->tl_qn, rg.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return OIT_ERR_DEVICE;
    for (uint32_t v : rg) tails += v;
  }
  std::memset(out, 0, sizeof *out);
  out->fragments = c->frame_frags;
  out->ll_overflow = h[1];
  out->tail_fragments = tails;
  out->appends = c->appends;
  out->device_errors = err;
  if (err != 0u) {
    // reported once: the word starts over for later frames, and the next
    // clear rebuilds the pool (walks that gave up left links behind)
    if (oit::launch_fill_u32(dev_err(c), 1, 0u, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      return OIT_ERR_DEVICE;
    c->ll_clean = false;
    c->ll_dirty_upto = c->ll_alloc;
    c->counters_clean = false;
    c->l64_clean = false;
    c->l32_clean = false;
    c->wb_clean = false;
    c->kb_clean = false;
    c->tl_clean = false;
    return OIT_ERR_DEVICE;
  }
  return OIT_OK;
}

int oit_debug_set_word(oit_ctx* c, uint32_t what, uint64_t index, uint32_t value, void* stream) {
  if (!c) return OIT_ERR_INVALID_ARG;
  if (!c->configured) return OIT_ERR_STATE;
  uint32_t* base = nullptr;
  uint64_t n = 0;
  switch (what) {
    case OIT_DEBUG_LL_HEAD: base = c->ll_head; n = c->ll_head ? c->cap_ll_head : 0; break;
    case OIT_DEBUG_LL_NEXT: base = c->ll_next; n = c->ll_next ? c->ll_alloc : 0; break;
    case OIT_DEBUG_SPIN_LOCK: base = c->kb_lock; n = c->kb_lock ? c->cap_kb_lock : 0; break;
    default: return OIT_ERR_INVALID_ARG;
  }
  if (!base || index >= n) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const hipError_t e = oit::launch_fill_u32(base + index, 1, value, s);
  mark_work(c, s);
  // a planted word is not the clear value a later clear may assume
  c->ll_clean = false;
  c->ll_dirty_upto = c->ll_alloc;
  c->kb_clean = false;
  return status_of(e);
}

int oit_codec_tables(float* thr257, float* dec256, float* alpha256) {
  if (!thr257 || !dec256 || !alpha256) return OIT_ERR_INVALID_ARG;
  oit::CodecTables t;
  oit::build_codec_tables(&t);
  std::memcpy(thr257, t.thr, sizeof t.thr);
  std::memcpy(dec256, t.dec, sizeof t.dec);
  std::memcpy(alpha256, t.alpha, sizeof t.alpha);
  return OIT_OK;
}

int oit_selftest_codec(oit_ctx* c, uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches, uint32_t* first_bad) {
  if (!c || !mismatches || !first_bad || hi_bits < lo_bits) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(c->device);
  struct {
    unsigned long long bad;
    uint32_t first;
    uint32_t pad;
  } h{0ull, 0xFFFFFFFFu, 0u};
  void* d = nullptr;
  if (hipMalloc(&d, sizeof h) != hipSuccess) return OIT_ERR_OOM;
  int rc = OIT_OK;
  if (hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) rc = OIT_ERR_DEVICE;
  if (!rc && hi_bits > lo_bits) {
    auto* dp = static_cast<unsigned long long*>(d);
    if (oit::launch_codec_selftest(c->d_tab, lo_bits, hi_bits, dp, reinterpret_cast<uint32_t*>(dp + 1), nullptr) !=
        hipSuccess)
      rc = OIT_ERR_DEVICE;
  }
  if (!rc && hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) rc = OIT_ERR_DEVICE;
  (void)hipFree(d);
  *mismatches = h.bad;
  *first_bad = h.first;
  return rc;
}

int oit_profile_enable(oit_ctx* c, int enable) {
  if (!c) return OIT_ERR_INVALID_ARG;
  c->profiling = enable != 0;
  return OIT_OK;
}

int oit_profile_read(oit_ctx* c, oit_profile* out) {
  if (!c || !out) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(c->device);
  for (auto& p : c->events) {
    if (hipEventSynchronize(p.b) != hipSuccess) return OIT_ERR_DEVICE;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->prof_ms[p.stage] += ms;
      c->prof_launches[p.stage] += 1;
    }
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->events.clear();
  std::memset(out, 0, sizeof *out);
  for (int i = 0; i < kStCount && i < OIT_PROFILE_STAGES; ++i) {
    out->ms[i] = c->prof_ms[i];
    out->launches[i] = c->prof_launches[i];
  }
  return OIT_OK;
}

int oit_profile_reset(oit_ctx* c) {
  if (!c) return OIT_ERR_INVALID_ARG;
  oit_profile tmp;
  int rc = oit_profile_read(c, &tmp);
  for (int i = 0; i < kStCount; ++i) {
    c->prof_ms[i] = 0;
    c->prof_launches[i] = 0;
  }
  return rc;
}

}  // extern "C"
