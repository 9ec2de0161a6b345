// oit_server.cpp - request loop over the OIT path (oit_server_*).
//
// Requests (a seeded scene, an algorithm, K) arrive at any time from any
// thread. The server owns `slots` HIP streams, each driven by a thread of
// its own; each slot keeps a scene generator and its fragment-stream
// buffers, and the OIT contexts form one pool whose device buffers stay
// resident and are reused by later requests. A slot's thread takes the
// oldest waiting request and:
//   1. stages the scene and counts its fragments on the device,
//   2. sizes the A-buffer from that exact count (Linked List pool = F + 1
//      nodes, Loop64/k-buffer tail arrays = F) and asks the HBM budget for
//      it: granted in arrival order (a ticket per request), idle contexts
//      and idle slots' stream buffers are trimmed under pressure, a request
//      larger than the whole budget is refused,
//   3. generates the fragments into the slot's stream buffers, runs clear,
//      append and resolve, and waits for its completion event.
// Slots run concurrently (continuous batching: a request that arrives while
// others run is admitted as soon as a slot and memory are free, and its
// kernels overlap theirs on the device), and their HIP calls are issued
// from their own threads. Latency is arrival to the slot seeing the
// completion event, on the host steady clock.
//
// No counterpart in the reference (a single interactive frame loop,
// Sample::onRender, oitRender.cpp:28-154); the per-request work is the
// reference's clear/draw pair through the same C ABI as a single frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "oit.h"
#include "oit_scene_gen.hpp"

struct PoolCtx;

namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kMaxSlots = 16;

int st(hipError_t e) {
  if (e == hipSuccess) return OIT_OK;
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory) return OIT_ERR_OOM;
  if (e == hipErrorInvalidValue) return OIT_ERR_INVALID_ARG;
  return OIT_ERR_DEVICE;
}

bool valid_layers(uint32_t k) { return k == 1 || k == 2 || k == 4 || k == 8 || k == 16 || k == 32; }

bool supported(uint32_t alg) { return alg <= OIT_WEIGHTED; }

// Device bytes a context holds for one request (the sizes oit_configure
// allocates; tail arrays 28 B per fragment and a 4 B queue counter per 1024
// fragments, tail heads 4 B per target).
uint64_t request_bytes(const oit_request& r, uint64_t F) {
  const uint64_t T = (uint64_t)r.scene.width * r.scene.height * r.scene.samples;
  const uint64_t K = r.algorithm == OIT_WEIGHTED ? 1 : r.layers;
  const bool tail = r.tail_blend != 0;
  const uint64_t tails = tail ? 4 * T + 28 * F + 4 * ((F + 1023) / 1024) : 0;
  switch (r.algorithm) {
    case OIT_LINKEDLIST: return 4 * T + 12 * (F + 1) + (tail ? 4 * T : 0);
    case OIT_INTERLOCK: return 4 * T + 28 * (F + 1);
    case OIT_LOOP: return 12 * K * T + tails;
    case OIT_LOOP64: return 8 * K * T + 4 * T + 8 * ((F + 63) / 64) + tails;  // slots, counts, chunk masks
    case OIT_SIMPLE: return 4 * T + 12 * K * T + tails;
    case OIT_SPINLOCK: return 12 * T + 12 * K * T + tails;
    default: return 20 * T;
  }
}

double micros(Clock::time_point t, Clock::time_point t0) {
  return std::chrono::duration<double, std::micro>(t - t0).count();
}

// A slot's stream buffers grow in stream order on the slot's own stream
// (hipFreeAsync / hipMallocAsync): only that stream's earlier work can still
// read the old buffer, so nothing waits for the other slots (hipFree would
// synchronise the whole device). Growth takes a quarter more than asked.
template <typename T>
int grow(T** p, uint64_t* cap, uint64_t need, hipStream_t s) {
  if (need <= *cap && *p) return OIT_OK;
  if (*p) (void)hipFreeAsync(*p, s);
  *p = nullptr;
  *cap = 0;
  const uint64_t n = need ? need + need / 4 : 1;
  const int rc = st(hipMallocAsync((void**)p, n * sizeof(T), s));
  if (!rc) *cap = n;
  return rc;
}

// Device bytes of a request's fragment stream and own output in a slot.
uint64_t stream_bytes(uint64_t F, uint64_t T, bool own_out) { return 24 * F + (own_out ? 16 * T : 0); }

struct Pending {
  oit_request req;
  Clock::time_point t_submit;
  uint64_t replay = 0;  // oit_server_replay call that submitted it (0: oit_server_submit)
};

struct Done {
  oit_completion c;
  uint64_t replay;
};

struct Slot {
  explicit Slot(int device) : gen(device) {}
  uint32_t index = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // device intervals of the request: count (ev[0] .. ev[1]) and the
  // generation plus OIT passes (ev[2] .. done)
  hipEvent_t ev[3] = {};
  oit::SceneGen gen;
  std::thread thread;
  // under the server's grant lock: the stream buffers (another slot's grant
  // may release them while this slot is not launching) and `launching`
  uint32_t* target = nullptr;
  float* depth = nullptr;
  float4* rgba = nullptr;
  uint64_t cap_target = 0, cap_depth = 0, cap_rgba = 0;
  float4* out_own = nullptr;
  uint64_t cap_out = 0;
  bool launching = false;  // granted and not finished: its buffers and context are in use
  int ctx = -1;            // index of the pooled context the request uses (grant .. finish)
  PoolCtx* pctx = nullptr;  // that context (a deque element: its address is stable)
  // granted growth not yet allocated (the stream buffers', the context's):
  // counted as held until the launch has grown them, so that a later grant
  // does not hand the same bytes out again
  uint64_t reserved_stream = 0, reserved_ctx = 0;
  uint64_t held_stream() const { return 4 * (cap_target + cap_depth) + 16 * (cap_rgba + cap_out); }
  // the slot thread's own
  Pending job{};
  uint64_t seq = 0;  // admission ticket (memory is granted in this order)
  uint64_t fragments = 0;
  uint64_t budgeted = 0;     // A-buffer bytes of the request's context
  uint64_t stream_need = 0;  // fragment-stream (and own output) bytes of the request
  const float* out = nullptr;
  Clock::time_point t_start;
};

}  // namespace

// A context of the server's pool: any slot takes an idle one of the
// request's algorithm (contexts keep their buffers between requests).
struct PoolCtx {
  oit_ctx* ctx = nullptr;
  uint32_t alg = 0;
  uint64_t resident = 0;  // device bytes it holds
  bool busy = false;
  uint64_t last_use = 0;  // admission number of its last request
};

struct oit_server {
  int device = 0;
  oit_server_config cfg{};
  Clock::time_point t0;
  std::vector<std::unique_ptr<Slot>> slots;
  // grants: the pool, the slots' buffers, the admission ticket
  std::mutex gmu;
  std::condition_variable grant_cv;  // a ticket was used up, or memory was given back
  std::deque<PoolCtx> ctxs;          // stable addresses: a slot uses its entry outside the lock
  uint64_t next_grant = 0;
  // shared with submitters
  std::mutex mu;
  std::condition_variable cv;       // work arrived / stop
  std::condition_variable idle_cv;  // everything finished
  std::deque<Pending> queue;
  std::vector<Done> done;
  uint64_t replay_calls = 0;
  bool stop = false;
  uint64_t submitted = 0, completed = 0, failed = 0;
  uint64_t held = 0, peak_held = 0;
  uint32_t in_flight = 0, peak_in_flight = 0;
  uint64_t next_seq = 0;
};

namespace {

// Device bytes the server holds for requests: contexts and the slots'
// stream buffers. Under gmu.
uint64_t held_bytes(const oit_server* s) {
  uint64_t h = 0;
  for (const auto& sl : s->slots) h += sl->held_stream() + sl->reserved_stream + sl->reserved_ctx;
  for (const auto& c : s->ctxs) h += c.resident;
  return h;
}

// The context a request of `alg` needing `need` bytes takes: the smallest
// idle one that already holds enough, else the largest idle one (it grows),
// else a new entry (created at launch). Returns its index. Under gmu.
int pick_ctx(oit_server* s, uint32_t alg, uint64_t need) {
  int fit = -1, big = -1;
  for (int i = 0; i < (int)s->ctxs.size(); ++i) {
    const PoolCtx& c = s->ctxs[i];
    if (c.busy || c.alg != alg) continue;
    if (c.resident >= need && (fit < 0 || c.resident < s->ctxs[fit].resident)) fit = i;
    if (big < 0 || c.resident > s->ctxs[big].resident) big = i;
  }
  if (fit >= 0) return fit;
  if (big >= 0) return big;
  PoolCtx c;
  c.alg = alg;
  s->ctxs.push_back(c);
  return (int)s->ctxs.size() - 1;
}

// Under gmu.
void publish_held(oit_server* s) {
  const uint64_t h = held_bytes(s);
  std::lock_guard<std::mutex> lk(s->mu);
  s->held = h;
  s->peak_held = std::max(s->peak_held, h);
}

void finish(oit_server* s, Slot& sl, int status) {
  oit_completion c;
  std::memset(&c, 0, sizeof c);
  c.tag = sl.job.req.tag;
  c.status = status;
  c.slot = sl.index;
  c.fragments = sl.fragments;
  c.abuffer_bytes = sl.budgeted;
  c.out = status == OIT_OK ? sl.out : nullptr;
  const auto now = Clock::now();
  c.t_submit_us = micros(sl.job.t_submit, s->t0);
  c.t_start_us = micros(sl.t_start, s->t0);
  c.t_done_us = micros(now, s->t0);
  if (status == OIT_OK) {
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, sl.ev[0], sl.ev[1]) == hipSuccess &&
        hipEventElapsedTime(&b, sl.ev[2], sl.done) == hipSuccess)
      c.device_us = 1e3 * ((double)a + (double)b);
    else
      (void)hipGetLastError();
  }
  {
    std::lock_guard<std::mutex> g(s->gmu);
    if (sl.ctx >= 0) s->ctxs[sl.ctx].busy = false;
    sl.ctx = -1;
    sl.pctx = nullptr;
    sl.launching = false;
    sl.reserved_stream = sl.reserved_ctx = 0;
  }
  s->grant_cv.notify_all();  // its context and buffers may be trimmed for a waiting request
  std::lock_guard<std::mutex> lk(s->mu);
  s->done.push_back(Done{c, sl.job.replay});
  s->completed++;
  if (status != OIT_OK) s->failed++;
  s->in_flight--;
  if (s->in_flight == 0 && s->queue.empty()) s->idle_cv.notify_all();
}

// Frees a slot's fragment-stream and output buffers, in order on its stream.
// Under gmu, for a slot that is not launching (or at teardown).
void release_stream(Slot& sl) {
  for (void* p : {(void*)sl.target, (void*)sl.depth, (void*)sl.rgba, (void*)sl.out_own})
    if (p) (void)hipFreeAsync(p, sl.stream);
  sl.target = nullptr;
  sl.depth = nullptr;
  sl.rgba = nullptr;
  sl.out_own = nullptr;
  sl.cap_target = sl.cap_depth = sl.cap_rgba = sl.cap_out = 0;
}

// Frees idle contexts (least recently used first) and the stream buffers
// of slots that are not launching until `need` more bytes fit the budget;
// `keep` is the context the request itself will use. Under gmu.
bool make_room(oit_server* s, Slot& self, int keep, uint64_t need) {
  uint64_t held = held_bytes(s);
  if (held + need <= s->cfg.budget_bytes) return true;
  std::vector<int> idle;
  for (int i = 0; i < (int)s->ctxs.size(); ++i)
    if (i != keep && !s->ctxs[i].busy && s->ctxs[i].resident) idle.push_back(i);
  std::sort(idle.begin(), idle.end(), [&](int a, int b) { return s->ctxs[a].last_use < s->ctxs[b].last_use; });
  for (int i : idle) {
    if (held + need <= s->cfg.budget_bytes) break;
    if (oit_trim(s->ctxs[i].ctx) != OIT_OK) continue;
    held -= s->ctxs[i].resident;
    s->ctxs[i].resident = 0;
  }
  for (auto& sp : s->slots) {
    if (held + need <= s->cfg.budget_bytes) break;
    Slot& sl = *sp;
    if (&sl == &self || sl.launching || !sl.held_stream()) continue;
    held -= sl.held_stream();
    release_stream(sl);
  }
  return held + need <= s->cfg.budget_bytes;
}

// Enqueues generation and the OIT passes of the slot's granted request.
int launch(oit_server* s, Slot& sl) {
  const oit_request& r = sl.job.req;
  const uint32_t alg = r.algorithm;
  const uint64_t F = sl.fragments;
  const uint64_t T = (uint64_t)r.scene.width * r.scene.height * r.scene.samples;
  int rc;
  if ((rc = st(hipEventRecord(sl.ev[2], sl.stream)))) return rc;
  float* out = r.out;
  {
    std::lock_guard<std::mutex> g(s->gmu);  // others read the caps for the budget
    if ((rc = grow(&sl.target, &sl.cap_target, F, sl.stream))) return rc;
    if ((rc = grow(&sl.depth, &sl.cap_depth, F, sl.stream))) return rc;
    if ((rc = grow(&sl.rgba, &sl.cap_rgba, F, sl.stream))) return rc;
    if (!out) {
      if ((rc = grow(&sl.out_own, &sl.cap_out, T, sl.stream))) return rc;
      out = reinterpret_cast<float*>(sl.out_own);
    }
    sl.reserved_stream = 0;  // now in the caps
  }
  sl.out = out;
  PoolCtx& pc = *sl.pctx;  // this slot's while busy
  if (!pc.ctx && (rc = oit_create(&pc.ctx, s->device))) return rc;
  oit_config c;
  std::memset(&c, 0, sizeof c);
  c.algorithm = alg;
  c.layers = alg == OIT_WEIGHTED ? 1 : r.layers;
  c.tail_blend = r.tail_blend;
  c.width = r.scene.width;
  c.height = r.scene.height;
  c.samples = r.scene.samples;
  c.ll_pool_nodes = (alg == OIT_LINKEDLIST || alg == OIT_INTERLOCK) ? F + 1 : 0;
  c.max_fragments = (r.tail_blend && alg != OIT_LINKEDLIST && alg != OIT_WEIGHTED) ? (F ? F : 1) : 0;
  c.background[0] = c.background[1] = c.background[2] = 0.5f;
  c.background[3] = 1.0f;
  c.wboit_view_near = 2.0f;
  c.wboit_view_span = 8.0f;
  if ((rc = oit_configure(pc.ctx, &c))) return rc;
  oit_sizes z;
  const bool sized = oit_query_sizes(pc.ctx, &z) == OIT_OK;
  {
    std::lock_guard<std::mutex> g(s->gmu);
    if (sized) pc.resident = z.resident_bytes;
    sl.reserved_ctx = 0;  // now in the context's resident bytes
    publish_held(s);
  }
  if ((rc = sl.gen.emit(sl.target, sl.depth, reinterpret_cast<float*>(sl.rgba), sl.cap_target, sl.stream))) return rc;
  oit_fragments f{sl.target, sl.depth, reinterpret_cast<const float*>(sl.rgba), F, nullptr};
  if ((rc = oit_clear(pc.ctx, sl.stream))) return rc;
  if ((rc = oit_append(pc.ctx, &f, sl.stream))) return rc;
  if ((rc = oit_resolve(pc.ctx, out, sl.stream))) return rc;
  if (oit_query_sizes(pc.ctx, &z) == OIT_OK) {  // the resolve may size per-frame buffers (Loop64's chunk masks)
    std::lock_guard<std::mutex> g(s->gmu);
    pc.resident = z.resident_bytes;
    publish_held(s);
  }
  return st(hipEventRecord(sl.done, sl.stream));
}

// Waits for the request's turn at the grants and for memory, then marks
// the slot launching with its context. Returns OIT_OK when granted, else
// the request's error; the ticket is used up either way.
int grant(oit_server* s, Slot& sl, bool counted) {
  std::unique_lock<std::mutex> g(s->gmu);
  s->grant_cv.wait(g, [&] { return s->next_grant == sl.seq; });
  int rc = OIT_OK;
  if (!counted) {
    rc = OIT_ERR_DEVICE;
  } else if (sl.budgeted + sl.stream_need > s->cfg.budget_bytes) {
    rc = OIT_ERR_OVER_BUDGET;
  } else {
    const int alg = (int)sl.job.req.algorithm;
    for (;;) {
      // the context's growth plus the slot's stream buffers' growth (with
      // the quarter of slack grow() adds)
      const int ci = pick_ctx(s, (uint32_t)alg, sl.budgeted);
      const uint64_t held_c = s->ctxs[ci].resident;
      const uint64_t held_s = sl.held_stream();
      const uint64_t want_s = sl.stream_need + sl.stream_need / 4;
      const uint64_t grow_c = sl.budgeted > held_c ? sl.budgeted - held_c : 0;
      const uint64_t grow_s = sl.stream_need > held_s ? want_s - held_s : 0;
      if (make_room(s, sl, ci, grow_c + grow_s)) {
        sl.reserved_ctx = grow_c;
        sl.reserved_stream = grow_s;
        sl.ctx = ci;
        sl.pctx = &s->ctxs[ci];
        s->ctxs[ci].busy = true;
        s->ctxs[ci].last_use = sl.seq;
        sl.launching = true;
        break;
      }
      // a later ticket waits behind this one; running requests give memory back
      s->grant_cv.wait(g);
    }
  }
  s->next_grant++;
  g.unlock();
  s->grant_cv.notify_all();
  return rc;
}

// One request on a slot, from its count to its completion.
void run_request(oit_server* s, Slot& sl) {
  sl.fragments = 0;
  sl.budgeted = 0;
  sl.stream_need = 0;
  sl.out = nullptr;
  sl.t_start = Clock::now();
  int rc = st(hipEventRecord(sl.ev[0], sl.stream));
  if (!rc) rc = sl.gen.begin(&sl.job.req.scene, sl.stream);
  if (!rc) rc = st(hipEventRecord(sl.ev[1], sl.stream));
  if (!rc) {
    while (!sl.gen.count_ready()) std::this_thread::yield();
    sl.fragments = sl.gen.fragments();
    sl.budgeted = request_bytes(sl.job.req, sl.fragments);
    const auto& sc = sl.job.req.scene;
    sl.stream_need = stream_bytes(sl.fragments, (uint64_t)sc.width * sc.height * sc.samples, !sl.job.req.out);
  }
  const int g = grant(s, sl, rc == OIT_OK);
  if (g != OIT_OK) {
    finish(s, sl, rc ? rc : g);
    return;
  }
  sl.t_start = Clock::now();
  rc = launch(s, sl);
  if (rc) {
    finish(s, sl, rc);
    return;
  }
  // this thread's only job is this request: poll its completion back to
  // back (a timed sleep of a few microseconds lasts tens on a loaded host)
  for (;;) {
    const hipError_t e = hipEventQuery(sl.done);
    if (e == hipSuccess) {
      finish(s, sl, OIT_OK);
      return;
    }
    if (e != hipErrorNotReady) {
      (void)hipGetLastError();
      finish(s, sl, OIT_ERR_DEVICE);
      return;
    }
    std::this_thread::yield();
  }
}

void slot_main(oit_server* s, Slot* slp) {
  Slot& sl = *slp;
  (void)hipSetDevice(s->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->queue.empty()) return;  // stopping, and nothing left to run
      sl.job = s->queue.front();
      s->queue.pop_front();
      sl.seq = s->next_seq++;  // tickets in queue order
      s->in_flight++;
      s->peak_in_flight = std::max(s->peak_in_flight, s->in_flight);
    }
    run_request(s, sl);
  }
}

void release(oit_server* s) {
  for (auto& c : s->ctxs)
    if (c.ctx) (void)oit_destroy(c.ctx);
  s->ctxs.clear();
  for (auto& sp : s->slots) {
    Slot& sl = *sp;
    if (sl.stream) {
      release_stream(sl);
      (void)hipStreamSynchronize(sl.stream);
    }
    if (sl.done) (void)hipEventDestroy(sl.done);
    for (hipEvent_t e : sl.ev)
      if (e) (void)hipEventDestroy(e);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  s->slots.clear();
}

int check_request(const oit_request* r) {
  if (!r) return OIT_ERR_INVALID_ARG;
  if (!supported(r->algorithm)) return OIT_ERR_INVALID_ARG;
  if (r->algorithm != OIT_WEIGHTED && !valid_layers(r->layers)) return OIT_ERR_INVALID_ARG;
  if (r->tail_blend > 1u || r->reserved != 0u) return OIT_ERR_INVALID_ARG;
  if (!oit::scene_params_valid(&r->scene)) return OIT_ERR_INVALID_ARG;
  if (r->out && ((uintptr_t)r->out & 15u)) return OIT_ERR_INVALID_ARG;
  return OIT_OK;
}

int submit_tagged(oit_server* s, const oit_request& r, uint64_t replay) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return OIT_ERR_STATE;
    s->queue.push_back(Pending{r, Clock::now(), replay});
    s->submitted++;
  }
  s->cv.notify_one();
  return OIT_OK;
}

}  // namespace

extern "C" {

int oit_server_create(oit_server** out, int device, const oit_server_config* cfg) {
  if (!out || !cfg) return OIT_ERR_INVALID_ARG;
  *out = nullptr;
  if (cfg->slots == 0 || cfg->slots > kMaxSlots || cfg->budget_bytes == 0) return OIT_ERR_INVALID_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return OIT_ERR_DEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) return OIT_ERR_DEVICE;
  oit_server* s = new (std::nothrow) oit_server();
  if (!s) return OIT_ERR_OOM;
  s->device = device;
  s->cfg = *cfg;
  s->t0 = Clock::now();
  int rc = OIT_OK;
  for (uint32_t i = 0; i < cfg->slots && !rc; ++i) {
    auto sl = std::make_unique<Slot>(device);
    sl->index = i;
    rc = st(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
    if (!rc) rc = st(hipEventCreate(&sl->done));
    for (int k = 0; k < 3 && !rc; ++k) rc = st(hipEventCreate(&sl->ev[k]));
    s->slots.push_back(std::move(sl));
  }
  if (rc) {
    release(s);
    delete s;
    return rc;
  }
  for (auto& sl : s->slots) sl->thread = std::thread(slot_main, s, sl.get());
  *out = s;
  return OIT_OK;
}

int oit_server_destroy(oit_server* s) {
  if (!s) return OIT_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto& sl : s->slots)
    if (sl->thread.joinable()) sl->thread.join();
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  release(s);
  delete s;
  return OIT_OK;
}

int oit_server_submit(oit_server* s, const oit_request* r) {
  if (!s) return OIT_ERR_INVALID_ARG;
  const int rc = check_request(r);
  if (rc) return rc;
  return submit_tagged(s, *r, 0);
}

int oit_server_poll(oit_server* s, oit_completion* out, uint32_t max, uint32_t* n) {
  if (!s || !n || (max && !out)) return OIT_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  uint32_t k = 0;  // completions of oit_server_submit only; replay calls collect their own
  for (auto it = s->done.begin(); it != s->done.end() && k < max;) {
    if (it->replay == 0) {
      out[k++] = it->c;
      it = s->done.erase(it);
    } else {
      ++it;
    }
  }
  *n = k;
  return OIT_OK;
}

int oit_server_drain(oit_server* s, uint32_t timeout_ms) {
  if (!s) return OIT_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->idle_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                                      [&] { return s->in_flight == 0 && s->queue.empty(); });
  return ok ? OIT_OK : OIT_ERR_NOT_READY;
}

int oit_server_replay(oit_server* s, const oit_request* reqs, const double* arrival_us, uint32_t n,
                      uint32_t timeout_ms, oit_completion* out) {
  if (!s || (n && (!reqs || !arrival_us || !out))) return OIT_ERR_INVALID_ARG;
  for (uint32_t i = 0; i < n; ++i) {
    const int rc = check_request(&reqs[i]);
    if (rc) return rc;
    if (i && arrival_us[i] < arrival_us[i - 1]) return OIT_ERR_INVALID_ARG;
  }
  uint64_t call;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    call = ++s->replay_calls;
  }
  const auto start = Clock::now();
  for (uint32_t i = 0; i < n; ++i) {
    const auto due = start + std::chrono::duration_cast<Clock::duration>(
                                 std::chrono::duration<double, std::micro>(arrival_us[i]));
    for (;;) {
      const auto now = Clock::now();
      if (now >= due) break;
      if (due - now > std::chrono::microseconds(300))
        std::this_thread::sleep_for(due - now - std::chrono::microseconds(200));
    }
    oit_request r = reqs[i];
    r.tag = i;  // index within this call; the caller's tag is restored below
    const int rc = submit_tagged(s, r, call);
    if (rc) return rc;
  }
  int rc = oit_server_drain(s, timeout_ms);
  if (rc) return rc;
  std::vector<bool> seen(n, false);
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto it = s->done.begin(); it != s->done.end();) {
    const uint64_t i = it->c.tag;
    if (it->replay == call && i < n && !seen[i]) {
      out[i] = it->c;
      out[i].tag = reqs[i].tag;
      seen[i] = true;
      it = s->done.erase(it);
    } else {
      ++it;
    }
  }
  for (uint32_t i = 0; i < n; ++i)
    if (!seen[i]) return OIT_ERR_STATE;
  return OIT_OK;
}

int oit_server_get_stats(oit_server* s, oit_server_stats* out) {
  if (!s || !out) return OIT_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof *out);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->failed = s->failed;
  out->held_bytes = s->held;
  out->peak_held_bytes = s->peak_held;
  out->peak_in_flight = s->peak_in_flight;
  out->slots = s->cfg.slots;
  return OIT_OK;
}

}  // extern "C"
