// oit_scene_gpu.hip - the synthetic sphere scene generated on the device.
//
// Produces exactly the stream of oit_scene_generate_host (same fragments,
// same order, same bits): the host builds the sphere table with the shared
// build_spheres(), and the device walks it. Work unit = one bounding-box row
// of one sphere (a "span"), handled by one wave:
//   1. k_span_walk<count>  : covered samples per span -> span_frags[]
//   2. exclusive scan of span_frags -> span_base[] (u64), total in span_base[n]
//   3. k_span_walk<emit>   : each span writes its fragments at span_base[]
// Within a span the order is column, then sample, then front/back, which the
// wave reproduces with one ballot per sample position: a lane's first slot is
// the number of covered samples of lower lanes, its samples follow in order.
//
// Replaces the reference's scene set-up and rasterisation of the transparent
// spheres (main.cpp:334-417, object.vert.glsl:32-39, shaderCommon.glsl:48-56)
// as the producer of the fragment stream that the OIT append consumes.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "oit.h"
#include "oit_device.hpp"
#include "oit_scene_gen.hpp"

namespace oit {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kMaxSpans = 0x7FFFFFFFu;
constexpr uint32_t kMaxSpheres = 0x7FFFFFFFu;


// One wave per span. kEmit=false writes the span's fragment count; kEmit=true
// writes its fragments at span_base[span].
template <int S, bool kEmit>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_span_walk(SpanArgs a, uint32_t* __restrict__ span_frags,
                                                                   const uint64_t* __restrict__ span_base,
                                                                   uint32_t* __restrict__ target,
                                                                   float* __restrict__ depth,
                                                                   float* __restrict__ rgba) {
  const uint32_t span = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + threadIdx.x / 64);
  if (span >= a.spans) return;
  const uint32_t lane = lane_id();
  // the sphere of this span: the last one whose first span is <= span
  uint32_t lo = 0, hi = a.nspheres;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.sphere_span[mid] <= span) lo = mid;
    else hi = mid;
  }
  const uint32_t sid = __builtin_amdgcn_readfirstlane(lo);
  const SphereDesc d = a.spheres[sid];
  const uint32_t y = (uint32_t)d.y0 + (span - a.sphere_span[sid]);
  const float fy = (float)y;
  const uint32_t plane = a.width * a.height;
  uint64_t first = 0;  // fragment offset of the span's stream
  if constexpr (kEmit) first = span_base[span];
  uint64_t base = 0;  // covered samples of the span's earlier columns
  float4 colour = make_float4(d.rgba[0], d.rgba[1], d.rgba[2], d.rgba[3]);
  for (int32_t xb = d.x0; xb <= d.x1; xb += 64) {
    const int32_t x = xb + (int32_t)lane;
    const bool in = x <= d.x1;
    const float fx = (float)x;
    float h[S];
    uint32_t mine = 0;    // this lane's covered samples so far
    uint32_t before = 0;  // covered samples of lower lanes
    uint32_t total = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float px = (fx + a.pos.xy[s][0]) / a.fw;
      const float py = (fy + a.pos.xy[s][1]) / a.fh;
      const float dx = px - d.cx;
      const float dy = py - d.cy;
      const float d2 = dx * dx + dy * dy;
      const bool cov = in && d2 < d.r2;
      h[s] = cov ? __builtin_sqrtf(d.r2 - d2) : -1.0f;
      const uint64_t m = __ballot(cov);
      before += lanes_below(m);
      total += (uint32_t)__builtin_popcountll(m);
    }
    if constexpr (kEmit) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (!(h[s] >= 0.0f)) continue;
        const uint64_t f = first + 2 * (base + before + mine);
        ++mine;
        const uint32_t t = (uint32_t)s * plane + y * a.width + (uint32_t)x;
        *reinterpret_cast<uint2*>(target + f) = make_uint2(t, t);
        *reinterpret_cast<float2*>(depth + f) = make_float2(d.cz - h[s], d.cz + h[s]);
        reinterpret_cast<float4*>(rgba)[f] = colour;
        reinterpret_cast<float4*>(rgba)[f + 1] = colour;
      }
    }
    base += total;
  }
  if constexpr (!kEmit) {
    if (lane == 0) span_frags[span] = (uint32_t)(2 * base);
  }
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t u = __shfl_up(v, o, 64);
    if (lane >= (uint32_t)o) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ uint64_t block_excl_scan(uint64_t v, uint64_t& excl) {
  __shared__ uint64_t wsum[kScanThreads / 64];
  const uint64_t incl = wave_incl_scan(v);
  const uint32_t w = threadIdx.x / 64;
  if (lane_id() == 63) wsum[w] = incl;
  __syncthreads();
  uint64_t off = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / 64; ++i) {
    if ((uint32_t)i < w) off += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  excl = off + incl - v;
  return all;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint32_t* __restrict__ v, uint32_t n,
                                                             uint64_t* __restrict__ tile_sum) {
  const uint64_t i0 = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (i0 + k < n) s += v[i0 + k];
  uint64_t excl;
  const uint64_t all = block_excl_scan(s, excl);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = all;
}

// One block: exclusive scan of the tile sums in place; the grand total goes
// to *total.
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(uint64_t* __restrict__ tile_sum, uint32_t tiles,
                                                                 uint64_t* __restrict__ total) {
  uint64_t carry = 0;
  for (uint32_t b = 0; b < tiles; b += kScanThreads) {
    const uint32_t i = b + threadIdx.x;
    const uint64_t v = i < tiles ? tile_sum[i] : 0;
    uint64_t excl;
    const uint64_t all = block_excl_scan(v, excl);
    if (i < tiles) tile_sum[i] = carry + excl;
    carry += all;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* __restrict__ v, uint32_t n,
                                                             const uint64_t* __restrict__ tile_off,
                                                             uint64_t* __restrict__ out) {
  const uint64_t i0 = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t x[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    x[k] = i0 + k < n ? v[i0 + k] : 0;
    s += x[k];
  }
  uint64_t excl;
  (void)block_excl_scan(s, excl);
  uint64_t run = tile_off[blockIdx.x] + excl;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (i0 + k < n) out[i0 + k] = run;
    run += x[k];
  }
}

template <bool kEmit>
hipError_t launch_walk(const SpanArgs& a, uint32_t S, uint32_t* span_frags, const uint64_t* span_base,
                       uint32_t* target, float* depth, float* rgba, hipStream_t s) {
  const dim3 grid((a.spans + kWavesPerBlock - 1) / kWavesPerBlock), block(64 * kWavesPerBlock);
  switch (S) {
    case 1: k_span_walk<1, kEmit><<<grid, block, 0, s>>>(a, span_frags, span_base, target, depth, rgba); break;
    case 2: k_span_walk<2, kEmit><<<grid, block, 0, s>>>(a, span_frags, span_base, target, depth, rgba); break;
    case 4: k_span_walk<4, kEmit><<<grid, block, 0, s>>>(a, span_frags, span_base, target, depth, rgba); break;
    case 8: k_span_walk<8, kEmit><<<grid, block, 0, s>>>(a, span_frags, span_base, target, depth, rgba); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int st(hipError_t e) {
  if (e == hipSuccess) return OIT_OK;
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory) return OIT_ERR_OOM;
  if (e == hipErrorInvalidValue) return OIT_ERR_INVALID_ARG;
  return OIT_ERR_DEVICE;
}

template <typename T>
int grow_device(T** p, uint64_t* cap, uint64_t need) {
  if (need <= *cap && *p) return OIT_OK;
  if (*p) {  // earlier scenes' launches may still read it
    (void)hipDeviceSynchronize();
    (void)hipFree(*p);
  }
  *p = nullptr;
  *cap = 0;
  const int rc = st(hipMalloc((void**)p, need * sizeof(T)));
  if (!rc) *cap = need;
  return rc;
}

template <typename T>
int grow_host(T** p, uint64_t* cap, uint64_t need) {
  if (need <= *cap && *p) return OIT_OK;
  if (*p) {
    (void)hipDeviceSynchronize();
    (void)hipHostFree(*p);
  }
  *p = nullptr;
  *cap = 0;
  const int rc = st(hipHostMalloc((void**)p, need * sizeof(T), hipHostMallocDefault));
  if (!rc) *cap = need;
  return rc;
}

}  // namespace

SceneGen::SceneGen(int device) : device_(device) {}

SceneGen::~SceneGen() {
  (void)hipFree(d_spheres_);
  (void)hipFree(d_sphere_span_);
  (void)hipFree(d_span_frags_);
  (void)hipFree(d_span_base_);
  (void)hipFree(d_tiles_);
  if (h_spheres_) (void)hipHostFree(h_spheres_);
  if (h_sphere_span_) (void)hipHostFree(h_sphere_span_);
  if (h_total_) (void)hipHostFree(h_total_);
  if (ev_staged_) (void)hipEventDestroy(ev_staged_);
  if (ev_counted_) (void)hipEventDestroy(ev_counted_);
}

int SceneGen::begin(const oit_scene_params* p, hipStream_t s) {
  if (!scene_params_valid(p)) return OIT_ERR_INVALID_ARG;
  int rc;
  if (!ev_staged_ && (rc = st(hipEventCreateWithFlags(&ev_staged_, hipEventDisableTiming)))) return rc;
  if (!ev_counted_ && (rc = st(hipEventCreateWithFlags(&ev_counted_, hipEventDisableTiming)))) return rc;
  if (staged_pending_) {  // the previous scene's copies have read the staging buffers
    if ((rc = st(hipEventSynchronize(ev_staged_)))) return rc;
    staged_pending_ = false;
  }
  spheres_.clear();
  build_spheres(p, spheres_);
  if (spheres_.size() > kMaxSpheres) return OIT_ERR_INVALID_ARG;
  const uint64_t ns = spheres_.size();
  if ((rc = grow_host(&h_spheres_, &cap_h_spheres_, ns ? ns : 1))) return rc;
  if ((rc = grow_host(&h_sphere_span_, &cap_h_sphere_span_, ns ? ns : 1))) return rc;
  if ((rc = grow_host(&h_total_, &cap_h_total_, 1))) return rc;
  uint64_t spans = 0;
  for (uint64_t i = 0; i < ns; ++i) {
    h_sphere_span_[i] = (uint32_t)spans;
    h_spheres_[i] = spheres_[i];
    const int64_t rows = (int64_t)spheres_[i].y1 - spheres_[i].y0 + 1;
    if (rows > 0) spans += (uint64_t)rows;
    if (spans > kMaxSpans) return OIT_ERR_INVALID_ARG;
  }
  samples_ = p->samples;
  SpanArgs& a = args_;
  a.nspheres = (uint32_t)ns;
  a.spans = (uint32_t)spans;
  a.width = p->width;
  a.height = p->height;
  a.fw = (float)p->width;
  a.fh = (float)p->height;
  a.pos = scene_positions(p->samples);
  fragments_ = 0;
  counting_ = false;
  if (spans == 0) return OIT_OK;
  const uint64_t tiles = (spans + kScanTile - 1) / kScanTile;
  if ((rc = grow_device(&d_spheres_, &cap_spheres_, ns))) return rc;
  if ((rc = grow_device(&d_sphere_span_, &cap_sphere_span_, ns))) return rc;
  if ((rc = grow_device(&d_span_frags_, &cap_spans_, spans))) return rc;
  if ((rc = grow_device(&d_span_base_, &cap_span_base_, spans + 1))) return rc;
  if ((rc = grow_device(&d_tiles_, &cap_tiles_, tiles))) return rc;
  if ((rc = st(hipMemcpyAsync(d_spheres_, h_spheres_, ns * sizeof(SphereDesc), hipMemcpyHostToDevice, s)))) return rc;
  if ((rc = st(hipMemcpyAsync(d_sphere_span_, h_sphere_span_, ns * sizeof(uint32_t), hipMemcpyHostToDevice, s))))
    return rc;
  if ((rc = st(hipEventRecord(ev_staged_, s)))) return rc;
  staged_pending_ = true;
  a.spheres = d_spheres_;
  a.sphere_span = d_sphere_span_;
  if ((rc = st(launch_walk<false>(a, samples_, d_span_frags_, nullptr, nullptr, nullptr, nullptr, s)))) return rc;
  k_scan_tiles<<<(uint32_t)tiles, kScanThreads, 0, s>>>(d_span_frags_, a.spans, d_tiles_);
  k_scan_tile_sums<<<1, kScanThreads, 0, s>>>(d_tiles_, (uint32_t)tiles, d_span_base_ + spans);
  k_scan_apply<<<(uint32_t)tiles, kScanThreads, 0, s>>>(d_span_frags_, a.spans, d_tiles_, d_span_base_);
  if ((rc = st(hipGetLastError()))) return rc;
  if ((rc = st(hipMemcpyAsync(h_total_, d_span_base_ + spans, sizeof(uint64_t), hipMemcpyDeviceToHost, s))))
    return rc;
  if ((rc = st(hipEventRecord(ev_counted_, s)))) return rc;
  counting_ = true;
  return OIT_OK;
}

bool SceneGen::count_ready() {
  if (!counting_) return true;
  const hipError_t e = hipEventQuery(ev_counted_);
  if (e == hipErrorNotReady) return false;
  (void)hipGetLastError();
  fragments_ = *h_total_;
  counting_ = false;
  return true;
}

int SceneGen::wait_count() {
  if (!counting_) return OIT_OK;
  const int rc = st(hipEventSynchronize(ev_counted_));
  if (rc) return rc;
  fragments_ = *h_total_;
  counting_ = false;
  return OIT_OK;
}

int SceneGen::emit(uint32_t* target, float* depth, float* rgba, uint64_t capacity, hipStream_t s) const {
  if (counting_) return OIT_ERR_STATE;
  if (fragments_ == 0) return OIT_OK;
  if (!target || !depth || !rgba) return OIT_ERR_INVALID_ARG;
  if (((uintptr_t)target & 7) || ((uintptr_t)depth & 7) || ((uintptr_t)rgba & 15)) return OIT_ERR_INVALID_ARG;
  if (capacity < fragments_) return OIT_ERR_CAPACITY;
  return st(launch_walk<true>(args_, samples_, nullptr, d_span_base_, target, depth, rgba, s));
}

}  // namespace oit

struct oit_scene_gen {
  explicit oit_scene_gen(int dev) : device(dev), gen(dev) {}
  int device;
  oit::SceneGen gen;
};

extern "C" {

int oit_scene_gen_create(oit_scene_gen** out, const oit_scene_params* p, int device) {
  if (!out) return OIT_ERR_INVALID_ARG;
  *out = nullptr;
  if (!oit::scene_params_valid(p)) return OIT_ERR_INVALID_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return OIT_ERR_DEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) return OIT_ERR_DEVICE;
  oit_scene_gen* g = new (std::nothrow) oit_scene_gen(device);
  if (!g) return OIT_ERR_OOM;
  int rc = g->gen.begin(p, nullptr);
  if (!rc) rc = g->gen.wait_count();
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return OIT_OK;
}

int oit_scene_gen_destroy(oit_scene_gen* g) {
  if (!g) return OIT_ERR_INVALID_ARG;
  if (hipSetDevice(g->device) != hipSuccess) return OIT_ERR_DEVICE;
  (void)hipDeviceSynchronize();
  delete g;
  return OIT_OK;
}

int oit_scene_gen_count(const oit_scene_gen* g, uint64_t* out_count) {
  if (!g || !out_count) return OIT_ERR_INVALID_ARG;
  *out_count = g->gen.fragments();
  return OIT_OK;
}

int oit_scene_gen_emit(oit_scene_gen* g, uint32_t* target, float* depth, float* rgba, uint64_t capacity,
                       void* stream) {
  if (!g) return OIT_ERR_INVALID_ARG;
  if (hipSetDevice(g->device) != hipSuccess) return OIT_ERR_DEVICE;
  return g->gen.emit(target, depth, rgba, capacity, static_cast<hipStream_t>(stream));
}

}  // extern "C"
