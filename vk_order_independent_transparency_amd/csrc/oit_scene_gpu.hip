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
#include "oit_scene.hpp"

namespace oit {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kMaxSpans = 0x7FFFFFFFu;

struct SpanArgs {
  const SphereDesc* spheres;
  const uint32_t* span_sphere;  // span -> sphere index
  const uint32_t* sphere_span;  // sphere -> index of its first span (row y0)
  uint32_t spans;
  uint32_t width, height;
  float fw, fh;
  ScenePositions pos;
};

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
  const uint32_t sid = __builtin_amdgcn_readfirstlane(a.span_sphere[span]);
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

}  // namespace
}  // namespace oit

struct oit_scene_gen {
  int device = 0;
  uint32_t samples = 1;
  oit::SpanArgs args{};
  oit::SphereDesc* d_spheres = nullptr;
  uint32_t* d_span_sphere = nullptr;
  uint32_t* d_sphere_span = nullptr;
  uint32_t* d_span_frags = nullptr;
  uint64_t* d_span_base = nullptr;  // spans + 1 entries; the last is the total
  uint64_t* d_tiles = nullptr;
  uint64_t fragments = 0;
};

namespace {

void release(oit_scene_gen* g) {
  (void)hipFree(g->d_spheres);
  (void)hipFree(g->d_span_sphere);
  (void)hipFree(g->d_sphere_span);
  (void)hipFree(g->d_span_frags);
  (void)hipFree(g->d_span_base);
  (void)hipFree(g->d_tiles);
  delete g;
}

int build(oit_scene_gen* g, const oit_scene_params* p) {
  using namespace oit;
  std::vector<SphereDesc> spheres;
  build_spheres(p, spheres);
  std::vector<uint32_t> sphere_span(spheres.size()), span_sphere;
  uint64_t spans = 0;
  for (size_t i = 0; i < spheres.size(); ++i) {
    sphere_span[i] = (uint32_t)spans;
    const int64_t rows = (int64_t)spheres[i].y1 - spheres[i].y0 + 1;
    if (rows <= 0) continue;
    if (spans + (uint64_t)rows > kMaxSpans) return OIT_ERR_INVALID_ARG;
    span_sphere.insert(span_sphere.end(), (size_t)rows, (uint32_t)i);
    spans += (uint64_t)rows;
  }
  g->samples = p->samples;
  SpanArgs& a = g->args;
  a.spans = (uint32_t)spans;
  a.width = p->width;
  a.height = p->height;
  a.fw = (float)p->width;
  a.fh = (float)p->height;
  a.pos = scene_positions(p->samples);
  if (spans == 0) return OIT_OK;
  const uint32_t tiles = (uint32_t)((spans + kScanTile - 1) / kScanTile);
  int rc;
  if ((rc = st(hipMalloc((void**)&g->d_spheres, spheres.size() * sizeof(SphereDesc))))) return rc;
  if ((rc = st(hipMalloc((void**)&g->d_span_sphere, spans * sizeof(uint32_t))))) return rc;
  if ((rc = st(hipMalloc((void**)&g->d_sphere_span, spheres.size() * sizeof(uint32_t))))) return rc;
  if ((rc = st(hipMalloc((void**)&g->d_span_frags, spans * sizeof(uint32_t))))) return rc;
  if ((rc = st(hipMalloc((void**)&g->d_span_base, (spans + 1) * sizeof(uint64_t))))) return rc;
  if ((rc = st(hipMalloc((void**)&g->d_tiles, tiles * sizeof(uint64_t))))) return rc;
  if ((rc = st(hipMemcpy(g->d_spheres, spheres.data(), spheres.size() * sizeof(SphereDesc), hipMemcpyHostToDevice))))
    return rc;
  if ((rc = st(hipMemcpy(g->d_span_sphere, span_sphere.data(), spans * sizeof(uint32_t), hipMemcpyHostToDevice))))
    return rc;
  if ((rc = st(hipMemcpy(g->d_sphere_span, sphere_span.data(), sphere_span.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice))))
    return rc;
  a.spheres = g->d_spheres;
  a.span_sphere = g->d_span_sphere;
  a.sphere_span = g->d_sphere_span;
  hipStream_t s = nullptr;
  if ((rc = st(launch_walk<false>(a, g->samples, g->d_span_frags, nullptr, nullptr, nullptr, nullptr, s)))) return rc;
  k_scan_tiles<<<tiles, kScanThreads, 0, s>>>(g->d_span_frags, a.spans, g->d_tiles);
  k_scan_tile_sums<<<1, kScanThreads, 0, s>>>(g->d_tiles, tiles, g->d_span_base + spans);
  k_scan_apply<<<tiles, kScanThreads, 0, s>>>(g->d_span_frags, a.spans, g->d_tiles, g->d_span_base);
  if ((rc = st(hipGetLastError()))) return rc;
  return st(hipMemcpy(&g->fragments, g->d_span_base + spans, sizeof(uint64_t), hipMemcpyDeviceToHost));
}

}  // namespace

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
  oit_scene_gen* g = new (std::nothrow) oit_scene_gen();
  if (!g) return OIT_ERR_OOM;
  g->device = device;
  const int rc = build(g, p);
  if (rc) {
    release(g);
    return rc;
  }
  *out = g;
  return OIT_OK;
}

int oit_scene_gen_destroy(oit_scene_gen* g) {
  if (!g) return OIT_ERR_INVALID_ARG;
  if (hipSetDevice(g->device) != hipSuccess) return OIT_ERR_DEVICE;
  release(g);
  return OIT_OK;
}

int oit_scene_gen_count(const oit_scene_gen* g, uint64_t* out_count) {
  if (!g || !out_count) return OIT_ERR_INVALID_ARG;
  *out_count = g->fragments;
  return OIT_OK;
}

int oit_scene_gen_emit(oit_scene_gen* g, uint32_t* target, float* depth, float* rgba, uint64_t capacity,
                       void* stream) {
  if (!g) return OIT_ERR_INVALID_ARG;
  if (g->fragments == 0) return OIT_OK;
  if (!target || !depth || !rgba) return OIT_ERR_INVALID_ARG;
  if (((uintptr_t)target & 7) || ((uintptr_t)depth & 7) || ((uintptr_t)rgba & 15)) return OIT_ERR_INVALID_ARG;
  if (capacity < g->fragments) return OIT_ERR_CAPACITY;
  if (hipSetDevice(g->device) != hipSuccess) return OIT_ERR_DEVICE;
  return oit::st(oit::launch_walk<true>(g->args, g->samples, nullptr, g->d_span_base, target, depth, rgba,
                                        static_cast<hipStream_t>(stream)));
}

}  // extern "C"
