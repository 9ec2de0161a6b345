// oit_scene.cpp - synthetic fragment stream: analytic spheres under an
// orthographic projection. Stands in for the reference's rasterised sphere
// mesh (Sample::initScene, main.cpp:334-417, random centres/radii/colours
// from a fixed seed) and its per-fragment shading output.
//
// Definition (bit-for-bit, built with -ffp-contract=off):
//  - PRNG: splitmix64 from `seed`; a uniform is (next() >> 40) * 2^-24 in
//    binary32, and a value in [lo, hi) is lo + u * (hi - lo).
//  - Draw order of uniforms: cluster origin ox, oy (each in [0, 1 - extent));
//    then per sphere: cx, cy, cz, radius, R, G, B, alpha.
//    The first floor(spheres * cluster_fraction) spheres have
//    cx = ox + cx * extent, cy = oy + cy * extent.
//  - Sample s of pixel (x, y) sits at ((x + sx) / W, (y + sy) / H) with the
//    standard sample positions (1x: centre; 2x/4x/8x: Vulkan standard).
//    d2 = dx*dx + dy*dy; covered iff d2 < r*r; h = sqrt(r*r - d2);
//    front depth cz - h, back depth cz + h.
//  - Emission order: sphere, row y, column x, sample, front then back.
//    target = s * W * H + y * W + x; colour premultiplied (R*a, G*a, B*a, a).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oit.h"
#include "oit_scene.hpp"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  float uniform() { return (float)(next() >> 40) * 0x1.0p-24f; }
  float range(float lo, float hi) { return lo + uniform() * (hi - lo); }
};

const float kPos1[1][2] = {{0.5f, 0.5f}};
const float kPos2[2][2] = {{0.75f, 0.75f}, {0.25f, 0.25f}};
const float kPos4[4][2] = {{0.375f, 0.125f}, {0.875f, 0.375f}, {0.125f, 0.625f}, {0.625f, 0.875f}};
const float kPos8[8][2] = {{0.5625f, 0.3125f}, {0.4375f, 0.6875f}, {0.8125f, 0.5625f}, {0.3125f, 0.1875f},
                           {0.1875f, 0.8125f}, {0.0625f, 0.4375f}, {0.6875f, 0.9375f}, {0.9375f, 0.0625f}};

const float (*sample_positions(uint32_t S))[2] {
  switch (S) {
    case 1: return kPos1;
    case 2: return kPos2;
    case 4: return kPos4;
    case 8: return kPos8;
    default: return nullptr;
  }
}

}  // namespace

namespace oit {

bool scene_params_valid(const oit_scene_params* p) {
  if (!p || p->width == 0 || p->height == 0) return false;
  if (!sample_positions(p->samples)) return false;
  if ((uint64_t)p->width * p->height * p->samples >= 0xFFFFFFFFull) return false;
  if (!(p->radius_min >= 0.f) || !(p->radius_max >= p->radius_min) || !(p->radius_max <= 1e6f)) return false;
  if (!(p->alpha_min >= 0.f) || !(p->alpha_max >= p->alpha_min) || p->alpha_max > 1.f) return false;
  if (!(p->cluster_fraction >= 0.f) || p->cluster_fraction > 1.f) return false;
  if (!(p->cluster_extent >= 0.f) || p->cluster_extent > 1.f) return false;
  return true;
}

ScenePositions scene_positions(uint32_t samples) {
  ScenePositions sp{};
  sp.samples = samples;
  const float(*pos)[2] = sample_positions(samples);
  for (uint32_t s = 0; pos && s < samples; ++s) {
    sp.xy[s][0] = pos[s][0];
    sp.xy[s][1] = pos[s][1];
  }
  return sp;
}

void build_spheres(const oit_scene_params* p, std::vector<SphereDesc>& out) {
  out.clear();
  out.reserve(p->spheres);
  SplitMix64 rng{p->seed};
  const float ext = p->cluster_extent;
  const float ox = rng.range(0.0f, 1.0f - ext);
  const float oy = rng.range(0.0f, 1.0f - ext);
  const uint32_t ncl = (uint32_t)std::floor((double)p->spheres * (double)p->cluster_fraction);
  const uint32_t W = p->width, H = p->height;
  for (uint32_t i = 0; i < p->spheres; ++i) {
    float cx = rng.uniform();
    float cy = rng.uniform();
    const float cz = rng.uniform();
    const float r = rng.range(p->radius_min, p->radius_max);
    const float R = rng.uniform();
    const float G = rng.uniform();
    const float B = rng.uniform();
    const float a = rng.range(p->alpha_min, p->alpha_max);
    if (i < ncl) {
      cx = ox + cx * ext;
      cy = oy + cy * ext;
    }
    SphereDesc d;
    d.cx = cx;
    d.cy = cy;
    d.cz = cz;
    d.r2 = r * r;
    d.rgba[0] = R * a;
    d.rgba[1] = G * a;
    d.rgba[2] = B * a;
    d.rgba[3] = a;
    // conservative pixel bounds, exact coverage test in the walkers
    long x0 = (long)std::floor(((double)cx - r) * W) - 1, x1 = (long)std::ceil(((double)cx + r) * W) + 1;
    long y0 = (long)std::floor(((double)cy - r) * H) - 1, y1 = (long)std::ceil(((double)cy + r) * H) + 1;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > (long)W - 1) x1 = (long)W - 1;
    if (y1 > (long)H - 1) y1 = (long)H - 1;
    d.x0 = (int32_t)x0;
    d.x1 = (int32_t)x1;
    d.y0 = (int32_t)y0;
    d.y1 = (int32_t)y1;
    out.push_back(d);
  }
}

}  // namespace oit

namespace {

// Walks the scene; calls emit(target, depth, rgba) in draw order.
template <typename Emit>
void walk(const oit_scene_params* p, Emit&& emit) {
  std::vector<oit::SphereDesc> spheres;
  oit::build_spheres(p, spheres);
  const uint32_t W = p->width, H = p->height, S = p->samples;
  const float(*pos)[2] = sample_positions(S);
  const float fW = (float)W, fH = (float)H;
  for (const oit::SphereDesc& d : spheres) {
    for (long y = d.y0; y <= d.y1; ++y) {
      for (long x = d.x0; x <= d.x1; ++x) {
        for (uint32_t s = 0; s < S; ++s) {
          const float px = ((float)x + pos[s][0]) / fW;
          const float py = ((float)y + pos[s][1]) / fH;
          const float dx = px - d.cx;
          const float dy = py - d.cy;
          const float d2 = dx * dx + dy * dy;
          if (!(d2 < d.r2)) continue;
          const float h = std::sqrt(d.r2 - d2);
          const uint32_t t = s * W * H + (uint32_t)y * W + (uint32_t)x;
          emit(t, d.cz - h, d.rgba);
          emit(t, d.cz + h, d.rgba);
        }
      }
    }
  }
}

}  // namespace

extern "C" {

int oit_scene_count(const oit_scene_params* p, uint64_t* out_count) {
  if (!oit::scene_params_valid(p) || !out_count) return OIT_ERR_INVALID_ARG;
  uint64_t n = 0;
  walk(p, [&](uint32_t, float, const float*) { ++n; });
  *out_count = n;
  return OIT_OK;
}

int oit_scene_generate_host(const oit_scene_params* p, uint32_t* target, float* depth, float* rgba, uint64_t capacity,
                            uint64_t* out_count) {
  if (!oit::scene_params_valid(p) || !target || !depth || !rgba || !out_count) return OIT_ERR_INVALID_ARG;
  uint64_t n = 0;
  bool over = false;
  walk(p, [&](uint32_t t, float d, const float* c) {
    if (n < capacity) {
      target[n] = t;
      depth[n] = d;
      std::memcpy(&rgba[4 * n], c, 4 * sizeof(float));
    } else {
      over = true;
    }
    ++n;
  });
  *out_count = n;
  return over ? OIT_ERR_CAPACITY : OIT_OK;
}

}  // extern "C"
