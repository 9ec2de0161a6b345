// oit_scene.hpp - the synthetic sphere scene's per-sphere table, shared by the
// host generator and the device generator (same values, same draw order).
#pragma once

#include <stdint.h>

#include <vector>

#include "oit.h"

namespace oit {

struct SphereDesc {
  float cx, cy, cz, r2;
  float rgba[4];          // premultiplied
  int32_t x0, x1, y0, y1;  // conservative pixel bounds (inclusive)
};

struct ScenePositions {
  float xy[8][2];
  uint32_t samples;
};

bool scene_params_valid(const oit_scene_params* p);
ScenePositions scene_positions(uint32_t samples);
// Sphere table in draw order (splitmix64 draws, cluster placement, bounds).
void build_spheres(const oit_scene_params* p, std::vector<SphereDesc>& out);

}  // namespace oit
