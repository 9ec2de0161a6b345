// oit_scene_gen.hpp - the device scene generator as a reusable engine: its
// device and pinned buffers grow to the largest scene seen and are reused,
// and the fragment count is read back asynchronously, so a host loop can
// stage one scene per request without blocking the device.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "oit_scene.hpp"

namespace oit {

struct SpanArgs {
  const SphereDesc* spheres;
  const uint32_t* sphere_span;  // [nspheres] first span of each sphere, ascending
  uint32_t nspheres;
  uint32_t spans;
  uint32_t width, height;
  float fw, fh;
  ScenePositions pos;
};

class SceneGen {
 public:
  explicit SceneGen(int device);
  ~SceneGen();
  SceneGen(const SceneGen&) = delete;
  SceneGen& operator=(const SceneGen&) = delete;

  // Builds the sphere table of p, stages it and enqueues the count on s.
  int begin(const oit_scene_params* p, hipStream_t s);
  // True once the count of the last begin() is known (non-blocking).
  bool count_ready();
  int wait_count();
  uint64_t fragments() const { return fragments_; }
  int device() const { return device_; }
  // Writes the stream (needs the count) into device arrays on s.
  int emit(uint32_t* target, float* depth, float* rgba, uint64_t capacity, hipStream_t s) const;

 private:
  int device_;
  SpanArgs args_{};
  uint32_t samples_ = 1;
  uint64_t fragments_ = 0;
  bool counting_ = false;
  bool staged_pending_ = false;
  std::vector<SphereDesc> spheres_;
  SphereDesc* d_spheres_ = nullptr;
  uint32_t* d_sphere_span_ = nullptr;
  uint32_t* d_span_frags_ = nullptr;
  uint64_t* d_span_base_ = nullptr;  // spans + 1 entries; the last is the total
  uint64_t* d_tiles_ = nullptr;
  uint64_t cap_spheres_ = 0, cap_sphere_span_ = 0, cap_spans_ = 0, cap_span_base_ = 0, cap_tiles_ = 0;
  SphereDesc* h_spheres_ = nullptr;
  uint32_t* h_sphere_span_ = nullptr;
  uint64_t* h_total_ = nullptr;
  uint64_t cap_h_spheres_ = 0, cap_h_sphere_span_ = 0, cap_h_total_ = 0;
  hipEvent_t ev_staged_ = nullptr;
  hipEvent_t ev_counted_ = nullptr;
};

}  // namespace oit
