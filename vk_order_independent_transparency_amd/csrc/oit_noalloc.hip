// oit_noalloc.hip - one Linked List frame that allocates nothing.
//
// The frame's structures (codec tables, frame counters, per-target heads,
// the node pool) live in `scratch`, a file-scope device array of
// OIT_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
// the process ends. Its device address is looked up once per device with
// hipGetSymbolAddress; host code uses only that address (the array's name in
// host code is a host-side stand-in, not a device address). The kernels are
// the single-frame Linked List's own (k_ll_append, k_ll_resolve_reset),
// given views into the array, after one fill kernel that writes the codec
// tables (passed by value as kernel arguments) and clears the heads, the
// links of the frame's nodes and the counters. No hipMalloc, hipFree,
// hipMemcpy or hipMemset is issued.
#include <hip/hip_runtime.h>

#include <mutex>

#include "oit.h"
#include "oit_kernels.hpp"

namespace {

static __device__ __align__(256) unsigned char scratch[OIT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices];  // device address of `scratch` per device, null until looked up
bool g_tables_built = false;
oit::CodecTables g_tables;  // host copy, built once per process

constexpr uint64_t kAlign = 256;
constexpr uint64_t round_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Byte offsets of the frame's structures in the array.
struct Layout {
  uint64_t tables, counters, head, cd, next, end;
};

Layout layout_of(uint64_t T, uint64_t nodes) {
  Layout l;
  l.tables = 0;
  l.counters = round_up(sizeof(oit::CodecTables));
  l.head = l.counters + round_up(32 * sizeof(uint32_t));
  l.cd = l.head + round_up(T * sizeof(uint32_t));
  l.next = l.cd + round_up(nodes * sizeof(uint64_t));
  l.end = l.next + round_up(nodes * sizeof(uint32_t));
  return l;
}

struct TablesArg {
  oit::CodecTables t;
};

// The codec tables from the kernel arguments; heads to 0 (empty), the links
// of nodes 0 .. nodes-1 to kUnset (the fused append's precondition), the 32
// frame counters to 0 (oitRender.cpp:206-212 clears head and counter).
__global__ __launch_bounds__(256) void k_noalloc_prepare(TablesArg tab, oit::CodecTables* dst, uint32_t* counters,
                                                         uint32_t* head, uint32_t T, uint32_t* next,
                                                         uint32_t nodes) {
  constexpr uint32_t kTab = (uint32_t)(sizeof(oit::CodecTables) / sizeof(float));
  const uint64_t total = (uint64_t)kTab + 32u + T + nodes;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    if (i < kTab) {
      reinterpret_cast<float*>(dst)[i] = reinterpret_cast<const float*>(&tab.t)[i];
    } else if (i < kTab + 32u) {
      counters[i - kTab] = 0u;
    } else if (i < kTab + 32u + T) {
      head[i - kTab - 32u] = 0u;
    } else {
      next[i - kTab - 32u - T] = oit::kUnset;
    }
  }
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

int status_of(hipError_t e) {
  if (e == hipSuccess) return OIT_OK;
  if (e == hipErrorInvalidValue) return OIT_ERR_INVALID_ARG;
  return OIT_ERR_DEVICE;
}

bool valid_layers(uint32_t k) { return k == 1 || k == 2 || k == 4 || k == 8 || k == 16 || k == 32; }

// Set-up, once per device with that device current: the array's device
// address (and, once per process, the host codec tables).
int scratch_of(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_tables_built) {
    oit::build_codec_tables(&g_tables);
    g_tables_built = true;
  }
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? status_of(e) : OIT_ERR_DEVICE;
    g_scratch[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[device];
  return OIT_OK;
}

}  // namespace

extern "C" int oit_ll_render_noalloc(int device, const oit_batch_frame* f, void* stream) {
  if (!f || !f->out) return OIT_ERR_INVALID_ARG;
  if (device < 0 || device >= kMaxDevices) return OIT_ERR_INVALID_ARG;
  if (f->width == 0 || f->height == 0 || f->samples == 0 || !valid_layers(f->layers) || f->reserved != 0)
    return OIT_ERR_INVALID_ARG;
  if (f->frags.coverage) return OIT_ERR_UNSUPPORTED;  // per-pixel fragments with sample masks: the context path
  const uint64_t T = (uint64_t)f->width * f->height * f->samples;
  const uint64_t F = f->frags.count;
  if (F > 0 && (!f->frags.target || !f->frags.depth || !f->frags.rgba)) return OIT_ERR_INVALID_ARG;
  if ((reinterpret_cast<uintptr_t>(f->out) & 15u) != 0) return OIT_ERR_INVALID_ARG;
  if (T > 0xFFFFFFFFull || F > 0x7FFFFFFEull) return OIT_ERR_CAPACITY;
  const uint64_t nodes = F + 1;  // node 0 is the terminator: the pool holds the frame, nothing overflows
  const Layout l = layout_of(T, nodes);
  if (l.end > OIT_NOALLOC_SCRATCH_BYTES) return OIT_ERR_CAPACITY;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device >= count) return OIT_ERR_INVALID_ARG;
  DeviceGuard g(device);
  unsigned char* base = nullptr;
  int rc = scratch_of(device, &base);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);

  auto* tab = reinterpret_cast<oit::CodecTables*>(base + l.tables);
  auto* counters = reinterpret_cast<uint32_t*>(base + l.counters);
  auto* head = reinterpret_cast<uint32_t*>(base + l.head);
  auto* cd = reinterpret_cast<uint64_t*>(base + l.cd);
  auto* next = reinterpret_cast<uint32_t*>(base + l.next);

  TablesArg ta;
  ta.t = g_tables;  // written once under the mutex before any device was set up; read-only since
  const uint64_t fill = sizeof(oit::CodecTables) / sizeof(float) + 32u + T + nodes;
  uint64_t blocks = (fill + 255u) / 256u;
  if (blocks > 2048u) blocks = 2048u;
  hipLaunchKernelGGL(k_noalloc_prepare, dim3((uint32_t)blocks), dim3(256), 0, s, ta, tab, counters, head,
                     (uint32_t)T, next, (uint32_t)nodes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return status_of(e);

  oit::Common k;
  k.T = (uint32_t)T;
  k.K = f->layers;
  k.tail = f->tail_blend ? 1 : 0;
  k.background = make_float4(f->background[0], f->background[1], f->background[2], f->background[3]);
  k.tab = tab;
  oit::LLBuf ll{head, cd, next, nodes, counters + 1, counters};
  const oit::TailBuf none{};
  if (F > 0) {
    oit::FragView v;
    v.target = f->frags.target;
    v.depth = f->frags.depth;
    v.rgba = reinterpret_cast<const float4*>(f->frags.rgba);
    v.count = (uint32_t)F;
    v.id_base = 0;
    e = oit::launch_ll_append(k, v, ll, none, 4, s);
    if (e != hipSuccess) return status_of(e);
  }
  oit::ResolveOut o{reinterpret_cast<float4*>(f->out), nullptr, nullptr, nullptr};
  e = oit::launch_ll_resolve_reset(k, ll, none, o, true, false, s);
  return status_of(e);
}
