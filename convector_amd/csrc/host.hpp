// host.hpp -- host-side helpers and object layouts shared by the C-ABI
// sources (api.hip: meshes, scenes, queries; render.hip: the wavefront
// pipeline, renderer and batches).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/convector.h"
#include "internal.hpp"

namespace cvh {

// Records the message returned by cv_last_error() and passes the status on.
cv_status fail(cv_status s, const std::string& msg);
cv_status hip_fail(hipError_t e, const char* what);
// Probes the HIP device once (CU count for the persistent launches).
cv_status ensure_device();
uint32_t num_cus();
const char* last_error();
// Copies between caller host memory and the device through a pinned bounce
// buffer, a chunk at a time, waiting for each chunk (and so for the work
// queued before it on st): the runtime never copies from or to the caller's
// pageable memory, and nothing touches that memory after the call returns.
cv_status copy_caller_memory(void* dst, const void* src, size_t bytes, bool to_device,
                             hipStream_t st);

inline uint32_t grid_for(size_t n, uint32_t block, uint32_t cap = 1u << 20) {
  size_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (uint32_t)b;
}

// Device allocation owned by a host object.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    reset();
    n = count;
    return hipMalloc((void**)&p, (count ? count : 1) * sizeof(T));
  }
  // Grows (without keeping the contents) when fewer than count elements.
  hipError_t reserve(size_t count) { return n >= count && p ? hipSuccess : alloc(count); }
};

// Pinned host allocation (asynchronous copies read it when they execute).
template <typename T>
struct HostBuf {
  T* p = nullptr;
  size_t n = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { reset(); }
  void reset() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t reserve(size_t count) {
    if (n >= count && p) return hipSuccess;
    reset();
    n = count;
    return hipHostMalloc((void**)&p, (count ? count : 1) * sizeof(T), hipHostMallocDefault);
  }
};

// Launch shape of the traversal kernels.
struct LaunchCfg {
  int stack = 8;              // per-lane LDS stack entries
  uint32_t blocks = 0;        // persistent 64-lane workgroups
  uint32_t refill = 16;       // 1..64 lane refill threshold, 65 whole 64-ray batches
  uint32_t fetch_batches = 1; // 64-ray batches per fetch in batch mode
  uint32_t block_waves = 1;   // waves per workgroup (lane refill)
  uint32_t frames_per_step = CV_FRAMES_AUTO;  // accumulation frames sharing launches
  uint32_t refill_interval = 16;              // traversal steps between refill checks
  uint32_t leaf_batch = 16;                   // 64ths of the active lanes per leaf round; 65 = off
  uint32_t width = 2;                         // BVH children per node: 2 (the reference's), 4, 8
  uint32_t sort_bits = 0;                     // ray-sort key bits after the camera rays (0: none)
  uint32_t parts = 8;                         // queue partitions of the traversal: 1 or 8
};

}  // namespace cvh

#define HIPTRY(x, what)                                        \
  do {                                                         \
    hipError_t e_ = (x);                                       \
    if (e_ != hipSuccess) return cvh::hip_fail(e_, what);      \
  } while (0)

struct cv_mesh {
  std::vector<float> v9;
  std::vector<float> uv6;
  std::vector<uint32_t> mat;
  size_t ntris() const { return mat.size(); }
};

// A resident scene: BVH nodes and triangles in BVH order, per-triangle
// material / uv / source id, the direct-sample light list, the camera and
// the traversal launch shape.
struct cv_scene {
  hipStream_t stream = nullptr;
  uint32_t ntris = 0;
  cvh::DevBuf<cvk::DevNode> nodes;
  cvh::DevBuf<cvk::DevTri> tris;
  cvh::DevBuf<uint32_t> tri_mat;
  cvh::DevBuf<float> tri_uv;
  cvh::DevBuf<uint32_t> tri_src;
  cvh::DevBuf<uint32_t> direct;
  cvh::DevBuf<cvk::DevLight> lights;
  uint32_t ndirect = 0;
  cv_bvh_info info{};
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounds of the triangles (the root's box)
  bool packed = true;
  cvk::CameraParams cam{};
  bool has_override = false;
  float emission[3] = {0, 0, 0};
  cvh::LaunchCfg cfg;              // traversal launch shape (cv_scene_set_tuning)
  cvh::DevBuf<uint2> spill;        // stack spill area of cv_bvh_intersect_nearest
  cvh::DevBuf<uint32_t> cursor;
  cvh::DevBuf<cvk::Ctrl> q_ctrl;   // hits-only queries: one-bounce step control block
  // the tree collapsed to 4 and 8 children per node (built on first use)
  cvh::DevBuf<float4> wide[2];
  uint32_t wide_nodes[2] = {0, 0};
  uint32_t wide_depth[2] = {0, 0};
  std::mutex mu;
};

namespace cvh {

cv_status validate_tuning(const cv_tuning& t);
// Applies the non-zero fields of t to c.
void apply_tuning(LaunchCfg& c, const cv_tuning& t);
cv_tuning tuning_of(const LaunchCfg& c);

// The job of a step for scene s (geometry, lights, camera, emission); its
// nodes are the tree of the given width (built by ensure_wide beforehand).
cvk::JobDesc job_for_scene(cv_scene* s, uint32_t width = 2);

// Builds the scene's tree of `width` children per node (4 or 8; 2 is the
// reference tree itself) once; ensure_wide_locked expects s->mu held.
cv_status ensure_wide(cv_scene* s, uint32_t width);
cv_status ensure_wide_locked(cv_scene* s, uint32_t width);
// The node array the traversal of that width reads.
const cvk::DevNode* traversal_nodes(const cv_scene* s, uint32_t width);
// What the traversal of that width needs in place of the tree depth when
// stack spill areas are sized (depth + 4 entries per lane); the wide tree
// must have been built.
uint32_t stack_depth(const cv_scene* s, uint32_t width);

// Denoises one accumulated image (hdr and features on the device, in the
// renderer's layouts) into dst (device or host, hdr layout) on the
// denoiser's own stream; allocates only when the image is larger than any
// before.
cv_status denoise_image(cv_denoiser* d, const float* hdr, const float* features, uint32_t width,
                        uint32_t height, uint32_t spp, float* dst, bool dst_device);

}  // namespace cvh
