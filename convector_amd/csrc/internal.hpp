// internal.hpp -- argument blocks and launch wrappers shared by the kernels
// and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"

namespace cvk {

constexpr int kMaxBounceSlots = 16;
// Shade workgroup: 16 waves, so that the compaction's one atomic per
// workgroup stays below the kernel's memory time (one cursor serialises at
// about 12 ns an atomic).
constexpr uint32_t kShadeBlock = 1024;

// Per-step control block in device memory: the step's path count and job
// count (written by the host before each step), queue sizes and work cursors
// per bounce, plus the optional traversal counters.
// Queue partitions of the lane-refill traversal: one per XCD.  Each
// partition's cursor has a 128-byte line of its own.
constexpr uint32_t kParts = 8;
constexpr uint32_t kPartStride = 32;

struct Ctrl {
  uint32_t step_paths;
  uint32_t step_jobs;
  uint32_t count[kMaxBounceSlots];
  uint32_t ext_cursor[kMaxBounceSlots];
  uint32_t shade_cursor[kMaxBounceSlots];
  unsigned long long sum_aabb;
  unsigned long long sum_tri;
  unsigned long long total_rays;
  alignas(128) uint32_t part_cursor[kMaxBounceSlots][kParts * kPartStride];
};

// Paths in HBM as structure of arrays; every field is one 16-B or 8-B load.
struct Queue {
  float4* r0;     // origin.xyz, pixel index bits
  float4* r1;     // direction.xyz, job index bits
  float4* col;    // path throughput rgb
  uint64_t* rng;  // PIXEL stream state
};

struct CameraParams {
  float pos[3];
  float dpos[3];
  float q[4];
  float dq[4];
  float screen_distance;
};

// One render job of a step: a resident scene, an image rectangle, a frame
// number and the outputs.  A step renders one frame of each of its jobs, so
// that a pixel receives at most one path per step and accumulation keeps
// frame order.  Jobs are sorted by path_base.
struct JobDesc {
  const DevNode* nodes;
  const DevTri* tris;
  const uint32_t* tri_mat;
  const float* tri_uv;
  const DevLight* lights;
  float* hdr;         // 16x4-block HDR accumulation, or null
  uint32_t* bitmap;   // RGBA8 rows, or null
  uint32_t* gbuffer;  // gbuffer rows, or null
  float* features;    // first-hit features summed over frames (8 f32 per pixel), or null
  CameraParams cam;
  float emission[3];
  uint32_t nlights;
  uint32_t has_override;
  uint32_t width, height;
  uint32_t rect_x, rect_y, rect_w, rect_h;
  uint32_t frame;
  uint32_t path_base;  // index of the job's first path in the step
  float lo[3];         // the scene's bounds: lower corner and 1 / extent per
  float inv_ext[3];    // axis (ray-sort keys)
};

struct GenerateArgs {
  Ctrl* ctrl;
  Queue q;
  const JobDesc* jobs;
};

struct ExtendArgs {
  Ctrl* ctrl;
  uint32_t bounce;
  const float4* r0;
  const float4* r1;
  uint4* hits;
  const DevNode* nodes;  // the scene when every job of the step shares one
  const DevTri* tris;
  const JobDesc* jobs;   // per-ray scenes otherwise (job index in r1.w)
  uint2* spill;
  uint32_t batches;  // 64-ray batches per cursor fetch (batch mode)
  uint32_t refill;   // > 0: lane refill when this many lanes are idle
  uint32_t interval; // lane refill: traversal steps between refill checks (>= 1)
  uint32_t leaf_batch;  // lane refill: > 0: lanes at leaves wait until leaf_batch / 64 of
                        // the active lanes do (or all)
  const float* rays6;         // non-null (batched queries): the rays as n x 6 floats, not r0/r1
  const uint32_t* tri_src;    // non-null (batched queries): hits name source triangles
  const uint32_t* order;      // non-null: queue position k traverses path order[k]
  uint32_t parts;             // lane refill: 1 (one cursor) or kParts queue partitions
};

// The ray sort before a bounce's traversal: a counting sort of the paths
// by key (job index, direction octant, Morton code of the origin) into
// `order`.  Buckets: 1 << key_bits counters; block_sums: one word per 4096
// buckets.
struct SortArgs {
  Ctrl* ctrl;
  uint32_t bounce;
  const float4* r0;
  const float4* r1;
  const JobDesc* jobs;
  uint32_t multi_job;   // job index in r1.w (else every path is job 0's)
  uint32_t axis_bits;   // Morton bits per axis
  uint32_t key_bits;    // job bits + 3 + 3 * axis_bits
  uint32_t* keys;
  uint32_t* buckets;
  uint32_t* block_sums;
  uint32_t* order;
};
constexpr uint32_t kSortScanBlock = 4096;  // buckets per scan workgroup

struct ShadeArgs {
  Ctrl* ctrl;
  uint32_t bounce;
  uint32_t max_bounces;
  Queue in;
  Queue out;
  const uint4* hits;
  const JobDesc* jobs;
};

// cv_bvh_intersect_nearest's work cursors: kIsectGroups words,
// kIsectCursorStride words (128 bytes) apart.
constexpr uint32_t kIsectGroups = 64;
constexpr uint32_t kIsectCursorStride = 32;

struct IntersectArgs {
  const float* rays;
  uint32_t n;
  uint32_t* cursor;
  const DevNode* nodes;
  const DevTri* tris;
  const uint32_t* tri_src;
  const uint32_t* tri_mat;
  const float* tri_uv;
  uint2* spill;
  uint32_t* hits;    // n x 4 u32 or null
  uint32_t* isects;  // n x 16 u32 or null
};

hipError_t launch_intersect(const IntersectArgs& a, bool count, bool packed, int stack,
                            uint32_t blocks, hipStream_t s);

// CV_RNG_REF_PATCH rendering: patches (x, y, width, 0) of the job's image,
// frames first_frame .. first_frame + nframes - 1 of each in order, into
// the job's hdr (added) or bitmap, and its gbuffer (the last frame's).
struct RefPatchArgs {
  JobDesc job;
  const uint4* patches;
  uint32_t npatches;
  uint32_t first_frame;
  uint32_t nframes;
  uint32_t max_bounces;
  uint2* spill;  // stack entries beyond the LDS part: stride gridDim.x * 64 lanes
};
// blocks: 64-lane workgroups, 8 patches each at a time (at most the
// spill area's waves).
hipError_t launch_ref_patch(const RefPatchArgs& a, bool packed, int stack, uint32_t blocks,
                            hipStream_t s);
// Hits-only queries through launch_extend: a zeroed control block with
// count[0] = n.
hipError_t launch_query_ctrl(Ctrl* ctrl, uint32_t n, hipStream_t s);
// multi: rays of the step belong to more than one scene (per-ray node and
// triangle pointers from the job table).
// blocks = persistent waves; block_waves (1, 2, 4) waves per workgroup in
// lane-refill mode.  width 2: the reference's node pairs (exact); 4 or 8:
// the collapsed wide tree of build_wide_bvh in place of the nodes (nodes and
// the jobs' node pointers then point at its float4 array).
hipError_t launch_extend(const ExtendArgs& a, bool count, bool packed, bool multi, int stack,
                         uint32_t blocks, uint32_t block_waves, hipStream_t s,
                         uint32_t width = 2);
hipError_t launch_generate(const GenerateArgs& a, uint32_t blocks, hipStream_t s);
// The ray sort of one bounce (four kernels: clear, keys, scan, scatter);
// `paths` bounds the step's path count (the grid of the per-path kernels).
hipError_t launch_sort(const SortArgs& a, uint32_t paths, hipStream_t s);
// multi: the step has more than one job.
// shared_scene: every job of the step has the same scene (tris, lights,
// emission), read once per wave from job 0.
hipError_t launch_shade(const ShadeArgs& a, bool multi, bool shared_scene, uint32_t max_paths,
                        hipStream_t s);
hipError_t launch_hdr_to_bitmap(const float* hdr, uint32_t* bitmap, uint32_t width,
                                uint32_t height, uint32_t num_samples, hipStream_t s);
hipError_t launch_prepare_tris(const float* v9, const uint32_t* order, uint32_t ntris, DevTri* out,
                               hipStream_t s);
hipError_t launch_prepare_lights(const DevTri* tris, const uint32_t* direct, uint32_t n,
                                 DevLight* out, hipStream_t s);
hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);
// hdr += slots[0] + ... + slots[count-1] (in that order) over the rectangle,
// clearing the slots; slot k starts at slots + k * slot_stride.
hipError_t launch_fold_hdr(float* hdr, float* slots, size_t slot_stride, uint32_t count,
                           uint32_t width, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                           hipStream_t s);
// features += slots[0] + ... + slots[count-1] (8 floats per pixel, in that
// order), clearing the slots.
hipError_t launch_fold_features(float* feat, float* slots, size_t slot_stride, uint32_t count,
                                uint32_t npix, hipStream_t s);

// ---------------------------------------------------------------- builder

struct BuildResult {
  uint32_t num_nodes;
  uint32_t depth;
  uint32_t max_leaf;
  uint32_t leaves;
  float avg_area_ratio;
  float avg_tris_per_leaf;
  float build_ms;
  int error;  // 0, or a cv_status value
};

// Builds the reference BVH (bvh.rs:91-545) of ntris triangles (device v9,
// source order) on the GPU, with nbins centroid bins per axis: 64 is the
// reference's build, 16 and 32 the same steps with bins.len() changed.
// nodes_out (device, 2*ntris DevNode) receives the depth-first pair layout
// of crystallize (bvh.rs:389-428), order_out (device, ntris) the
// BVH-position -> source-index permutation.
hipError_t build_reference_bvh(const float* d_v9, uint32_t ntris, DevNode* nodes_out,
                               uint32_t* order_out, BuildResult* res, hipStream_t s,
                               uint32_t nbins = 64);

}  // namespace cvk
