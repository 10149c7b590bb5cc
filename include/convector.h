/*
 * convector.h -- C ABI of the MI355X (gfx950) per-ray path of Convector.
 *
 * The reference (ruuda/convector) is a Rust crate with no FFI; these entry
 * points are what a Rust crate would bind with `extern "C"` to replace its
 * hot path.  Each declaration names the reference item it stands in for.
 *
 * Conventions
 *  - Every reference panic/assert on the path becomes a cv_status; the text
 *    of the last error of the calling thread is in cv_last_error().
 *  - Handles are opaque.  cv_scene and cv_renderer calls that render are
 *    serialised internally, so one handle may be shared between threads
 *    (the reference shares `&Renderer` between its patch workers).
 *  - Buffers are caller owned.  With CV_PTR_DEVICE in `flags` the pointers
 *    are device pointers (hipMalloc'd); otherwise host pointers.
 *  - Materials are the reference's 32-bit encoding (material.rs:10-30).
 *  - Triangles are 9 floats v0,v1,v2; texture coordinates 6 floats
 *    uv0,uv1,uv2.  Triangle ids are source indices in load order: faces in
 *    file order after fan triangulation, meshes concatenated
 *    (wavefront.rs:125-140, bvh.rs:547-568).
 *  - Barycentrics follow triangle.rs:109-134: u weights v2, v weights v1,
 *    w = 1 - u - v weights v0.
 */
#ifndef CONVECTOR_H
#define CONVECTOR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cv_status {
    CV_OK = 0,
    CV_ERR_INVALID_ARGUMENT = 1,
    CV_ERR_IO = 2,                  /* FileBuffer::open(..).expect, wavefront.rs:86 */
    CV_ERR_PARSE = 3,               /* from_str(..).unwrap / missing index, wavefront.rs:46-53,100 */
    CV_ERR_DEGENERATE_TRIANGLE = 4, /* assert_nondegenerate panic, wavefront.rs:30-46 */
    CV_ERR_UNKNOWN_MATERIAL = 5,    /* usemtl lookup panic, wavefront.rs:116-123 */
    CV_ERR_NOT_MULTIPLE_OF_16 = 6,  /* assert_eq!(w & 15, 0), renderer.rs:46-47,313,346 */
    CV_ERR_LIGHT_COUNT = 7,         /* 8 direct-sample triangles needed, scene.rs:147-160 */
    CV_ERR_BVH_NO_SPLIT = 8,        /* assert!(!is_first) / root split, bvh.rs:303,512 */
    CV_ERR_NAN = 9,                 /* partial_cmp(..).unwrap() on NaN, bvh.rs:186 */
    CV_ERR_DEVICE = 10,             /* HIP runtime failure */
    CV_ERR_OUT_OF_MEMORY = 11,
    CV_ERR_NO_DEVICE = 12,
    CV_ERR_OUT_OF_RANGE = 13        /* patch outside the frame */
} cv_status;

/* ---------------------------------------------------------------- meshes */

typedef struct cv_mesh cv_mesh;

/* One entry of the name -> material dictionary of
 * Mesh::load_with_materials (wavefront.rs:83-85). */
typedef struct cv_material_entry {
    const char* name;
    uint32_t material;
} cv_material_entry;

/* Mesh::load / Mesh::load_with_materials (wavefront.rs:79-152): `v`, `vt`,
 * `usemtl` and `f` lines, convex n-gons fan triangulated, degenerate
 * triangles rejected, default material SMaterial::white().
 * table may be NULL (n = 0) for Mesh::load. */
cv_status cv_mesh_load_obj(const char* path, const cv_material_entry* table, size_t n,
                           cv_mesh** out);
/* Builds a mesh from a triangle soup (no reference counterpart; the
 * generator's path).  uv may be NULL.  Degenerate triangles are rejected as
 * in wavefront.rs:30-46. */
cv_status cv_mesh_from_triangles(const float* v9, const float* uv6, const uint32_t* materials,
                                 size_t ntris, cv_mesh** out);
size_t cv_mesh_num_triangles(const cv_mesh* mesh);
/* Copies the triangles out (any pointer may be NULL). */
cv_status cv_mesh_copy_triangles(const cv_mesh* mesh, float* v9, float* uv6, uint32_t* materials);
void cv_mesh_destroy(cv_mesh* mesh);

/* ---------------------------------------------------------------- scenes */

typedef struct cv_scene cv_scene;

/* The reference's BvhNode (bvh.rs:20-30): 32 bytes, siblings in pairs.
 * len == 0: internal node, children at index, index+1.
 * len > 0: leaf holding triangles [index, index+len) of the BVH order. */
typedef struct cv_bvh_node {
    float min[3];
    float max[3];
    uint32_t index;
    uint32_t len;
} cv_bvh_node;

#define CV_BUILDER_REFERENCE 0u /* 64 bins, TreeSAH, sorted fallback: bvh.rs:91-545 */
/* The same build with 32 or 16 centroid bins per axis (bvh.rs:275 with
 * bins.len() changed: sorted bins for at most bins/2 triangles or more than
 * n/8 of n > bins in one bin, chunks of ceil(n/bins)).  A different tree:
 * nearest hits equal the reference's except at exact ties in t; the
 * traversal counters are this tree's. */
#define CV_BUILDER_BINS_32 32u
#define CV_BUILDER_BINS_16 16u

typedef struct cv_scene_opts {
    uint32_t builder;               /* CV_BUILDER_REFERENCE, CV_BUILDER_BINS_32 or _16 */
    uint32_t has_emission_override; /* 0: emissive hits give sky_intensity(dir) */
    float emission_override[3];     /* radiance of emissive triangle hits when set */
} cv_scene_opts;

typedef struct cv_bvh_info {
    uint32_t num_nodes;       /* nodes array length (pairs, bvh.rs:514-533) */
    uint32_t num_triangles;
    uint32_t depth;           /* levels below the implicit root */
    uint32_t max_leaf;        /* largest leaf */
    uint32_t num_direct;      /* Scene::direct_sample length (scene.rs:114-129) */
    float avg_area_ratio;     /* bvh.rs:536-544 */
    float avg_tris_per_leaf;
    float build_ms;           /* device time of the GPU build */
    uint32_t bins;            /* centroid bins of the build (64: the reference's; 0: a
                                 tree given by the caller) */
} cv_bvh_info;

/* Scene::from_meshes + Bvh::from_meshes + Bvh::build (scene.rs:114-129,
 * bvh.rs:487-568): concatenates the meshes, builds the reference BVH on the
 * GPU and uploads the scene; the camera starts as Camera::new (scene.rs:33-42).
 * opts may be NULL. */
cv_status cv_scene_create(const cv_mesh* const* meshes, size_t n, const cv_scene_opts* opts,
                          cv_scene** out);
/* Same, with a BVH given by the caller (a serialised tree in the reference
 * layout plus the BVH-order -> source-id permutation) instead of building. */
cv_status cv_scene_create_with_bvh(const cv_mesh* const* meshes, size_t n,
                                   const cv_scene_opts* opts, const cv_bvh_node* nodes,
                                   size_t num_nodes, const uint32_t* order, cv_scene** out);
/* Camera::set_position + set_orientation (scene.rs:47-58) and the screen
 * distance of Camera::set_fov (scene.rs:61-63).  q = (a, b, c, d). */
cv_status cv_scene_set_camera(cv_scene* scene, const float pos[3], const float dpos[3],
                              const float q[4], const float dq[4], float screen_distance);
/* Bvh::print_stats / Scene::print_stats as numbers (bvh.rs:570-583). */
cv_status cv_scene_get_bvh_info(const cv_scene* scene, cv_bvh_info* info);
/* Launch shape of the kernels, the knobs of the autotuner (no reference
 * counterpart; the reference tunes its patch size by hand, main.rs:91-98).
 * Zero fields keep their current value. */
typedef struct cv_tuning {
    uint32_t stack_entries;       /* per-lane LDS stack entries: 8, 16, 24 or 32 */
    uint32_t extend_waves_per_cu; /* persistent traversal waves per CU */
    uint32_t refill_threshold;    /* 1..64: idle lanes of a wave take new rays when at least
                                     this many are idle; 65: whole 64-ray batches */
    uint32_t fetch_batches;       /* 64-ray batches per cursor fetch (whole-batch mode), or
                                     rays a wave reserves per fetch / 64 (lane refill) */
    uint32_t block_waves;         /* lane refill: 1, 2 or 4 waves (64-256 threads) per block */
    uint32_t frames_per_step;     /* frame accumulation: 1..64 frames rendered in the same
                                     launches (each into its own buffer, added in frame order),
                                     or CV_FRAMES_AUTO (about 2^25 paths per step) */
    uint32_t refill_interval;     /* lane refill: 1..64 traversal steps between refill checks */
    uint32_t leaf_batch;          /* lane refill: 1..64: a lane at a leaf waits until
                                     leaf_batch / 64 of the wave's active lanes wait (or all
                                     do) and they test their triangles together (16: a
                                     quarter); 65: every lane tests its leaf at once */
    uint32_t bvh_width;           /* children per traversed node: 2, the reference's node
                                     pairs (bvh.rs:20-30), the default; 4 or 8, the same tree
                                     with inner levels collapsed, visited nearest child first:
                                     hits equal the reference's except at exact ties and
                                     last-bit box decisions, counters those of the wide tree
                                     (needs leaves of at most 31 triangles) */
    uint32_t ray_sort_bits;       /* rendering: before each bounce after the camera rays, the
                                     paths are traversed in the order of a key of this many
                                     bits (3 to 24, a multiple of 3): the direction octant and
                                     a Morton code of the origin in the scene's bounds
                                     ((bits - 3) / 3 bits per axis), under the job index in a
                                     multi-job step; CV_RAY_SORT_NONE (the default): queue
                                     order.  Only the order of the work changes: results are
                                     the same bits */
    uint32_t queue_parts;         /* traversal: 1, every wave takes rays from one queue
                                     cursor; 8, the queue is cut into chunks of 2^12..2^18
                                     rays dealt to 8 partitions in turn, and a wave takes rays
                                     from the partition of its XCD (each XCD has its own L2)
                                     until it is used up, then from the others.  Only which
                                     lane traverses which ray changes: results are the same
                                     bits */
} cv_tuning;
#define CV_FRAMES_AUTO 255u
#define CV_RAY_SORT_NONE 255u
cv_status cv_scene_set_tuning(cv_scene* scene, const cv_tuning* tuning);
cv_status cv_scene_get_tuning(const cv_scene* scene, cv_tuning* tuning);
/* Copies the BVH out: nodes (num_nodes) and order (num_triangles). */
cv_status cv_scene_copy_bvh(const cv_scene* scene, cv_bvh_node* nodes, uint32_t* order);
void cv_scene_destroy(cv_scene* scene);

/* -------------------------------------------------------------- traversal */

typedef struct cv_ray {
    float origin[3];
    float direction[3];
} cv_ray;

#define CV_NO_HIT 0xffffffffu

typedef struct cv_hit {
    uint32_t tri_id; /* source triangle id or CV_NO_HIT (sky, distance 1e5) */
    float t;
    float u;
    float v;
} cv_hit;

/* MIntersection (ray.rs:35-50) plus the debug counters of
 * Bvh::intersect_debug (bvh.rs:604-669). */
typedef struct cv_isect {
    float position[3];
    float normal[3];
    float distance;
    uint32_t material;
    float tex_coords[2];
    uint32_t tri_id;
    float u;
    float v;
    uint32_t numi_aabb;
    uint32_t numi_tri;
    uint32_t reserved;
} cv_isect;

#define CV_PTR_DEVICE 1u /* buffers are device pointers */

/* Scene::intersect_nearest / Bvh::intersect_nearest / intersect_debug
 * (scene.rs:220-244, bvh.rs:589-669), batched, one ray per GPU lane.  Each
 * ray is traversed as the reference traverses a packet holding that ray in
 * every lane.  hits and isects may each be NULL.  A call with hits and no
 * isects traverses with the scene's tuning, taking at least 4 batches of 64
 * rays per fetch and checking for refills every refill_interval / 4 steps
 * (at least 1).  hits must be 16-byte aligned.  Returns after the results
 * are written. */
cv_status cv_bvh_intersect_nearest(const cv_scene* scene, const cv_ray* rays, size_t n,
                                   cv_hit* hits, cv_isect* isects, uint32_t flags);

/* The same query with no allocation at all.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  rays, hits and isects are device pointers (hits and isects
 * may each be NULL); the results equal cv_bvh_intersect_nearest's with
 * CV_PTR_DEVICE bit for bit.  The traversal's scratch (work cursors and the
 * stack spill area) is a fixed array of CV_NOALLOC_SCRATCH_BYTES in the
 * device memory of every GPU the library runs on, held until the process
 * ends; scenes whose BVH is deeper than CV_NOALLOC_MAX_DEPTH levels are
 * refused with CV_ERR_OUT_OF_RANGE.  Every call on a GPU shares that GPU's
 * array and overwrites it: calls are serialised inside the library and
 * return after the results are written, so calls that overlapped on one GPU
 * would race on the array. */
#define CV_NOALLOC_MAX_DEPTH 44u
#define CV_NOALLOC_SCRATCH_BYTES (8192u + 16u * 1024u * 1024u)
cv_status cv_bvh_intersect_nearest_noalloc(const cv_scene* scene, const cv_ray* rays, size_t n,
                                           cv_hit* hits, cv_isect* isects);

/* --------------------------------------------------------------- renderer */

typedef struct cv_renderer cv_renderer;

#define CV_RNG_PIXEL 0u     /* xorshift64 per pixel, seed splitmix64(pixel ^ frame << 32) */
#define CV_RNG_REF_PATCH 1u /* the reference Rng per 32x32 patch (random.rs, renderer.rs:316) */
#define CV_APPROX_PORTABLE 0u /* recip_fast = 1/x, rsqrt = 1/sqrt(x), correctly rounded */

typedef struct cv_render_opts {
    uint32_t rng_mode;    /* CV_RNG_PIXEL */
    uint32_t approx_mode; /* CV_APPROX_PORTABLE */
    uint32_t max_bounces; /* 5 (renderer.rs:422); 0 selects 5 */
} cv_render_opts;

typedef struct cv_render_stats {
    uint64_t paths;            /* camera paths traced by the last call */
    uint64_t extend_rays;      /* intersect_nearest queries of the last call */
    uint64_t extend_launches;  /* traversal kernel launches of the last call */
    double extend_ms;          /* summed device time of those launches */
    double shade_ms;           /* summed device time of the shading launches */
    double generate_ms;        /* summed device time of the camera-ray launches */
    double total_ms;           /* device time from first to last launch */
    uint64_t sum_numi_aabb;    /* with CV_RENDER_COUNT: summed box tests */
    uint64_t sum_numi_tri;     /* with CV_RENDER_COUNT: summed triangle tests */
    double sort_ms;            /* summed device time of the ray-sort launches */
} cv_render_stats;

#define CV_RENDER_COUNT 2u /* count box/triangle tests (slower) */
#define CV_RENDER_TIMING 4u /* record HIP events around every launch */
#define CV_RENDER_GRAPH 8u  /* capture a step's launches once as a HIP graph and replay it
                               (ignored with CV_RENDER_TIMING) */

/* Renderer::new (renderer.rs:121-130).  The renderer keeps a reference to
 * the scene (which must outlive it).  width must be a multiple of 16 and
 * height of 4: the 16x4-block HDR layout (renderer.rs:366-385).  The
 * reference's RenderBuffer further asks for heights divisible by 16
 * (renderer.rs:45-47); patch calls keep that rule.  Fails with CV_ERR_LIGHT_COUNT when the scene
 * has fewer than 8 direct-sample triangles.  opts may be NULL. */
cv_status cv_renderer_create(cv_scene* scene, uint32_t width, uint32_t height,
                             const cv_render_opts* opts, cv_renderer** out);
/* Renderer::render_patch_u8 (renderer.rs:306-327): renders the
 * patch_width x patch_width square at (x, y) (bottom-left) of frame
 * `frame` into the RGBA bitmap and gbuffer (width*height u32 each, row y
 * at offset y*width, y = 0 the bottom row). */
cv_status cv_renderer_render_patch_u8(cv_renderer* r, uint32_t* bitmap, uint32_t* gbuffer,
                                      uint32_t patch_width, uint32_t x, uint32_t y,
                                      uint32_t frame, uint32_t flags);
/* Renderer::accumulate_patch_f32 (renderer.rs:339-362): adds the patch's
 * colours to the HDR buffer in the reference's 16x4 block layout
 * ((height/4)*(width/16) blocks of [8 groups][3 channels][8 lanes] f32). */
cv_status cv_renderer_accumulate_patch_f32(cv_renderer* r, float* hdr, uint32_t* gbuffer,
                                           uint32_t patch_width, uint32_t x, uint32_t y,
                                           uint32_t frame, uint32_t flags);
/* The whole frame at once: every patch of main.rs:174-203 in one launch
 * sequence.  gbuffer may be NULL. */
cv_status cv_renderer_render_frame_u8(cv_renderer* r, uint32_t* bitmap, uint32_t* gbuffer,
                                      uint32_t frame, uint32_t flags);
/* Accumulates frames first_frame .. first_frame+n_frames-1 into hdr
 * (block layout) in frame order; gbuffer (may be NULL) receives the last
 * frame's gbuffer. */
cv_status cv_renderer_accumulate_frames_f32(cv_renderer* r, float* hdr, uint32_t* gbuffer,
                                            uint32_t first_frame, uint32_t n_frames,
                                            uint32_t flags);
/* accumulate_frames_f32 plus the denoiser's first-hit features (the cv_job
 * features layout, accumulated into; no reference counterpart). */
cv_status cv_renderer_accumulate_aov_f32(cv_renderer* r, float* hdr, float* features,
                                         uint32_t first_frame, uint32_t n_frames,
                                         uint32_t flags);
/* Renderer::buffer_f32_into_render_buffer (renderer.rs:378-410). */
cv_status cv_renderer_buffer_f32_into_bitmap(cv_renderer* r, const float* hdr, uint32_t* bitmap,
                                             uint32_t num_samples, uint32_t flags);
/* Counters and device times of the last render call. */
cv_status cv_renderer_get_stats(const cv_renderer* r, cv_render_stats* stats);
void cv_renderer_destroy(cv_renderer* r);

/* ---------------------------------------------------------------- batches */

/* Many render jobs in shared launches: the patch loop of main.rs:174-214
 * generalised to jobs of different scenes, sizes and sample counts (no
 * reference counterpart beyond that loop).  Step s renders frame
 * first_frame + s of every job with more than s frames, all of them in the
 * same generate / extend / shade launches, so a job's result is identical
 * to cv_renderer_accumulate_frames_f32 of the same scene, size and
 * frames, alone or in any mix. */
typedef struct cv_batch cv_batch;

typedef struct cv_job {
    cv_scene* scene;       /* resident scene; its camera at the call is used */
    uint32_t width;        /* multiple of 16 */
    uint32_t height;       /* multiple of 4 */
    uint32_t first_frame;  /* frames first_frame .. first_frame + num_frames - 1 */
    uint32_t num_frames;   /* samples per pixel, >= 1 */
    float* hdr;            /* accumulated into: 16x4-block layout, width*height*3 f32 */
    uint32_t* gbuffer;     /* NULL, or the last frame's gbuffer (width*height u32) */
    float* features;       /* NULL, or accumulated into: first-hit denoiser features summed
                              over the frames, 8 f32 per pixel in rows (albedo rgb, shading
                              normal xyz, hit distance (1e5 for sky), count) */
} cv_job;

/* max_paths_per_step bounds the path queues (0: 2^27, about 17 GB of queues
 * when a batch needs that many; queues grow only to the largest step a call
 * renders); a step whose jobs hold more pixels is split into several steps
 * by jobs (a job larger than the bound gets a step of its own).  Fewer,
 * larger steps mean fewer traversal launches, each of which ends with a
 * drain.  opts may be NULL. */
cv_status cv_batch_create(const cv_render_opts* opts, uint64_t max_paths_per_step,
                          cv_batch** out);
/* Launch shape of the batch's traversal (see cv_tuning). */
cv_status cv_batch_set_tuning(cv_batch* b, const cv_tuning* tuning);
cv_status cv_batch_get_tuning(const cv_batch* b, cv_tuning* tuning);
/* Renders all jobs and returns when their outputs are written.  flags:
 * CV_PTR_DEVICE (hdr and gbuffer are device pointers), CV_RENDER_COUNT,
 * CV_RENDER_TIMING, CV_RENDER_GRAPH. */
cv_status cv_batch_render(cv_batch* b, const cv_job* jobs, size_t n, uint32_t flags);
/* The autotuner: times launch shapes (stack entries 8/16/24/32, 16-48
 * persistent waves per CU, refill thresholds 16-64 and whole batches, 1/2/4
 * waves per workgroup, 1/4/16/32 steps between refill checks, leaf rounds,
 * ray-sort keys of 0/12/24 bits) on a sample step -- the first frame of the
 * jobs in order, up to max_sample_paths paths (0: 2^22), rendered into
 * scratch buffers -- by two rounds of coordinate descent, and keeps the
 * fastest on the batch.  Every one of these knobs leaves the results
 * unchanged.  With CV_TUNE_TOLERANCE in flags it also sweeps BVH width 2/4/8,
 * whose wide trees give the reference's hits except at exact ties (see
 * cv_tuning.bvh_width); without it the batch's width is kept.  chosen and
 * chosen_ms (may be NULL) receive the choice and its step time. */
#define CV_TUNE_TOLERANCE 1u
cv_status cv_batch_autotune(cv_batch* b, const cv_job* jobs, size_t n, uint64_t max_sample_paths,
                            uint32_t flags, cv_tuning* chosen, double* chosen_ms);
/* Counters and device times of the last cv_batch_render. */
cv_status cv_batch_get_stats(const cv_batch* b, cv_render_stats* stats);
void cv_batch_destroy(cv_batch* b);

/* ----------------------------------------------------------------- server */

/* A request loop over resident scenes with continuous batching (no
 * reference counterpart; the reference renders one interactive view,
 * main.rs:124-217).  A worker thread runs steps: each step admits waiting
 * requests in arrival order while the memory budget (resident scenes, path
 * queues and the admitted requests' buffers) and the step's path capacity
 * allow, then renders the next frame of every admitted request in shared
 * launches (a captured HIP graph).  A request is finished after spp steps
 * (then denoised if asked) and its result delivered; its image is the same
 * as when rendered alone (PIXEL streams are keyed by pixel and frame). */
typedef struct cv_server cv_server;
typedef struct cv_denoiser cv_denoiser; /* see the denoiser section */

typedef struct cv_request {
    uint32_t scene;            /* index returned by cv_server_add_scene */
    uint32_t width;            /* multiple of 16 (64 with denoise) */
    uint32_t height;           /* multiple of 4 */
    uint32_t spp;              /* frames 1..spp */
    uint32_t denoise;          /* 1: run the server's denoiser on the result */
    uint32_t flags;            /* CV_PTR_DEVICE: hdr is a device pointer */
    float camera_pos[3];       /* camera position */
    float camera_q[4];         /* camera orientation quaternion (scene.rs:47-63) */
    float screen_distance;     /* 0: the scene's */
    float* hdr;                /* result: 16x4-block HDR holding spp x radiance */
} cv_request;

typedef struct cv_server_stats {
    uint64_t submitted, completed, steps;
    double p50_ms, p99_ms, mean_ms; /* submit -> delivered latency of completed requests */
    double busy_ms;                 /* summed device time of the steps */
    uint64_t peak_bytes;            /* peak of resident + admitted bytes */
    uint64_t resident_bytes;        /* scenes + path queues */
} cv_server_stats;

/* hbm_budget_bytes: 0 selects 200 GiB; max_paths_per_step: 0 selects 2^23. */
cv_status cv_server_create(const cv_render_opts* opts, uint64_t hbm_budget_bytes,
                           uint64_t max_paths_per_step, cv_server** out);
/* Registers a resident scene (it must outlive the server). */
cv_status cv_server_add_scene(cv_server* sv, cv_scene* scene, uint32_t* id);
/* The denoiser used by requests with denoise = 1 (must outlive the server).
   The server's finisher thread runs it; the caller must not use the same
   denoiser from another thread while denoised requests are in flight. */
cv_status cv_server_set_denoiser(cv_server* sv, cv_denoiser* d);
cv_status cv_server_set_tuning(cv_server* sv, const cv_tuning* tuning);
/* Queues a request and returns at once; thread safe. */
cv_status cv_server_submit(cv_server* sv, const cv_request* req, uint64_t* ticket);
/* done = 1 when the request's result is delivered (or it failed). */
cv_status cv_server_poll(cv_server* sv, uint64_t ticket, int* done);
/* Blocks until the request is delivered; returns its status.  The ticket is
 * released. */
cv_status cv_server_wait(cv_server* sv, uint64_t ticket, double* latency_ms);
cv_status cv_server_get_stats(cv_server* sv, cv_server_stats* stats);
cv_status cv_server_reset_stats(cv_server* sv);
/* Finishes every queued request, then stops the worker. */
void cv_server_destroy(cv_server* sv);

/* --------------------------------------------------------------- denoiser */

/* The learned denoiser that stands where the reference runs its median
 * filter (src/gpu/median.glsl:51-75; no reference counterpart): a 3-level
 * U-Net (32/64/128 channels, 3x3 convolutions, ReLU, 2x2 average pooling,
 * nearest upsampling, skip concatenation), BF16 weights and activations on
 * the matrix cores with FP32 accumulation, FP32 master weights and Adam.
 * Input: 10 f32 per pixel in rows (log1p radiance rgb, albedo rgb,
 * octahedral normal xy, log1p distance, spp; see
 * cv_denoiser_prepare_input); output: 3 f32 per pixel, log1p radiance.
 * Batches are nb images of height x width, height a multiple of 4 and width
 * of 64.  Pointers are host memory unless flags has CV_PTR_DEVICE. */
typedef struct cv_denoiser cv_denoiser;

/* Number of f32 parameters (weights [cout][3][3][cin] then bias, per layer). */
size_t cv_denoiser_num_params(void);
/* He-normal weights from a splitmix64 stream seeded with seed, zero bias. */
cv_status cv_denoiser_create(uint64_t seed, cv_denoiser** out);
cv_status cv_denoiser_get_params(cv_denoiser* d, float* params, float* grads);
/* Replaces the parameters and resets the Adam state. */
cv_status cv_denoiser_set_params(cv_denoiser* d, const float* params);
/* flags: CV_PTR_DEVICE; CV_DENOISE_MATERIALIZE runs the separate pooling
 * and upsampling kernels (the pooled encoder outputs and the upsampled
 * decoder inputs written out by kernels of their own) instead of pooling in
 * the encoder convolutions' epilogues and reading the upsampled channels
 * inside the decoder convolutions: the same values and outputs, slower. */
#define CV_DENOISE_MATERIALIZE 2u
cv_status cv_denoiser_forward(cv_denoiser* d, const float* input, float* output, uint32_t nb,
                              uint32_t height, uint32_t width, uint32_t flags);
/* Forward, L1 loss against log1p(target) (3 f32 per pixel, linear
 * radiance), backward into the gradients, then one Adam step with learning
 * rate lr (lr = 0: gradients only).  loss (may be NULL) receives the mean
 * absolute error before the step, summed in a fixed order (the same value
 * for the same inputs and weights on every call).  flags: CV_PTR_DEVICE;
 * CV_DENOISE_MATERIALIZE runs the separate pooling, upsampling-gradient and
 * ReLU-gradient kernels instead of their fused forms (the same bits). */
cv_status cv_denoiser_train_step(cv_denoiser* d, const float* input, const float* target,
                                 uint32_t nb, uint32_t height, uint32_t width, float lr,
                                 uint32_t flags, float* loss);
/* Network input from a renderer's HDR accumulation (16x4-block layout, spp
 * frames) and its summed features (cv_renderer_accumulate_aov_f32). */
cv_status cv_denoiser_prepare_input(const float* hdr, const float* features, uint32_t width,
                                    uint32_t height, uint32_t spp, float* input, uint32_t flags);
/* Network output -> 16x4-block HDR holding spp x radiance, so that
 * cv_renderer_buffer_f32_into_bitmap(..., spp) displays it. */
cv_status cv_denoiser_output_to_hdr(const float* output, uint32_t width, uint32_t height,
                                    uint32_t spp, float* hdr, uint32_t flags);
void cv_denoiser_destroy(cv_denoiser* d);

/* ------------------------------------------------------------------ trace */

/* TraceLog (trace.rs:22-152) for the device: while enabled, every
 * generate / extend / shade launch of renderers, batches and servers is
 * timed with HIP events (steps are not replayed from graphs) and kept in a
 * ring of the last `limit` spans; cv_trace_export writes them in the Chrome
 * trace format of trace.rs:122-146 (name, ph "X", ts and dur in
 * microseconds from cv_trace_enable, tid = kernel kind, args frame and id =
 * bounce). */
cv_status cv_trace_enable(size_t limit);
cv_status cv_trace_disable(void);
cv_status cv_trace_export(const char* path);

/* ----------------------------------------------------------------- device */

/* Waits for all work of this library on the current device. */
cv_status cv_device_synchronize(void);
/* Text of the last error raised on the calling thread ("" if none). */
const char* cv_last_error(void);
const char* cv_status_string(cv_status s);
/* Library version string. */
const char* cv_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CONVECTOR_H */
