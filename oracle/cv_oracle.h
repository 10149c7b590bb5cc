/*
 * cv_oracle.h -- CPU restatement of Convector's per-ray path, used ONLY as
 * test infrastructure (the checker for the HIP path and the timed CPU
 * baseline).  Nothing in the product links or loads this library.
 *
 * Every function restates a function of the reference (ruuda/convector, Rust)
 * lane by lane, with the exact floating-point operation order, fused
 * multiply-adds where the reference calls mul_add/mul_sub/neg_mul_add, x86
 * minps/maxps/cmpps semantics, and the reference's sign-bit masks.  The
 * citations in cv_oracle.c give the reference file:line for each function.
 *
 * Modes
 *   approx  OR_APPROX_PORTABLE: recip_fast(x) = 1.0f/x and rsqrt(x) =
 *             1.0f/sqrtf(x), both correctly rounded (reproducible anywhere).
 *           OR_APPROX_HOST_X86: the host's rcpps / rsqrtps, which is what the
 *             reference executes (vendor specific, ~12 bits).
 *   rng     OR_RNG_REF_PATCH: the reference's multiplicative-hash Rng, seeded
 *             per 32x32 patch and shared by the 8 lanes of a group
 *             (random.rs:43-136, renderer.rs:316,349).
 *           OR_RNG_PIXEL: one xorshift64 stream per pixel, seeded with
 *             splitmix64(pixel_index ^ (frame << 32)); the draws are turned
 *             into floats with the reference's conversions and consumed in
 *             the reference's per-path order.
 */
#ifndef CV_ORACLE_H
#define CV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_APPROX_PORTABLE 0
#define OR_APPROX_HOST_X86 1

#define OR_RNG_PIXEL 0
#define OR_RNG_REF_PATCH 1

#define OR_NO_HIT 0xffffffffu

/* BvhNode (bvh.rs:20-30): 32 bytes, siblings stored in pairs. */
typedef struct {
    float min[3];
    float max[3];
    uint32_t index;
    uint32_t len;
} or_node;

typedef struct or_scene or_scene;

/* ---- primitives exported for known-answer tests ---- */
float or_rcp(float x, int approx);
float or_rsqrt(float x, int approx);
float or_sin_poly(float x);
float or_cos_poly(float x);
float or_acos_poly(float x);
float or_acos_fast_poly(float x);
float or_recip_precise(float x, int approx);
int32_t or_cvt_f32_i32(float x);
float or_x86_min(float a, float b);
float or_x86_max(float a, float b);
uint32_t or_geq_mask(float a, float b);

/* Reference Rng (random.rs:33-159): writes `ndraws` x 8 u32 lanes. */
void or_rng_ref_draws(uint32_t x, uint32_t y, uint32_t i, uint32_t ndraws, uint32_t* out);
/* Reference float conversions of one 8-lane draw. kind: 0 unit, 1 biunit, 2 angle. */
void or_rng_ref_floats(uint32_t x, uint32_t y, uint32_t i, uint32_t ndraws, int kind, float* out);
/* Reference cosine-weighted hemisphere vectors: ndraws x 8 lanes x 3. */
void or_rng_ref_hemisphere(uint32_t x, uint32_t y, uint32_t i, uint32_t n, float* out);
/* PIXEL stream: first ndraws u32 of the stream of (pixel_index, frame). */
void or_rng_pixel_draws(uint32_t pixel_index, uint32_t frame, uint32_t ndraws, uint32_t* out);

/* MVector3 helpers (vector3.rs) on arrays of n vectors. */
void or_rotate_hemisphere(const float* v, const float* normal, size_t n, int approx, float* out);
void or_quat_rotate(const float* v, const float* q, size_t n, float* out);
void or_quat_interpolate(const float* q, const float* dq, const float* t, size_t n, float* out);

/* Aabb::intersect (aabb.rs:105-141) for n rays against one box.
 * rays: n x 6 (origin, direction).  out: n x 3 (tmin, tmax, mask bits as float). */
void or_aabb_intersect(const float box[6], const float* rays, size_t n, int approx, float* tmin,
                       float* tmax, uint32_t* mask);
float or_aabb_area(const float box[6]);

/* Triangle::intersect (triangle.rs:59-148) for one lane.
 * tri: v0,v1,v2 (9 floats).  Returns 1 when the candidate replaces best_t.
 * out (if non-null): t, u, v, w. */
int or_triangle_intersect(const float tri[9], const float ray[6], float best_t, float* out4);
/* Triangle::intersect_direct (triangle.rs:152-185).  out: normal[3], area,
 * distance, mask (sign bit 1 = miss) */
void or_triangle_intersect_direct(const float tri[9], const float ray[6], int approx, float* out6);

/* ---- BVH build (bvh.rs:91-545) ---- */
/* Builds the reference BVH over ntris triangles (v: ntris x 9).
 * nodes_out must hold 2*ntris nodes; order_out ntris entries (sorted position
 * -> source triangle index).  Returns 0 on success, negative on the
 * reference's panics: -1 fewer than one split possible (root is a leaf),
 * -2 no valid split axis in a node (all centroids coincide), -3 NaN input. */
int or_bvh_build(const float* v, uint32_t ntris, or_node* nodes_out, uint32_t* nnodes_out,
                 uint32_t* order_out, float* avg_area_ratio, float* avg_tris_per_leaf);
/* The same build with nbins (16, 32 or 64) centroid bins per axis in place
 * of the reference's 64 (bvh.rs:275 with bins.len() changed: the uniform
 * threshold bins/2, the fallback test n > bins and the chunk size
 * ceil(n/bins) follow it).  -4 for another nbins. */
int or_bvh_build_bins(const float* v, uint32_t ntris, uint32_t nbins, or_node* nodes_out,
                      uint32_t* nnodes_out, uint32_t* order_out, float* avg_area_ratio,
                      float* avg_tris_per_leaf);
/* TreeSurfaceAreaHeuristic::aabb_cost (bvh.rs:452-469) */
float or_tree_sah_cost(const float parent_box[6], const float box[6], uint32_t num_tris);

/* ---- scene ---- */
/* tris: ntris x 9 vertices, uv: ntris x 6 or NULL, mat: ntris materials.
 * nodes/order: a BVH from or_bvh_build (or NULL to build here). */
or_scene* or_scene_create(const float* v, const float* uv, const uint32_t* mat, uint32_t ntris,
                          const or_node* nodes, uint32_t nnodes, const uint32_t* order,
                          int approx, int* status);
void or_scene_destroy(or_scene* s);
void or_scene_set_camera(or_scene* s, const float pos[3], const float dpos[3], const float q[4],
                         const float dq[4], float screen_distance);
void or_scene_set_emission_override(or_scene* s, int enable, const float rgb[3]);
void or_scene_set_approx(or_scene* s, int approx);
uint32_t or_scene_num_nodes(const or_scene* s);
uint32_t or_scene_num_direct(const or_scene* s);
void or_scene_copy_bvh(const or_scene* s, or_node* nodes, uint32_t* order);

/* Scene::intersect_nearest through the single-ray traversal (bvh.rs:589-657
 * with one active lane).  rays: n x 6.  Any output may be NULL.
 * isect (n x 12): position[3], normal[3], distance, material bits, tex u, tex v, pad, pad */
void or_intersect_nearest(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                          float* tuv, uint32_t* numi_aabb, uint32_t* numi_tri, float* isect,
                          int nthreads);
/* The reference's 8-lane packet traversal (bvh.rs:589-657 verbatim): rays are
 * grouped in packets of 8 consecutive rays (n multiple of 8). */
void or_intersect_nearest_packet(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                                 float* tuv, uint32_t* numi_aabb, uint32_t* numi_tri);
/* Brute-force nearest hit over all triangles (t in float64 using the same
 * formula) for checking traversal. */
void or_intersect_brute_f64(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                            double* tuv);

/* ---- shading, one lane at a time (material.rs, scene.rs) ---- */
/* microfacet_normal_dist (material.rs:475-513); gloss index 0..5 */
float or_microfacet_normal_dist(const float h[3], const float n[3], int32_t gloss);
/* microfacet_fresnel (material.rs:453-469): out = Schlick colour, raw = ct5 */
void or_microfacet_fresnel(const float incoming[3], const float h[3], const float color[3],
                           float out[3], float* raw);
/* microfacet_brdf (material.rs:415-450) */
void or_microfacet_brdf(uint32_t material, const float in_dir[3], const float out_dir[3],
                        const float normal[3], int ignore_fresnel, int approx, float out[3],
                        float* raw);
/* pd_brdf (material.rs:245-250) */
float or_pd_brdf(const float normal[3], const float dir[3]);
/* sky_intensity (material.rs:205-215) */
void or_sky_intensity_lane(const float dir[3], float out[3]);
/* pd_direct_sample (material.rs:281-320) over the scene's direct-sample
 * triangles for n rays (n x 6).  Returns -1 when the scene has none. */
int or_pd_direct(const or_scene* s, const float* rays, size_t n, float* out);
/* Scene::get_direct_sample (scene.rs:143-202) for n lanes; draws3 holds the
 * lane's sample_u32 draw and its two sample_unit draws (raw u32).  normal
 * and area may be NULL.  Returns -1 unless the scene has exactly 8
 * direct-sample triangles. */
int or_direct_sample(const or_scene* s, const uint32_t* draws3, size_t n, float* pos,
                     float* normal, float* area);
/* continue_path (material.rs:327-411) for n independent lanes.
 * ray7: origin, direction, active bits; isect7: position, normal, material
 * bits; draws6: the six draws in the reference's order (angle, unit, u32,
 * unit, unit, biunit).  Outputs the new ray (same layout), the colour
 * modulation and the raw Fresnel factor (may be NULL).  Returns -1 unless
 * the scene has exactly 8 direct-sample triangles. */
int or_continue_path(const or_scene* s, size_t n, const float* ray7, const float* isect7,
                     const uint32_t* draws6, int ignore_fresnel, float* out_ray7,
                     float* color_mod3, float* fresnel_raw);

/* ---- renderer (renderer.rs) ---- */
/* render_patch_u8 (renderer.rs:306-327). bitmap/gbuffer: width*height u32. */
int or_render_patch_u8(const or_scene* s, uint32_t width, uint32_t height, uint32_t* bitmap,
                       uint32_t* gbuffer, uint32_t patch_width, uint32_t x, uint32_t y,
                       uint32_t frame, int rng_mode);
/* accumulate_patch_f32 (renderer.rs:339-362).  hdr: (height/4)*(width/16)
 * blocks of [8 groups][3 comps][8 lanes] floats. */
int or_accumulate_patch_f32(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                            uint32_t* gbuffer, uint32_t patch_width, uint32_t x, uint32_t y,
                            uint32_t frame, int rng_mode);
/* buffer_f32_into_render_buffer (renderer.rs:378-410). */
int or_buffer_f32_into_bitmap(uint32_t width, uint32_t height, const float* hdr,
                              uint32_t* bitmap, uint32_t num_samples);
/* Renders whole frames with the patch loop of main.rs:174-214 over nthreads
 * workers (32x32 patches; the frame need only be a multiple of 16: edge
 * patches are 16 wide).  Accumulates frames first..first+n-1 into hdr. */
int or_accumulate_frames(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                         uint32_t* gbuffer, uint32_t first_frame, uint32_t n_frames, int rng_mode,
                         int nthreads);
int or_render_frame_u8(const or_scene* s, uint32_t width, uint32_t height, uint32_t* bitmap,
                       uint32_t* gbuffer, uint32_t frame, int rng_mode, int nthreads);
/* Per-pixel path statistics of one frame (PIXEL or REF_PATCH): number of
 * extend queries (intersect_nearest calls on active lanes) and their summed
 * box/triangle test counts. */
int or_frame_work(const or_scene* s, uint32_t width, uint32_t height, uint32_t frame, int rng_mode,
                  int nthreads, uint64_t* n_queries, uint64_t* sum_aabb, uint64_t* sum_tri);

/* CPU-baseline sample: renders the listed patches (column-major patch index)
 * of one frame; out3 = {extend queries, box tests, triangle tests}. */
int or_sample_patches(const or_scene* s, uint32_t width, uint32_t height, uint32_t frame,
                      int rng_mode, int nthreads, const uint32_t* patch_ids, uint32_t n_ids,
                      uint64_t* out3);

/* Converts between the 16x4-block HDR layout and row-major RGB float. */
void or_hdr_blocks_to_rows(uint32_t width, uint32_t height, const float* hdr, float* rows);

/* ---- AVX2 packet baseline (cv_oracle_avx.c) ---- */
/* Read-only view of a scene's arrays (BVH order) and camera. */
typedef struct {
    const float* v;          /* 9 per triangle: v0, v1, v2 */
    const float* uv;         /* 6 per triangle */
    const uint32_t* mat;
    const or_node* nodes;
    const uint32_t* direct;  /* light triangles (BVH indices) */
    const uint32_t* src;     /* source triangle index of each BVH position */
    uint32_t ntris, nnodes, ndirect;
    float cam_pos[3], cam_dpos[3], cam_q[4], cam_dq[4], screen_distance;
    int emission_override;
    float emission[3];
    int approx;
} or_scene_view;
void or_scene_get_view(const or_scene* s, or_scene_view* out);

/* 1 when the library was built for AVX2 + FMA and the host has both. */
int or_avx_available(void);
#define OR_AVX_PACKET 0   /* the reference's packet traversal (bvh.rs:589-657) */
#define OR_AVX_PER_LANE 1 /* each lane traversed alone: the scalar oracle's */
/* The reference's 8-wide packet path (REF_PATCH rng, the scene's approx
 * mode) over whole frames first_frame .. first_frame + n_frames - 1 in the
 * patch order of main.rs:174-214, accumulated into hdr (16x4 block layout).
 * n_queries (optional) receives the number of extend queries (active rays
 * traced).  Returns 0 or a negative status. */
int or_avx_accumulate_frames(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                             uint32_t first_frame, uint32_t n_frames, int mode, int nthreads,
                             uint64_t* n_queries);
/* The packet path over the listed 32x32 patches of one frame (a bounded
 * timing sample); the image is discarded. */
int or_avx_sample_patches(const or_scene* s, uint32_t width, uint32_t height, uint32_t frame,
                          int nthreads, const uint32_t* patch_ids, uint32_t n_ids,
                          uint64_t* n_queries);
/* Bvh::intersect_nearest on 8-ray packets (bvh.rs:589-657), the shape of the
 * reference's bench_intersect_* benches (bvh.rs:681-739): rays (n x 6, n a
 * multiple of 8) taken 8 consecutive rays to a packet, max distance 1e5.
 * tri_id (source index or 0xffffffff) and dist may each be NULL.  Returns 0
 * or a negative status. */
int or_avx_intersect_packets(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                             float* dist, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
