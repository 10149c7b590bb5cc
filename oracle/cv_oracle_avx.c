/*
 * cv_oracle_avx.c -- the reference's 8-wide AVX2 packet path, restated with
 * intrinsics: TEST INFRASTRUCTURE AND CPU BASELINE ONLY (nothing in the
 * product links it).
 *
 * The reference renders 8-ray packets (one per group of a 16x4 block) with
 * Mf32 = __m256 arithmetic.  This file does the same work in the same
 * structure so that its time is a fair host baseline:
 *   - packet BVH traversal with a stack of (tmin, mask, node) entries:
 *     any_masked / should_try_before / is_further_away_than over the 8
 *     lanes (bvh.rs:589-657, aabb.rs:105-183);
 *   - Triangle::intersect per test with the full MIntersection pick
 *     (position, normal, distance, material, texture coordinates;
 *     triangle.rs:59-148);
 *   - continue_path 8-wide: BRDF ray, direct-light ray, one-sample MIS over
 *     the light list, microfacet BRDF with a per-lane gloss table pick
 *     (material.rs:228-513, scene.rs:143-215);
 *   - render_pixels with the all-lanes-emissive break (renderer.rs:413-468),
 *     the 16x4 blocks and the per-patch Rng (renderer.rs:171-206,339-360,
 *     random.rs:43-101), the patch loop of main.rs:174-214.
 * Every floating-point operation is the one the scalar restatement
 * (cv_oracle.c) performs for a lane, in the same order.  With each lane
 * traversed alone (OR_AVX_PER_LANE) the images and query counts are those
 * of or_accumulate_frames in OR_RNG_REF_PATCH mode, bit for bit.  The
 * reference's packet traversal (OR_AVX_PACKET, the timed baseline) differs
 * from a single-ray traversal for rare rays: a box is visited when any lane
 * hits it, so a lane can hit a triangle in a box its own slab test missed,
 * and ties between triangles at the same distance are decided in another
 * visit order.
 */
#include "cv_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#if defined(__AVX2__) && defined(__FMA__)
#include <immintrin.h>

typedef __m256 f8;
typedef __m256i i8;
typedef struct {
    f8 x, y, z;
} v8;

/* ---------------------------------------------------------------- basics */

static inline f8 bc(float x) { return _mm256_set1_ps(x); }
static inline f8 add(f8 a, f8 b) { return _mm256_add_ps(a, b); }
static inline f8 sub(f8 a, f8 b) { return _mm256_sub_ps(a, b); }
static inline f8 mul(f8 a, f8 b) { return _mm256_mul_ps(a, b); }
/* mul_add (a*b + c), mul_sub (a*b - c), neg_mul_add (c - a*b) (simd.rs:81-94) */
static inline f8 fma8(f8 a, f8 b, f8 c) { return _mm256_fmadd_ps(a, b, c); }
static inline f8 fms8(f8 a, f8 b, f8 c) { return _mm256_fmsub_ps(a, b, c); }
static inline f8 fnma8(f8 a, f8 b, f8 c) { return _mm256_fnmadd_ps(a, b, c); }
/* x86 minps / maxps: (a<b)?a:b, (a>b)?a:b (simd.rs:318-326) */
static inline f8 vmin(f8 a, f8 b) { return _mm256_min_ps(a, b); }
static inline f8 vmax(f8 a, f8 b) { return _mm256_max_ps(a, b); }
/* geq = cmpps NLT_UQ (simd.rs:329-333) */
static inline f8 geq(f8 a, f8 b) { return _mm256_cmp_ps(a, b, _CMP_NLT_UQ); }
/* pick: b where the mask's sign bit is set (blendvps, simd.rs:376-378) */
static inline f8 pick(f8 a, f8 b, f8 m) { return _mm256_blendv_ps(a, b, m); }
static inline f8 bor(f8 a, f8 b) { return _mm256_or_ps(a, b); }
static inline f8 band(f8 a, f8 b) { return _mm256_and_ps(a, b); }
static inline f8 bxor(f8 a, f8 b) { return _mm256_xor_ps(a, b); }
/* Mf32 negation is 0 - x (simd.rs:576-588) */
static inline f8 neg0(f8 x) { return sub(_mm256_setzero_ps(), x); }
static inline f8 fabs8(f8 x) { return _mm256_and_ps(x, _mm256_castsi256_ps(_mm256_set1_epi32(0x7fffffff))); }
static inline f8 ibits(i8 x) { return _mm256_castsi256_ps(x); }
static inline i8 fbits(f8 x) { return _mm256_castps_si256(x); }

/* recip_fast / rsqrt: rcpps / rsqrtps, or correctly rounded (PORTABLE) */
static inline f8 rcp8(f8 x, int approx) {
    return approx == OR_APPROX_HOST_X86 ? _mm256_rcp_ps(x) : _mm256_div_ps(bc(1.0f), x);
}
static inline f8 rsqrt8(f8 x, int approx) {
    return approx == OR_APPROX_HOST_X86 ? _mm256_rsqrt_ps(x)
                                        : _mm256_div_ps(bc(1.0f), _mm256_sqrt_ps(x));
}

static inline v8 mk(f8 x, f8 y, f8 z) {
    v8 r = {x, y, z};
    return r;
}
static inline v8 bc3(const float* p) { return mk(bc(p[0]), bc(p[1]), bc(p[2])); }
static inline v8 sub3(v8 a, v8 b) { return mk(sub(a.x, b.x), sub(a.y, b.y), sub(a.z, b.z)); }
static inline v8 scale3(v8 a, f8 s) { return mk(mul(a.x, s), mul(a.y, s), mul(a.z, s)); }
static inline v8 mulc3(v8 a, v8 b) { return mk(mul(a.x, b.x), mul(a.y, b.y), mul(a.z, b.z)); }
static inline v8 neg3(v8 a) { return mk(neg0(a.x), neg0(a.y), neg0(a.z)); }
/* dot_naive (vector3.rs:201-217) and cross_fma (vector3.rs:185-199) */
static inline f8 dot3(v8 a, v8 b) { return add(add(mul(a.x, b.x), mul(a.y, b.y)), mul(a.z, b.z)); }
static inline v8 cross3(v8 a, v8 b) {
    return mk(fms8(a.y, b.z, mul(a.z, b.y)), fms8(a.z, b.x, mul(a.x, b.z)),
              fms8(a.x, b.y, mul(a.y, b.x)));
}
static inline v8 muladd3(v8 a, f8 f, v8 o) {
    return mk(fma8(a.x, f, o.x), fma8(a.y, f, o.y), fma8(a.z, f, o.z));
}
static inline v8 negmuladd3(v8 a, f8 f, v8 o) {
    return mk(fnma8(a.x, f, o.x), fnma8(a.y, f, o.y), fnma8(a.z, f, o.z));
}
static inline v8 normalized3(v8 a, int approx) { return scale3(a, rsqrt8(dot3(a, a), approx)); }
static inline v8 pick3(v8 a, v8 b, f8 m) {
    return mk(pick(a.x, b.x, m), pick(a.y, b.y, m), pick(a.z, b.z, m));
}

/* ------------------------------------------------------------------- rng */

/* Rng::next + the Mu64 -> 8 x u32 view (random.rs:73-101) */
typedef struct {
    uint64_t st[4];
} rng_ref;

static void rng_seed(rng_ref* r, uint32_t x, uint32_t y, uint32_t i) {
    static const uint64_t primes[4] = {14491630826648200009ull, 13149596372461506851ull,
                                       6119410235796056053ull, 14990141545859273719ull};
    uint64_t seed = (uint64_t)x * 12276630456901467871ull + (uint64_t)y * 7661526868048087387ull +
                    (uint64_t)i * 2268244495640532043ull;
    seed = seed + seed % 9358246936573323101ull;
    for (int k = 0; k < 4; k++) r->st[k] = seed * primes[k];
}
static inline i8 rng_next(rng_ref* r) {
    static const uint64_t step[4] = {3ull * 1073243692214514217ull, 5ull * 3335100457702756523ull,
                                     7ull * 8789056573444181ull, 11ull * 781436371140792079ull};
    i8 out = _mm256_loadu_si256((const __m256i*)r->st);
    for (int k = 0; k < 4; k++) r->st[k] *= step[k];
    return out;
}
/* random.rs:104-136 */
static inline f8 as_unit(i8 u) {
    return fma8(_mm256_cvtepi32_ps(u), bc(0.5f / -2147483648.0f), bc(0.5f));
}
static inline f8 as_biunit(i8 u) { return mul(_mm256_cvtepi32_ps(u), bc(1.0f / -2147483648.0f)); }
static inline f8 as_angle(i8 u) {
    return mul(_mm256_cvtepi32_ps(u), bc(3.14159274f / -2147483648.0f));
}

/* sin / cos polynomials (simd.rs:190-233) */
static inline f8 sin_poly(f8 x) {
    f8 x2 = mul(x, x), x3 = mul(x, x2), x5 = mul(x3, x2);
    return fma8(x5, bc(0.00533738081738135917f), fma8(x3, bc(-0.152175915294f), mul(x, bc(0.982006669405f))));
}
static inline f8 cos_poly(f8 x) {
    f8 x2 = mul(x, x), x4 = mul(x2, x2), x6 = mul(x2, x4);
    return fma8(x6, bc(-0.000966231179636657107f),
                fma8(x4, bc(0.0392596924214f), fma8(x2, bc(-0.496000299455f), bc(1.0f))));
}

/* -------------------------------------------------------------- geometry */

typedef struct {
    v8 o, d;
    f8 active; /* sign bit 1 = inactive */
} ray8;

typedef struct {
    v8 pos, normal;
    f8 dist, mat, tu, tv, tri;
} isect8;

typedef struct {
    f8 tmin, mask;
    uint32_t node;
} entry;

typedef struct {
    const or_scene_view* s;
    int approx;
    int per_lane; /* OR_AVX_PER_LANE: each lane traversed as its own packet */
    uint64_t queries;
    int status;
} ctx;

/* Scene::intersect_nearest's far-away sky hit (scene.rs:220-230) */
static inline isect8 far_away(const ray8* r) {
    isect8 is;
    is.pos = muladd3(r->d, bc(1.0e5f), r->o);
    is.normal = r->d;
    is.dist = bc(1.0e5f);
    is.mat = ibits(_mm256_set1_epi32((int)0xC0000000u));
    is.tu = is.tv = _mm256_setzero_ps();
    is.tri = ibits(_mm256_set1_epi32(-1));
    return is;
}

/* Aabb::intersect (aabb.rs:105-141) */
static inline void aabb8(const or_node* nd, const v8* o, const v8* inv, f8* tmin, f8* tmax,
                         f8* mask) {
    f8 tx1 = mul(sub(bc(nd->min[0]), o->x), inv->x), tx2 = mul(sub(bc(nd->max[0]), o->x), inv->x);
    f8 ty1 = mul(sub(bc(nd->min[1]), o->y), inv->y), ty2 = mul(sub(bc(nd->max[1]), o->y), inv->y);
    f8 tz1 = mul(sub(bc(nd->min[2]), o->z), inv->z), tz2 = mul(sub(bc(nd->max[2]), o->z), inv->z);
    f8 txmin = vmin(tx1, tx2), txmax = vmax(tx1, tx2);
    f8 tymin = vmin(ty1, ty2), tymax = vmax(ty1, ty2);
    f8 tzmin = vmin(tz1, tz2), tzmax = vmax(tz1, tz2);
    *tmin = vmax(txmin, vmax(tymin, tzmin));
    *tmax = vmin(txmax, vmin(tymax, tzmax));
    *mask = geq(*tmax, *tmin);
}

/* any_masked (aabb.rs:146-158): some active lane hits with tmax sign 0 */
static inline int any_masked(f8 tmax, f8 mask, f8 active) {
    f8 m = band(bxor(mask, active), mask);
    return !_mm256_testc_ps(tmax, m);
}

/* Triangle::intersect with the MIntersection pick (triangle.rs:59-148) */
static inline void tri8(const ctx* c, uint32_t k, const ray8* r, isect8* is) {
    const or_scene_view* s = c->s;
    const float* tv = s->v + 9 * (size_t)k;
    v8 v0 = bc3(tv), v1 = bc3(tv + 3), v2 = bc3(tv + 6);
    v8 e1 = sub3(v0, v2);
    v8 e2 = sub3(v1, v0);
    v8 n = cross3(e1, e2);
    v8 fr = sub3(v0, r->o);
    f8 denom = _mm256_div_ps(bc(1.0f), dot3(r->d, n));
    f8 t = mul(dot3(fr, n), denom);
    v8 cr = cross3(r->d, fr);
    f8 u = mul(dot3(cr, e2), denom);
    f8 v = mul(dot3(cr, e1), denom);
    f8 w = sub(sub(bc(1.0f), u), v);
    f8 mask = bor(bor(bor(t, u), bor(v, w)), bor(r->active, geq(t, is->dist)));
    /* the reference computes every field of the new intersection per test
     * and blends, whether or not a lane takes it */
    const float* uv = s->uv + 6 * (size_t)k;
    v8 pos = muladd3(r->d, t, r->o);
    v8 nrm = normalized3(n, c->approx);
    f8 tu = fma8(bc(uv[0]), w, fma8(bc(uv[2]), v, mul(bc(uv[4]), u)));
    f8 tvv = fma8(bc(uv[1]), w, fma8(bc(uv[3]), v, mul(bc(uv[5]), u)));
    /* new where the mask's sign bit is 0 */
    is->pos = pick3(pos, is->pos, mask);
    is->normal = pick3(nrm, is->normal, mask);
    is->dist = pick(t, is->dist, mask);
    is->mat = pick(ibits(_mm256_set1_epi32((int)s->mat[k])), is->mat, mask);
    is->tu = pick(tu, is->tu, mask);
    is->tv = pick(tvv, is->tv, mask);
    is->tri = pick(ibits(_mm256_set1_epi32((int)k)), is->tri, mask);
}

/* Bvh::intersect_nearest_impl for one packet (bvh.rs:589-657) */
static void intersect_nearest(ctx* c, const ray8* r, isect8* is) {
    const or_node* nodes = c->s->nodes;
    v8 inv = mk(rcp8(r->d.x, c->approx), rcp8(r->d.y, c->approx), rcp8(r->d.z, c->approx));
    entry stack[256];
    int sp = 0;
#define VISIT_PAIR(base)                                                                 \
    do {                                                                                 \
        const uint32_t b_ = (base);                                                      \
        f8 tmin0, tmax0, m0, tmin1, tmax1, m1;                                           \
        aabb8(&nodes[b_], &r->o, &inv, &tmin0, &tmax0, &m0);                             \
        aabb8(&nodes[b_ + 1], &r->o, &inv, &tmin1, &tmax1, &m1);                         \
        const int h0 = any_masked(tmax0, m0, r->active);                                 \
        const int h1 = any_masked(tmax1, m1, r->active);                                 \
        if (sp + 2 > 256) {                                                              \
            c->status = -20;                                                             \
            return;                                                                      \
        }                                                                                \
        if (_mm256_movemask_ps(sub(tmin0, tmin1)) == 0) { /* should_try_before */        \
            if (h0) stack[sp++] = (entry){tmin0, m0, b_};                                \
            if (h1) stack[sp++] = (entry){tmin1, m1, b_ + 1};                            \
        } else {                                                                         \
            if (h1) stack[sp++] = (entry){tmin1, m1, b_ + 1};                            \
            if (h0) stack[sp++] = (entry){tmin0, m0, b_};                                \
        }                                                                                \
    } while (0)

    VISIT_PAIR(0u);
    while (sp > 0) {
        const entry e = stack[--sp];
        /* is_further_away_than (aabb.rs:161-170) */
        if (_mm256_testc_ps(bor(geq(e.tmin, is->dist), r->active), e.mask)) continue;
        const or_node* nd = &nodes[e.node];
        if (nd->len == 0) {
            VISIT_PAIR(nd->index);
        } else {
            for (uint32_t k = nd->index; k < nd->index + nd->len; k++) tri8(c, k, r, is);
        }
    }
#undef VISIT_PAIR
}

/* The same query with every active lane traversed alone (its ray broadcast
 * to all 8 lanes), which is the single-ray traversal of the scalar oracle. */
static void intersect_lanes(ctx* c, const ray8* r, isect8* is, int act) {
    float ox[8], oy[8], oz[8], dx[8], dy[8], dz[8];
    _mm256_storeu_ps(ox, r->o.x);
    _mm256_storeu_ps(oy, r->o.y);
    _mm256_storeu_ps(oz, r->o.z);
    _mm256_storeu_ps(dx, r->d.x);
    _mm256_storeu_ps(dy, r->d.y);
    _mm256_storeu_ps(dz, r->d.z);
    const i8 lane = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
    for (int l = 0; l < 8; l++) {
        if (!(act >> l & 1)) continue;
        ray8 b;
        b.o = mk(bc(ox[l]), bc(oy[l]), bc(oz[l]));
        b.d = mk(bc(dx[l]), bc(dy[l]), bc(dz[l]));
        b.active = _mm256_setzero_ps();
        isect8 bi = far_away(&b);
        intersect_nearest(c, &b, &bi);
        const f8 m = ibits(_mm256_cmpeq_epi32(lane, _mm256_set1_epi32(l)));
        is->pos = pick3(is->pos, bi.pos, m);
        is->normal = pick3(is->normal, bi.normal, m);
        is->dist = pick(is->dist, bi.dist, m);
        is->mat = pick(is->mat, bi.mat, m);
        is->tu = pick(is->tu, bi.tu, m);
        is->tv = pick(is->tv, bi.tv, m);
        is->tri = pick(is->tri, bi.tri, m);
    }
}

/* -------------------------------------------------------------- shading */

/* MVector3::rotate_hemisphere (vector3.rs:219-251) */
static inline v8 rotate_hemisphere(v8 v, v8 normal, int approx) {
    f8 flip = normal.z;
    v8 n = pick3(normal, neg3(normal), flip);
    f8 rz = rcp8(add(bc(1.0f), n.z), approx);
    f8 c = mul(mul(n.x, n.y), rz);
    f8 x = fms8(v.x, fma8(n.y, mul(n.y, rz), n.z), fms8(v.y, c, mul(v.z, n.x)));
    f8 y = fnma8(v.x, c, fma8(v.y, fma8(n.x, mul(n.x, rz), n.z), mul(v.z, n.y)));
    f8 z = fnma8(v.x, n.x, fnma8(v.y, n.y, mul(v.z, n.z)));
    v8 r = mk(x, y, z);
    return pick3(r, neg3(r), flip);
}

/* sky_intensity (material.rs:205-215) */
static inline v8 sky_intensity(v8 dir) {
    v8 up = mk(_mm256_setzero_ps(), _mm256_setzero_ps(), bc(1.0f));
    f8 d = fma8(dot3(dir, up), bc(0.5f), bc(0.5f));
    v8 rgb = mk(d, mul(d, d), mul(d, mul(d, d)));
    return muladd3(rgb, bc(2.0f), mk(bc(0.5f), bc(0.5f), bc(0.5f)));
}

/* Triangle::intersect_direct folded into pd_direct_sample
 * (triangle.rs:152-185, material.rs:281-320) */
static inline f8 pd_direct(const ctx* c, v8 o, v8 d) {
    const or_scene_view* s = c->s;
    const int approx = c->approx;
    f8 total = _mm256_setzero_ps();
    for (uint32_t k = 0; k < s->ndirect; k++) {
        const float* tv = s->v + 9 * (size_t)s->direct[k];
        v8 v0 = bc3(tv), v1 = bc3(tv + 3), v2 = bc3(tv + 6);
        v8 e1 = sub3(v0, v2);
        v8 e2 = sub3(v1, v0);
        v8 n = cross3(e1, e2);
        f8 rnorm = rsqrt8(dot3(n, n), approx);
        f8 area = mul(bc(0.5f), rcp8(rnorm, approx));
        v8 fr = sub3(v0, o);
        f8 denom = rcp8(dot3(d, n), approx);
        f8 t = mul(dot3(fr, n), denom);
        v8 cr = cross3(d, fr);
        f8 u = mul(dot3(cr, e2), denom);
        f8 v = mul(dot3(cr, e1), denom);
        f8 mask = bor(bor(u, v), bor(t, geq(add(u, v), bc(1.0f))));
        v8 normal = scale3(n, rnorm);
        f8 dist_sqr = mul(t, t);
        f8 dot_em = add(fabs8(dot3(normal, d)), bc(0.0001f));
        f8 pd = mul(dist_sqr, rcp8(mul(area, dot_em), approx));
        total = pick(add(total, pd), total, mask);
    }
    return mul(total, bc(1.0f / (float)s->ndirect));
}

/* microfacet_brdf with the normal distribution table and Schlick's
 * approximation (material.rs:415-513) */
static inline v8 microfacet_brdf(f8 mat, v8 in_dir, v8 out_dir, v8 normal, int ignore_fresnel,
                                 int approx) {
    const float pi = 3.14159274f;
    i8 mi = fbits(mat);
    const f8 k255 = bc(1.0f / 255.0f);
    const i8 ff = _mm256_set1_epi32(0xff);
    v8 color = mk(mul(_mm256_cvtepi32_ps(_mm256_and_si256(mi, ff)), k255),
                  mul(_mm256_cvtepi32_ps(_mm256_and_si256(_mm256_srai_epi32(mi, 8), ff)), k255),
                  mul(_mm256_cvtepi32_ps(_mm256_and_si256(_mm256_srai_epi32(mi, 16), ff)), k255));
    i8 gloss = _mm256_and_si256(_mm256_srai_epi32(mi, 26), _mm256_set1_epi32(7));
    v8 h = normalized3(sub3(in_dir, out_dir), approx);
    f8 c1 = dot3(h, normal);
    f8 c2 = mul(c1, c1), c4 = mul(c2, c2), c8 = mul(c4, c4), c16 = mul(c8, c8);
    /* the reference picks one table entry per lane; indices 5..7 use entry 5 */
    f8 dist = mul(c16, bc(8.5f / pi));
    dist = pick(dist, mul(c8, bc(4.5f / pi)), ibits(_mm256_cmpeq_epi32(gloss, _mm256_set1_epi32(4))));
    dist = pick(dist, mul(c4, bc(2.5f / pi)), ibits(_mm256_cmpeq_epi32(gloss, _mm256_set1_epi32(3))));
    dist = pick(dist, mul(c2, bc(1.5f / pi)), ibits(_mm256_cmpeq_epi32(gloss, _mm256_set1_epi32(2))));
    dist = pick(dist, mul(vmax(c1, _mm256_setzero_ps()), bc(1.0f / pi)),
                ibits(_mm256_cmpeq_epi32(gloss, _mm256_set1_epi32(1))));
    dist = pick(dist, bc(0.5f / pi), ibits(_mm256_cmpeq_epi32(gloss, _mm256_setzero_si256())));
    /* microfacet_fresnel (material.rs:453-469) */
    v8 r1 = sub3(mk(bc(1.0f), bc(1.0f), bc(1.0f)), color);
    f8 ct = sub(bc(1.0f), fabs8(dot3(h, in_dir)));
    f8 ct2 = mul(ct, ct);
    f8 ct3 = mul(ct2, ct);
    f8 ct5 = fabs8(mul(ct2, ct3));
    v8 fc = muladd3(r1, ct5, color);
    if (ignore_fresnel) {
        f8 has_tex = ibits(_mm256_or_si256(_mm256_slli_epi32(mi, 6), _mm256_slli_epi32(mi, 7)));
        fc = pick3(fc, mk(bc(1.0f), bc(1.0f), bc(1.0f)), has_tex);
    }
    return scale3(fc, dist);
}

/* continue_path with the BRDF and direct-light candidates and one-sample
 * MIS (material.rs:228-411); the draws come in the reference's order. */
static inline void continue_path(const ctx* c, ray8* ray, const isect8* is, const i8 dr[6],
                                 int ignore_fresnel, v8* color_mod) {
    const or_scene_view* s = c->s;
    const int approx = c->approx;
    f8 active = bor(ray->active, is->mat);
    /* continue_path_brdf (material.rs:228-242), sample_hemisphere_vector
     * (random.rs:140-159) */
    f8 phi = as_angle(dr[0]), r_sqr = as_unit(dr[1]);
    f8 rr_ = _mm256_sqrt_ps(r_sqr);
    v8 dir_z = mk(mul(sin_poly(phi), rr_), mul(cos_poly(phi), rr_),
                  _mm256_sqrt_ps(sub(bc(1.0f), r_sqr)));
    v8 dir_b = rotate_hemisphere(dir_z, is->normal, approx);
    v8 org_b = muladd3(dir_b, bc(1.0e-4f), is->pos);
    /* Scene::get_direct_sample: a serial per-lane gather (scene.rs:143-202) */
    uint32_t bits[8];
    float gv[9][8];
    _mm256_storeu_si256((__m256i*)bits, dr[2]);
    for (int l = 0; l < 8; l++) {
        const float* tv = s->v + 9 * (size_t)s->direct[bits[l] >> 29];
        for (int q = 0; q < 9; q++) gv[q][l] = tv[q];
    }
    v8 v0 = mk(_mm256_loadu_ps(gv[0]), _mm256_loadu_ps(gv[1]), _mm256_loadu_ps(gv[2]));
    v8 v1 = mk(_mm256_loadu_ps(gv[3]), _mm256_loadu_ps(gv[4]), _mm256_loadu_ps(gv[5]));
    v8 v2 = mk(_mm256_loadu_ps(gv[6]), _mm256_loadu_ps(gv[7]), _mm256_loadu_ps(gv[8]));
    v8 e1 = sub3(v0, v2);
    v8 e2 = sub3(v1, v0);
    f8 u = as_unit(dr[3]), v = as_unit(dr[4]);
    f8 fold = sub(sub(bc(1.0f), u), v);
    f8 uu = pick(u, sub(bc(1.0f), u), fold);
    f8 vv = pick(v, sub(bc(1.0f), v), fold);
    v8 p = muladd3(e2, vv, negmuladd3(e1, uu, v0));
    /* continue_path_direct_sample (material.rs:253-265) */
    v8 dir_d = normalized3(sub3(p, is->pos), approx);
    v8 org_d = muladd3(dir_d, bc(1.0e-4f), is->pos);
    f8 rr = fma8(is->normal.z, bc(0.8f), as_biunit(dr[5]));
    v8 n_org = pick3(org_b, org_d, rr);
    v8 n_dir = pick3(dir_b, dir_d, rr);
    f8 pd_b = mul(vmax(dot3(is->normal, n_dir), _mm256_setzero_ps()), bc(1.0f / 3.14159274f));
    f8 pd_d = pd_direct(c, n_org, n_dir);
    f8 weight_denom = add(add(pd_b, pd_d), bc(0.01f));
    f8 p_brdf = fma8(is->normal.z, bc(0.4f), bc(0.5f));
    f8 p_direct = fnma8(is->normal.z, bc(0.4f), bc(0.5f));
    f8 pp = pick(p_brdf, p_direct, rr);
    f8 modulation = rcp8(mul(weight_denom, pp), approx);
    f8 cos_theta = vmax(dot3(is->normal, n_dir), _mm256_setzero_ps());
    v8 brdf = microfacet_brdf(is->mat, n_dir, ray->d, is->normal, ignore_fresnel, approx);
    v8 cm = scale3(brdf, mul(modulation, cos_theta));
    cm = mk(vmin(cm.x, bc(2.0f)), vmin(cm.y, bc(2.0f)), vmin(cm.z, bc(2.0f)));
    ray->o = pick3(n_org, ray->o, active);
    ray->d = pick3(n_dir, ray->d, active);
    ray->active = active;
    *color_mod = pick3(cm, mk(bc(1.0f), bc(1.0f), bc(1.0f)), active);
}

/* ------------------------------------------------------------- renderer */

/* Camera::get_ray (scene.rs:81-99, quaternion.rs:62-153) */
static inline ray8 camera_ray(const ctx* c, f8 x, f8 y, f8 t) {
    const or_scene_view* s = c->s;
    ray8 r;
    r.o = muladd3(bc3(s->cam_dpos), t, bc3(s->cam_pos));
    f8 qa = fma8(bc(s->cam_dq[0]), t, bc(s->cam_q[0]));
    f8 qb = fma8(bc(s->cam_dq[1]), t, bc(s->cam_q[1]));
    f8 qc = fma8(bc(s->cam_dq[2]), t, bc(s->cam_q[2]));
    f8 qd = fma8(bc(s->cam_dq[3]), t, bc(s->cam_q[3]));
    f8 ns = add(fma8(qa, qa, mul(qb, qb)), fma8(qc, qc, mul(qd, qd)));
    f8 rn = _mm256_div_ps(bc(1.0f), _mm256_sqrt_ps(ns));
    qa = mul(qa, rn);
    qb = mul(qb, rn);
    qc = mul(qc, rn);
    qd = mul(qd, rn);
    v8 src = normalized3(mk(x, y, bc(-s->screen_distance)), c->approx);
    f8 vx = src.x, vy = src.y, vz = src.z;
    f8 pa = add(add(mul(qb, vx), mul(qc, vy)), mul(qd, vz));
    f8 pb = add(add(add(sub(mul(qb, vx), mul(qc, vy)), mul(qd, vz)), mul(sub(qa, qb), vx)),
                mul(sub(qc, qd), add(vy, vz)));
    f8 pc = sub(add(mul(qd, vx), mul(qa, vy)), mul(qb, vz));
    f8 pd = add(sub(sub(sub(mul(qd, vx), mul(qa, vy)), mul(qb, vz)), mul(add(qc, qd), vx)),
                mul(add(qa, qb), add(vy, vz)));
    f8 rb = sub(add(sub(sub(sub(mul(add(pa, pb), add(qa, qb)), mul(sub(pc, pd), add(qc, qd))),
                            mul(pa, qa)),
                        mul(pb, qb)),
                    mul(pc, qc)),
                mul(pd, qd));
    f8 rc = add(add(sub(mul(pc, qa), mul(pd, qb)), mul(pa, qc)), mul(pb, qd));
    f8 rd = add(sub(sub(sub(add(mul(add(pc, pd), add(qa, qb)), mul(sub(pa, pb), add(qc, qd))),
                            mul(pc, qa)),
                        mul(pd, qb)),
                    mul(pa, qc)),
                mul(pb, qd));
    r.d = mk(rb, rc, rd);
    r.active = _mm256_setzero_ps();
    return r;
}

/* render_pixels for one 8-lane group (renderer.rs:413-468) */
static v8 render_pixels(ctx* c, f8 x, f8 y, rng_ref* rng) {
    const or_scene_view* s = c->s;
    f8 t = as_unit(rng_next(rng));
    ray8 ray = camera_ray(c, x, y, t);
    v8 color = mk(bc(1.0f), bc(1.0f), bc(1.0f));
    f8 hit_emissive = _mm256_setzero_ps();
    f8 term_tri = _mm256_setzero_ps();
    for (int i = 0; i < 5; i++) {
        isect8 is = far_away(&ray);
        const int act = ~_mm256_movemask_ps(ray.active) & 0xff;
        if (act) {
            c->queries += (uint64_t)__builtin_popcount(act);
            if (c->per_lane)
                intersect_lanes(c, &ray, &is, act);
            else
                intersect_nearest(c, &ray, &is);
            /* an active lane that ends on an emissive triangle (not the sky) */
            f8 upd = _mm256_andnot_ps(ray.active, is.mat);
            f8 is_tri = ibits(_mm256_xor_si256(
                _mm256_cmpeq_epi32(fbits(is.tri), _mm256_set1_epi32(-1)), _mm256_set1_epi32(-1)));
            term_tri = pick(term_tri, is_tri, upd);
        }
        hit_emissive = is.mat;
        if (_mm256_movemask_ps(is.mat) == 0xff) break;
        i8 dr[6];
        for (int k = 0; k < 6; k++) dr[k] = rng_next(rng);
        v8 cm;
        continue_path(c, &ray, &is, dr, i == 0, &cm);
        color = mulc3(color, cm);
    }
    v8 em = sky_intensity(ray.d);
    if (s->emission_override) em = pick3(em, bc3(s->emission), term_tri);
    v8 col = mulc3(color, em);
    return pick3(mk(_mm256_setzero_ps(), _mm256_setzero_ps(), _mm256_setzero_ps()), col,
                 hit_emissive);
}

/* One patch (accumulate_patch_f32, renderer.rs:339-360): 16x4 blocks with
 * the per-patch Rng; get_pixel_coords_16x4 + render_block_16x4
 * (renderer.rs:171-206,292-300). */
static void render_patch(ctx* c, uint32_t width, uint32_t height, float* hdr, uint32_t pw,
                         uint32_t x, uint32_t y, uint32_t frame) {
    rng_ref rng;
    rng_seed(&rng, x, y, frame);
    const float scale = 2.0f / (float)width;
    const float sm[4] = {2.0f * scale, 4.0f * scale, 8.0f * scale, 12.0f * scale};
    const f8 off_x = _mm256_setr_ps(0, 1, 2, 3, 0, 1, 2, 3);
    const f8 off_y = _mm256_setr_ps(0, 0, 0, 0, 1, 1, 1, 1);
    for (uint32_t i = 0; i < pw / 16; i++)
        for (uint32_t j = 0; j < pw / 4; j++) {
            const uint32_t xb = x + i * 16, yb = y + j * 4;
            const f8 bx = mul(bc(scale), add(off_x, bc((float)xb - (float)width * 0.5f)));
            const f8 by = mul(bc(scale), add(off_y, bc((float)yb - (float)height * 0.5f)));
            f8 xs[8], ys[8];
            for (int g = 0; g < 8; g++) {
                xs[g] = g / 2 == 0 ? bx : add(bx, bc(sm[g / 2]));
                ys[g] = g % 2 == 0 ? by : add(by, bc(sm[0]));
            }
            for (int g = 0; g < 8; g++) xs[g] = fma8(as_unit(rng_next(&rng)), bc(scale), xs[g]);
            for (int g = 0; g < 8; g++) ys[g] = fma8(as_unit(rng_next(&rng)), bc(scale), ys[g]);
            float* blk =
                hdr ? hdr + ((size_t)(y / 4 + j) * (width / 16) + (x / 16 + i)) * 192 : NULL;
            for (int g = 0; g < 8; g++) {
                v8 col = render_pixels(c, xs[g], ys[g], &rng);
                if (blk) {
                    float* o = blk + g * 24;
                    _mm256_storeu_ps(o, add(_mm256_loadu_ps(o), col.x));
                    _mm256_storeu_ps(o + 8, add(_mm256_loadu_ps(o + 8), col.y));
                    _mm256_storeu_ps(o + 16, add(_mm256_loadu_ps(o + 16), col.z));
                }
            }
        }
}

/* ------------------------------------------------------------ patch loop */

typedef struct {
    or_scene_view view;
    int mode;
    uint32_t width, height, first_frame, n_frames;
    float* hdr;
    const uint32_t* ids; /* NULL: every patch */
    uint32_t n;
    volatile uint32_t next;
    pthread_mutex_t lock;
    uint64_t queries;
    int status;
} avx_job;

static void* avx_worker(void* arg) {
    avx_job* j = (avx_job*)arg;
    ctx c = {&j->view, j->view.approx, j->mode == OR_AVX_PER_LANE, 0, 0};
    const uint32_t npy = (j->height + 31) / 32;
    for (;;) {
        pthread_mutex_lock(&j->lock);
        uint32_t k = j->next++;
        pthread_mutex_unlock(&j->lock);
        if (k >= j->n) break;
        const uint32_t p = j->ids ? j->ids[k] : k;
        /* columns outer, rows inner (main.rs:180-181); edge patches that are
         * narrower than 32 are tiled with 16-wide patches */
        const uint32_t x = (p / npy) * 32, y = (p % npy) * 32;
        const uint32_t pw = (x + 32 <= j->width && y + 32 <= j->height) ? 32 : 16;
        const uint32_t xe = x + 32 <= j->width ? x + 32 : j->width;
        const uint32_t ye = y + 32 <= j->height ? y + 32 : j->height;
        for (uint32_t f = 0; f < j->n_frames; f++)
            for (uint32_t yy = y; yy < ye; yy += pw)
                for (uint32_t xx = x; xx < xe; xx += pw)
                    render_patch(&c, j->width, j->height, j->hdr, pw, xx, yy, j->first_frame + f);
    }
    pthread_mutex_lock(&j->lock);
    j->queries += c.queries;
    if (c.status) j->status = c.status;
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

static int avx_run(avx_job* j, int nthreads, uint64_t* n_queries) {
    if ((j->width & 15) || (j->height & 15)) return -10;
    if (j->view.ndirect < 8) return -12;
    if (!or_avx_available()) return -30;
    pthread_mutex_init(&j->lock, NULL);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, avx_worker, j);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j->lock);
    if (n_queries) *n_queries = j->queries;
    return j->status;
}

int or_avx_available(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
}

int or_avx_accumulate_frames(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                             uint32_t first_frame, uint32_t n_frames, int mode, int nthreads,
                             uint64_t* n_queries) {
    if (mode != OR_AVX_PACKET && mode != OR_AVX_PER_LANE) return -13;
    avx_job j;
    memset(&j, 0, sizeof(j));
    or_scene_get_view(s, &j.view);
    j.mode = mode;
    j.width = width;
    j.height = height;
    j.first_frame = first_frame;
    j.n_frames = n_frames;
    j.hdr = hdr;
    j.n = ((width + 31) / 32) * ((height + 31) / 32);
    return avx_run(&j, nthreads, n_queries);
}

int or_avx_sample_patches(const or_scene* s, uint32_t width, uint32_t height, uint32_t frame,
                          int nthreads, const uint32_t* patch_ids, uint32_t n_ids,
                          uint64_t* n_queries) {
    avx_job j;
    memset(&j, 0, sizeof(j));
    or_scene_get_view(s, &j.view);
    j.width = width;
    j.height = height;
    j.first_frame = frame;
    j.n_frames = 1;
    j.ids = patch_ids;
    j.n = n_ids;
    const uint32_t npatch = ((width + 31) / 32) * ((height + 31) / 32);
    for (uint32_t k = 0; k < n_ids; k++)
        if (patch_ids[k] >= npatch) return -11;
    return avx_run(&j, nthreads, n_queries);
}

/* ------------------------------------------------ batched packet queries */

typedef struct {
    const or_scene_view* view;
    const float* rays;
    size_t npackets;
    uint32_t* tri_id;
    float* dist;
    volatile size_t next;
    pthread_mutex_t lock;
    int status;
} pk_job;

static void* pk_worker(void* arg) {
    pk_job* j = (pk_job*)arg;
    ctx c = {j->view, j->view->approx, 0, 0, 0};
    const size_t chunk = 256;
    for (;;) {
        pthread_mutex_lock(&j->lock);
        size_t p0 = j->next;
        j->next += chunk;
        pthread_mutex_unlock(&j->lock);
        if (p0 >= j->npackets) break;
        size_t p1 = p0 + chunk < j->npackets ? p0 + chunk : j->npackets;
        for (size_t p = p0; p < p1 && !c.status; p++) {
            float l[6][8];
            const float* r = j->rays + 48 * p;
            for (int k = 0; k < 8; k++)
                for (int q = 0; q < 6; q++) l[q][k] = r[6 * k + q];
            ray8 ray;
            ray.o = mk(_mm256_loadu_ps(l[0]), _mm256_loadu_ps(l[1]), _mm256_loadu_ps(l[2]));
            ray.d = mk(_mm256_loadu_ps(l[3]), _mm256_loadu_ps(l[4]), _mm256_loadu_ps(l[5]));
            ray.active = _mm256_setzero_ps();
            isect8 is = far_away(&ray); /* MIntersection::with_max_distance(1e5) */
            intersect_nearest(&c, &ray, &is);
            uint32_t tri[8];
            _mm256_storeu_si256((i8*)tri, fbits(is.tri));
            if (j->tri_id)
                for (int k = 0; k < 8; k++)
                    j->tri_id[8 * p + k] = tri[k] == 0xffffffffu ? tri[k] : j->view->src[tri[k]];
            if (j->dist) _mm256_storeu_ps(j->dist + 8 * p, is.dist);
        }
    }
    if (c.status) {
        pthread_mutex_lock(&j->lock);
        j->status = c.status;
        pthread_mutex_unlock(&j->lock);
    }
    return NULL;
}

int or_avx_intersect_packets(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                             float* dist, int nthreads) {
    if (n & 7) return -14;
    if (!or_avx_available()) return -30;
    or_scene_view view;
    or_scene_get_view(s, &view);
    pk_job j;
    memset(&j, 0, sizeof(j));
    j.view = &view;
    j.rays = rays;
    j.npackets = n / 8;
    j.tri_id = tri_id;
    j.dist = dist;
    pthread_mutex_init(&j.lock, NULL);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, pk_worker, &j);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j.lock);
    return j.status;
}

#else /* no AVX2 / FMA at build time */

int or_avx_available(void) { return 0; }
int or_avx_accumulate_frames(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                             uint32_t first_frame, uint32_t n_frames, int mode, int nthreads,
                             uint64_t* n_queries) {
    (void)s; (void)width; (void)height; (void)hdr; (void)first_frame; (void)n_frames;
    (void)mode; (void)nthreads; (void)n_queries;
    return -30;
}
int or_avx_sample_patches(const or_scene* s, uint32_t width, uint32_t height, uint32_t frame,
                          int nthreads, const uint32_t* patch_ids, uint32_t n_ids,
                          uint64_t* n_queries) {
    (void)s; (void)width; (void)height; (void)frame; (void)nthreads; (void)patch_ids;
    (void)n_ids; (void)n_queries;
    return -30;
}

int or_avx_intersect_packets(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                             float* dist, int nthreads) {
    (void)s; (void)rays; (void)n; (void)tri_id; (void)dist; (void)nthreads;
    return -30;
}

#endif
