/*
 * cv_oracle.c -- scalar CPU restatement of Convector's per-ray path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the oracle that tests/, the smoke check
 * and the CPU-baseline leg of bench.py compare against.  It is not part of
 * the product and the product never links it.
 *
 * Reference: ruuda/convector (Rust, AVX).  Each function below restates the
 * cited reference function for ONE lane of its 8-lane Mf32 code, keeping the
 * exact operation order:
 *   - mul_add(a,b,c) = fmaf(a,b,c), mul_sub(a,b,c) = fmaf(a,b,-c),
 *     neg_mul_add(a,b,c) = fmaf(-a,b,c)                  (simd.rs:81-94)
 *   - Mf32 negation is 0 - x, not a sign flip           (simd.rs:576-588)
 *   - min/max are x86 minps/maxps: (a<b)?a:b, (a>b)?a:b  (simd.rs:318-326)
 *   - geq is cmpps NLT_UQ: !(a<b), true when unordered   (simd.rs:329-333)
 *   - masks live in float sign bits; pick(a,b,m) takes b when m's sign bit
 *     is 1 (blendvps)                                     (simd.rs:376-378)
 *   - dot = (x*x' + y*y') + z*z' without FMA             (vector3.rs:201-217)
 *   - cross = FMA form                                   (vector3.rs:185-199)
 * Compile with -ffp-contract=off so that no other product is fused.
 *
 * Parity status of each part is recorded in the header of tests/ and in the
 * design notes; the x86 rcpps/rsqrtps instructions are reproduced only by
 * running them (OR_APPROX_HOST_X86).
 */
#include "cv_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#if defined(__x86_64__) || defined(__i386__)
#include <immintrin.h>
#define OR_HAVE_X86 1
#else
#define OR_HAVE_X86 0
#endif

/* ------------------------------------------------------------------------ */
/* Bit helpers                                                              */
/* ------------------------------------------------------------------------ */

static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t sgn(float f) { return f2u(f) >> 31; }

/* x86 minps / maxps (simd.rs:318-326). */
static inline float xmin(float a, float b) { return a < b ? a : b; }
static inline float xmax(float a, float b) { return a > b ? a : b; }
/* cmpps predicate 21, NLT_UQ (simd.rs:329-333): all ones when !(a < b). */
static inline uint32_t geq(float a, float b) { return (a < b) ? 0u : 0xffffffffu; }
/* blendvps: pick b when the sign bit of m is set (simd.rs:376-378). */
static inline float pickf(float a, float b, uint32_t m) { return (m >> 31) ? b : a; }
/* Mf32 negation = 0 - x (simd.rs:576-588). */
static inline float neg0(float x) { return 0.0f - x; }
static inline float fabs_bits(float x) { return u2f(f2u(x) & 0x7fffffffu); }

float or_x86_min(float a, float b) { return xmin(a, b); }
float or_x86_max(float a, float b) { return xmax(a, b); }
uint32_t or_geq_mask(float a, float b) { return geq(a, b); }

/* recip_fast = rcpps (simd.rs:254-256), rsqrt = rsqrtps (simd.rs:309-311). */
float or_rcp(float x, int approx) {
#if OR_HAVE_X86
    if (approx == OR_APPROX_HOST_X86) return _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x)));
#endif
    (void)approx;
    return 1.0f / x;
}
float or_rsqrt(float x, int approx) {
#if OR_HAVE_X86
    if (approx == OR_APPROX_HOST_X86) return _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x)));
#endif
    (void)approx;
    return 1.0f / sqrtf(x);
}
/* recip_precise (simd.rs:264-267) */
float or_recip_precise(float x, int approx) {
    float x0 = or_rcp(x, approx);
    return fmaf(-x, x0 * x0, x0 + x0);
}

/* cvtps2dq (simd.rs:382-384): round to nearest even; NaN and out of range
 * give the integer indefinite 0x80000000. */
int32_t or_cvt_f32_i32(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return (int32_t)0x80000000u;
    return (int32_t)nearbyintf(x);
}

/* Polynomial sine (simd.rs:214-233) and cosine (simd.rs:190-203). */
float or_sin_poly(float x) {
    const float a = 0.982006669405f;
    const float b = -0.152175915294f;
    const float c = 0.00533738081738135917f;
    float x2 = x * x;
    float x3 = x * x2;
    float x5 = x3 * x2;
    return fmaf(x5, c, fmaf(x3, b, x * a));
}
float or_cos_poly(float x) {
    const float a = -0.496000299455f;
    const float b = 0.0392596924214f;
    const float c = -0.000966231179636657107f;
    float x2 = x * x;
    float x4 = x2 * x2;
    float x6 = x2 * x4;
    return fmaf(x6, c, fmaf(x4, b, fmaf(x2, a, 1.0f)));
}
/* acos (simd.rs:143-171) and acos_fast (simd.rs:104-133); not on the render
 * path, kept for the reference's accuracy tests. */
float or_acos_poly(float x) {
    const float z = 1.57079637f;
    const float a = -0.939115566365855f;
    const float b = 0.9217841528914573f;
    const float c = -1.2845906244690837f;
    const float d = 0.295624144969963174f;
    float x2 = x * x;
    float x3 = x * x2;
    float x4 = x2 * x2;
    float numer = fmaf(x3, b, a * x);
    float denom = fmaf(x4, d, fmaf(x2, c, 1.0f));
    return fmaf(numer, or_rcp(denom, OR_APPROX_HOST_X86), z);
}
float or_acos_fast_poly(float x) {
    const float z = 1.57079637f;
    const float a = -1.010967650783108321215466002435f;
    const float b = 0.7187490034089083954058416974792f;
    const float c = -0.8139678312270743216751354761334f;
    float x2 = x * x;
    float x3 = x * x2;
    float numer = fmaf(x3, b, a * x);
    float denom = fmaf(x2, c, 1.0f);
    return fmaf(numer, or_rcp(denom, OR_APPROX_HOST_X86), z);
}

/* ------------------------------------------------------------------------ */
/* Vectors (vector3.rs)                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    float x, y, z;
} v3;

static inline v3 mk3(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static inline v3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 scale3(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline v3 mulc3(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 neg3(v3 a) { return mk3(neg0(a.x), neg0(a.y), neg0(a.z)); }
/* MVector3::dot = dot_naive (vector3.rs:201-217) */
static inline float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* MVector3::cross = cross_fma (vector3.rs:185-199) */
static inline v3 cross3(v3 a, v3 b) {
    return mk3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
               fmaf(a.x, b.y, -(a.y * b.x)));
}
/* MVector3::mul_add (vector3.rs:254-260): self*factor + other */
static inline v3 muladd3(v3 a, float f, v3 o) {
    return mk3(fmaf(a.x, f, o.x), fmaf(a.y, f, o.y), fmaf(a.z, f, o.z));
}
/* MVector3::neg_mul_add (vector3.rs:263-269) */
static inline v3 negmuladd3(v3 a, float f, v3 o) {
    return mk3(fmaf(-a.x, f, o.x), fmaf(-a.y, f, o.y), fmaf(-a.z, f, o.z));
}
/* MVector3::normalized via rsqrt (vector3.rs:290-306) */
static inline v3 normalized3(v3 a, int approx) {
    float r = or_rsqrt(dot3(a, a), approx);
    return scale3(a, r);
}
static inline v3 pick3(v3 a, v3 b, uint32_t m) {
    return (m >> 31) ? b : a;
}

/* MVector3::rotate_hemisphere (vector3.rs:219-251) */
static v3 rotate_hemisphere(v3 v, v3 normal, int approx) {
    uint32_t flip = f2u(normal.z);
    v3 n = pick3(normal, neg3(normal), flip);
    float rz = or_rcp(1.0f + n.z, approx);
    float c = (n.x * n.y) * rz;
    float x = fmaf(v.x, fmaf(n.y, n.y * rz, n.z), -fmaf(v.y, c, -(v.z * n.x)));
    float y = fmaf(-v.x, c, fmaf(v.y, fmaf(n.x, n.x * rz, n.z), v.z * n.y));
    float z = fmaf(-v.x, n.x, fmaf(-v.y, n.y, v.z * n.z));
    v3 r = mk3(x, y, z);
    return pick3(r, neg3(r), flip);
}

void or_rotate_hemisphere(const float* v, const float* normal, size_t n, int approx, float* out) {
    for (size_t i = 0; i < n; i++) {
        v3 r = rotate_hemisphere(ld3(v + 3 * i), ld3(normal + 3 * i), approx);
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}

/* ------------------------------------------------------------------------ */
/* Quaternions (quaternion.rs:62-153)                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    float a, b, c, d;
} quat;

/* MQuaternion::interpolate (quaternion.rs:62-107) */
static quat quat_interpolate(quat q, quat dq, float t) {
    float a = fmaf(dq.a, t, q.a);
    float b = fmaf(dq.b, t, q.b);
    float c = fmaf(dq.c, t, q.c);
    float d = fmaf(dq.d, t, q.d);
    float ns = fmaf(a, a, b * b) + fmaf(c, c, d * d);
    float rnorm = 1.0f / sqrtf(ns);
    quat r = {a * rnorm, b * rnorm, c * rnorm, d * rnorm};
    return r;
}

/* quaternion::rotate (quaternion.rs:110-153): left-to-right evaluation. */
static v3 quat_rotate(v3 v, quat q) {
    float pa = (q.b * v.x + q.c * v.y) + q.d * v.z;
    float pb = (((q.b * v.x - q.c * v.y) + q.d * v.z) + (q.a - q.b) * v.x) +
               (q.c - q.d) * (v.y + v.z);
    float pc = (q.d * v.x + q.a * v.y) - q.b * v.z;
    float pd = (((q.d * v.x - q.a * v.y) - q.b * v.z) - (q.c + q.d) * v.x) +
               (q.a + q.b) * (v.y + v.z);
    float rb = (((((pa + pb) * (q.a + q.b) - (pc - pd) * (q.c + q.d)) - pa * q.a) - pb * q.b) +
                pc * q.c) -
               pd * q.d;
    float rc = ((pc * q.a - pd * q.b) + pa * q.c) + pb * q.d;
    float rd = (((((pc + pd) * (q.a + q.b) + (pa - pb) * (q.c + q.d)) - pc * q.a) - pd * q.b) -
                pa * q.c) +
               pb * q.d;
    return mk3(rb, rc, rd);
}

void or_quat_rotate(const float* v, const float* q, size_t n, float* out) {
    for (size_t i = 0; i < n; i++) {
        quat qq = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
        v3 r = quat_rotate(ld3(v + 3 * i), qq);
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}

void or_quat_interpolate(const float* q, const float* dq, const float* t, size_t n, float* out) {
    for (size_t i = 0; i < n; i++) {
        quat a = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
        quat b = {dq[4 * i], dq[4 * i + 1], dq[4 * i + 2], dq[4 * i + 3]};
        quat r = quat_interpolate(a, b, t[i]);
        out[4 * i] = r.a;
        out[4 * i + 1] = r.b;
        out[4 * i + 2] = r.c;
        out[4 * i + 3] = r.d;
    }
}

/* ------------------------------------------------------------------------ */
/* Random numbers (random.rs)                                               */
/* ------------------------------------------------------------------------ */

static const uint64_t RNG_PRIMES[4] = {14491630826648200009ull, 13149596372461506851ull,
                                       6119410235796056053ull, 14990141545859273719ull};
static const uint64_t RNG_STEP[4] = {3ull * 1073243692214514217ull, 5ull * 3335100457702756523ull,
                                     7ull * 8789056573444181ull, 11ull * 781436371140792079ull};

/* Rng::with_seed (random.rs:43-70) */
static void rng_ref_seed(uint64_t st[4], uint32_t x, uint32_t y, uint32_t i) {
    uint64_t a = (uint64_t)x * 12276630456901467871ull;
    uint64_t b = (uint64_t)y * 7661526868048087387ull;
    uint64_t c = (uint64_t)i * 2268244495640532043ull;
    uint64_t seed = a + b + c;
    seed = seed + seed % 9358246936573323101ull;
    for (int k = 0; k < 4; k++) st[k] = seed * RNG_PRIMES[k];
}

/* Rng::next + the Mu64 -> 8 x u32 transmute (random.rs:73-101): lane 2k is
 * the low half of u64 lane k, lane 2k+1 the high half (little endian). */
static void rng_ref_next(uint64_t st[4], uint32_t out[8]) {
    for (int k = 0; k < 4; k++) {
        out[2 * k] = (uint32_t)st[k];
        out[2 * k + 1] = (uint32_t)(st[k] >> 32);
        st[k] = st[k] * RNG_STEP[k];
    }
}

/* PIXEL stream: splitmix64 premix of (pixel_index ^ frame<<32), then
 * xorshift64 (13, 7, 17); each draw returns the high 32 bits. */
static uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t rng_pixel_seed(uint32_t pixel_index, uint32_t frame) {
    uint64_t s = splitmix64((uint64_t)pixel_index ^ ((uint64_t)frame << 32));
    return s ? s : 0x9E3779B97F4A7C15ull;
}
static inline uint32_t rng_pixel_next(uint64_t* s) {
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return (uint32_t)(x >> 32);
}

/* Conversions (random.rs:104-136) */
static inline float u32_unit(uint32_t u) {
    return fmaf((float)(int32_t)u, 0.5f / -2147483648.0f, 0.5f);
}
static inline float u32_biunit(uint32_t u) { return (float)(int32_t)u * (1.0f / -2147483648.0f); }
static inline float u32_angle(uint32_t u) {
    return (float)(int32_t)u * (3.14159274f / -2147483648.0f);
}

/* sample_hemisphere_vector (random.rs:140-159), from the angle and unit draws */
static inline v3 hemisphere_from(float phi, float r_sqr) {
    float r = sqrtf(r_sqr);
    float x = or_sin_poly(phi) * r;
    float y = or_cos_poly(phi) * r;
    float z = sqrtf(1.0f - r_sqr);
    return mk3(x, y, z);
}

void or_rng_ref_draws(uint32_t x, uint32_t y, uint32_t i, uint32_t ndraws, uint32_t* out) {
    uint64_t st[4];
    rng_ref_seed(st, x, y, i);
    for (uint32_t k = 0; k < ndraws; k++) rng_ref_next(st, out + 8 * k);
}

void or_rng_ref_floats(uint32_t x, uint32_t y, uint32_t i, uint32_t ndraws, int kind, float* out) {
    uint64_t st[4];
    uint32_t d[8];
    rng_ref_seed(st, x, y, i);
    for (uint32_t k = 0; k < ndraws; k++) {
        rng_ref_next(st, d);
        for (int l = 0; l < 8; l++)
            out[8 * k + l] = kind == 0 ? u32_unit(d[l]) : kind == 1 ? u32_biunit(d[l]) : u32_angle(d[l]);
    }
}

void or_rng_ref_hemisphere(uint32_t x, uint32_t y, uint32_t i, uint32_t n, float* out) {
    uint64_t st[4];
    uint32_t a[8], b[8];
    rng_ref_seed(st, x, y, i);
    for (uint32_t k = 0; k < n; k++) {
        rng_ref_next(st, a);
        rng_ref_next(st, b);
        for (int l = 0; l < 8; l++) {
            v3 v = hemisphere_from(u32_angle(a[l]), u32_unit(b[l]));
            float* o = out + (size_t)(8 * k + l) * 3;
            o[0] = v.x;
            o[1] = v.y;
            o[2] = v.z;
        }
    }
}

void or_rng_pixel_draws(uint32_t pixel_index, uint32_t frame, uint32_t ndraws, uint32_t* out) {
    uint64_t s = rng_pixel_seed(pixel_index, frame);
    for (uint32_t k = 0; k < ndraws; k++) out[k] = rng_pixel_next(&s);
}

/* An 8-lane draw source in either mode. */
typedef struct {
    int mode;
    uint64_t ref[4];
    uint64_t pix[8];
} rng8;

static inline void rng8_next(rng8* r, uint32_t out[8]) {
    if (r->mode == OR_RNG_REF_PATCH) {
        rng_ref_next(r->ref, out);
    } else {
        for (int l = 0; l < 8; l++) out[l] = rng_pixel_next(&r->pix[l]);
    }
}
static inline void rng8_unit(rng8* r, float out[8]) {
    uint32_t d[8];
    rng8_next(r, d);
    for (int l = 0; l < 8; l++) out[l] = u32_unit(d[l]);
}

/* ------------------------------------------------------------------------ */
/* AABB (aabb.rs)                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    float tmin, tmax;
    uint32_t mask;
} aabbx;

/* Aabb::intersect (aabb.rs:105-141) with the reciprocals hoisted per ray. */
static inline aabbx aabb_isect(const float* bmin, const float* bmax, v3 o, v3 inv) {
    float d1x = bmin[0] - o.x, d1y = bmin[1] - o.y, d1z = bmin[2] - o.z;
    float d2x = bmax[0] - o.x, d2y = bmax[1] - o.y, d2z = bmax[2] - o.z;
    float tx1 = d1x * inv.x, tx2 = d2x * inv.x;
    float txmin = xmin(tx1, tx2), txmax = xmax(tx1, tx2);
    float ty1 = d1y * inv.y, ty2 = d2y * inv.y;
    float tymin = xmin(ty1, ty2), tymax = xmax(ty1, ty2);
    float tz1 = d1z * inv.z, tz2 = d2z * inv.z;
    float tzmin = xmin(tz1, tz2), tzmax = xmax(tz1, tz2);
    aabbx r;
    r.tmin = xmax(txmin, xmax(tymin, tzmin));
    r.tmax = xmin(txmax, xmin(tymax, tzmax));
    r.mask = geq(r.tmax, r.tmin);
    return r;
}

void or_aabb_intersect(const float box[6], const float* rays, size_t n, int approx, float* tmin,
                       float* tmax, uint32_t* mask) {
    for (size_t i = 0; i < n; i++) {
        v3 o = ld3(rays + 6 * i), d = ld3(rays + 6 * i + 3);
        v3 inv = mk3(or_rcp(d.x, approx), or_rcp(d.y, approx), or_rcp(d.z, approx));
        aabbx r = aabb_isect(box, box + 3, o, inv);
        tmin[i] = r.tmin;
        tmax[i] = r.tmax;
        mask[i] = r.mask;
    }
}

/* Aabb::area (aabb.rs:97-103) */
static inline float aabb_area(const float* bmin, const float* bmax) {
    float sx = bmax[0] - bmin[0], sy = bmax[1] - bmin[1], sz = bmax[2] - bmin[2];
    float x = sy * sz;
    float y = sz * sx;
    float z = sx * sy;
    return 2.0f * ((x + y) + z);
}
float or_aabb_area(const float box[6]) { return aabb_area(box, box + 3); }

/* ------------------------------------------------------------------------ */
/* Triangles (triangle.rs)                                                  */
/* ------------------------------------------------------------------------ */

/* Core of Triangle::intersect (triangle.rs:59-127).  Returns the pick mask
 * (sign bit 0 => the new intersection replaces the old one). */
static inline uint32_t tri_core(const float* tv, v3 o, v3 d, uint32_t active, float best_t,
                                float* t_out, float* u_out, float* v_out, float* w_out) {
    v3 v0 = ld3(tv), v1 = ld3(tv + 3), v2 = ld3(tv + 6);
    v3 e1 = sub3(v0, v2);
    v3 e2 = sub3(v1, v0);
    v3 n = cross3(e1, e2);
    v3 from_ray = sub3(v0, o);
    float denom = 1.0f / dot3(d, n);
    float t = dot3(from_ray, n) * denom;
    v3 c = cross3(d, from_ray);
    float u = dot3(c, e2) * denom;
    float v = dot3(c, e1) * denom;
    float w = (1.0f - u) - v;
    uint32_t mask_positive = (f2u(t) | f2u(u)) | (f2u(v) | f2u(w));
    uint32_t mask_closer = geq(t, best_t);
    *t_out = t;
    *u_out = u;
    *v_out = v;
    *w_out = w;
    return mask_positive | (active | mask_closer);
}

int or_triangle_intersect(const float tri[9], const float ray[6], float best_t, float* out4) {
    float t, u, v, w;
    uint32_t m = tri_core(tri, ld3(ray), ld3(ray + 3), 0u, best_t, &t, &u, &v, &w);
    if (out4) {
        out4[0] = t;
        out4[1] = u;
        out4[2] = v;
        out4[3] = w;
    }
    return (m >> 31) == 0;
}

typedef struct {
    v3 normal;
    float area, distance;
    uint32_t mask;
} direct_isect;

/* Triangle::intersect_direct (triangle.rs:152-185) */
static inline direct_isect tri_direct(const float* tv, v3 o, v3 d, int approx) {
    v3 v0 = ld3(tv), v1 = ld3(tv + 3), v2 = ld3(tv + 6);
    v3 e1 = sub3(v0, v2);
    v3 e2 = sub3(v1, v0);
    v3 n = cross3(e1, e2);
    float norm_sqr = dot3(n, n);
    float rnorm = or_rsqrt(norm_sqr, approx);
    float area = 0.5f * or_rcp(rnorm, approx);
    v3 from_ray = sub3(v0, o);
    float denom = or_rcp(dot3(d, n), approx);
    float t = dot3(from_ray, n) * denom;
    v3 c = cross3(d, from_ray);
    float u = dot3(c, e2) * denom;
    float v = dot3(c, e1) * denom;
    uint32_t mask_uv = geq(u + v, 1.0f);
    direct_isect r;
    r.mask = (f2u(u) | f2u(v)) | (f2u(t) | mask_uv);
    r.normal = scale3(n, rnorm);
    r.distance = t;
    r.area = area;
    return r;
}

void or_triangle_intersect_direct(const float tri[9], const float ray[6], int approx, float* out6) {
    direct_isect r = tri_direct(tri, ld3(ray), ld3(ray + 3), approx);
    out6[0] = r.normal.x;
    out6[1] = r.normal.y;
    out6[2] = r.normal.z;
    out6[3] = r.area;
    out6[4] = r.distance;
    out6[5] = u2f(r.mask);
}

/* ------------------------------------------------------------------------ */
/* BVH build (bvh.rs:91-545)                                                */
/* ------------------------------------------------------------------------ */

/* Total order on floats used for every min/max of the build (f32::min /
 * f32::max in the reference); differs from the reference only in which of
 * -0.0 / +0.0 is kept, which the reference leaves unspecified. */
static inline uint32_t fkey(float f) {
    uint32_t u = f2u(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static inline float tmin_f(float a, float b) { return fkey(b) < fkey(a) ? b : a; }
static inline float tmax_f(float a, float b) { return fkey(b) > fkey(a) ? b : a; }

typedef struct {
    float bmin[3], bmax[3], c[3];
    uint32_t idx;
} tref;

typedef struct {
    float mn[3], mx[3];
} box6;

static inline void box_empty(box6* b) {
    for (int k = 0; k < 3; k++) {
        b->mn[k] = INFINITY;
        b->mx[k] = -INFINITY;
    }
}
static inline void box_grow(box6* b, const float* mn, const float* mx) {
    for (int k = 0; k < 3; k++) {
        b->mn[k] = tmin_f(b->mn[k], mn[k]);
        b->mx[k] = tmax_f(b->mx[k], mx[k]);
    }
}

/* TreeSurfaceAreaHeuristic (bvh.rs:452-474) with the reference's constants
 * aabb 40, triangle 120, p 0.8 (bvh.rs:500-504). */
static float tree_sah_cost(const box6* parent, const box6* b, uint32_t num_tris) {
    const float p = 0.8f;
    float ac_ap = aabb_area(b->mn, b->mx) / aabb_area(parent->mn, parent->mx);
    float n = (float)num_tris;
    float m = log2f(n);
    float aabb_term = 1.0f + (ac_ap * (2.0f * p - n * powf(p, m))) / (p - (2.0f * p) * p);
    float tri_term = (n * powf(p, m - 1.0f)) * ac_ap;
    return aabb_term * 40.0f + tri_term * 120.0f;
}

float or_tree_sah_cost(const float parent_box[6], const float box[6], uint32_t num_tris) {
    box6 p, b;
    memcpy(p.mn, parent_box, 12);
    memcpy(p.mx, parent_box + 3, 12);
    memcpy(b.mn, box, 12);
    memcpy(b.mx, box + 3, 12);
    return tree_sah_cost(&p, &b, num_tris);
}

typedef struct {
    box6 outer, inner;
    uint32_t begin, count;
    int32_t child[2];
} inode;

typedef struct {
    tref* refs;
    tref* tmp;
    inode* nodes;
    uint32_t nnodes, cap;
    /* scratch */
    uint32_t* bin_of;
    uint32_t* perm;
    float* keys;
    int error;
    uint32_t nbins; /* bins.len() of InterimNode::split (bvh.rs:275): 64, or 16 / 32 */
} builder;

static int32_t new_inode(builder* B, const tref* refs, uint32_t begin, uint32_t count) {
    if (B->nnodes == B->cap) {
        B->cap = B->cap ? 2 * B->cap : 1024;
        B->nodes = (inode*)realloc(B->nodes, sizeof(inode) * B->cap);
    }
    inode* nd = &B->nodes[B->nnodes];
    /* InterimNode::from_triangle_refs (bvh.rs:120-128) */
    box_empty(&nd->outer);
    box_empty(&nd->inner);
    for (uint32_t k = 0; k < count; k++) {
        const tref* r = &refs[begin + k];
        box_grow(&nd->outer, r->bmin, r->bmax);
        box_grow(&nd->inner, r->c, r->c);
    }
    nd->begin = begin;
    nd->count = count;
    nd->child[0] = nd->child[1] = -1;
    return (int32_t)B->nnodes++;
}

/* Stable sort of positions [0,n) of the segment by barycenter coordinate
 * (bin_triangles_uniform, bvh.rs:176-201; sort_by is a stable merge sort and
 * partial_cmp treats -0.0 == +0.0). */
static void stable_sort_by_key(uint32_t* perm, const float* keys, uint32_t n, uint32_t* tmp) {
    /* bottom-up merge sort */
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    uint32_t* src = perm;
    uint32_t* dst = tmp;
    for (uint32_t width = 1; width < n; width *= 2) {
        for (uint32_t lo = 0; lo < n; lo += 2 * width) {
            uint32_t mid = lo + width < n ? lo + width : n;
            uint32_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            uint32_t a = lo, b = mid, k = lo;
            while (a < mid && b < hi) {
                /* take from the right run only when strictly smaller */
                if (keys[src[b]] < keys[src[a]])
                    dst[k++] = src[b++];
                else
                    dst[k++] = src[a++];
            }
            while (a < mid) dst[k++] = src[a++];
            while (b < hi) dst[k++] = src[b++];
        }
        uint32_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != perm) memcpy(perm, src, n * sizeof(uint32_t));
}

typedef struct {
    uint32_t count;
    box6 box;
} binacc;

/* Bins one axis into B->nbins bins (the reference's 64, or 16 / 32: the
 * same steps with bins.len() changed).  Returns 1 when uniform (sorted)
 * binning was used; fills bins[0, nbins) and, for every position of the
 * segment, its bin (bin_of) and for uniform binning the sorted permutation
 * (perm).  bins[nbins, 64) stay empty. */
static int bin_axis(builder* B, const inode* nd, int axis, binacc bins[64], uint32_t* bin_of,
                    uint32_t* perm, uint32_t* perm_tmp) {
    const uint32_t n = nd->count;
    const uint32_t nb = B->nbins;
    const tref* R = B->refs + nd->begin;
    int uniform = n <= nb / 2; /* bvh.rs:145-147: len <= bins/2 */
    if (!uniform) {
        /* bin_triangles (bvh.rs:142-171) */
        float mn = nd->inner.mn[axis];
        float size = nd->inner.mx[axis] - mn;
        uint32_t counts[64] = {0};
        for (uint32_t k = 0; k < n; k++) {
            float coord = R[k].c[axis];
            float f = floorf(((float)nb * (coord - mn)) / size);
            /* `as usize`: NaN becomes out of range on x86-64, then clamped */
            uint32_t idx = (f >= 0.0f && f < (float)nb) ? (uint32_t)f : nb - 1u;
            if (f != f) idx = nb - 1u;
            bin_of[k] = idx;
            counts[idx]++;
        }
        uint32_t mx = 0;
        for (uint32_t b = 0; b < nb; b++) mx = counts[b] > mx ? counts[b] : mx;
        if (mx > n / 8 && n > nb) uniform = 1; /* bvh.rs:159-169 */
    }
    for (int b = 0; b < 64; b++) {
        bins[b].count = 0;
        box_empty(&bins[b].box);
    }
    if (uniform) {
        for (uint32_t k = 0; k < n; k++) B->keys[k] = R[k].c[axis];
        stable_sort_by_key(perm, B->keys, n, perm_tmp);
        uint32_t tpb = (n + nb - 1) / nb;
        if (tpb == 0) tpb = 1;
        for (uint32_t s = 0; s < n; s++) {
            uint32_t b = s / tpb;
            bin_of[perm[s]] = b;
        }
    }
    for (uint32_t k = 0; k < n; k++) {
        binacc* bb = &bins[bin_of[k]];
        bb->count++;
        box_grow(&bb->box, R[k].bmin, R[k].bmax);
    }
    return uniform;
}

/* find_cheapest_split (bvh.rs:218-261).  Returns the split bin. */
static int cheapest_split(const inode* nd, const binacc bins[64], float* cost_out,
                          uint32_t* left_count) {
    int first = -1, last = -1;
    for (int b = 0; b < 64; b++)
        if (bins[b].count) {
            if (first < 0) first = b;
            last = b;
        }
    first += 1;
    last += 1;
    box6 pre[65], suf[65];
    uint32_t precnt[65];
    box_empty(&pre[0]);
    precnt[0] = 0;
    for (int b = 0; b < 64; b++) {
        pre[b + 1] = pre[b];
        precnt[b + 1] = precnt[b] + bins[b].count;
        if (bins[b].count) box_grow(&pre[b + 1], bins[b].box.mn, bins[b].box.mx);
    }
    box_empty(&suf[64]);
    for (int b = 63; b >= 0; b--) {
        suf[b] = suf[b + 1];
        if (bins[b].count) box_grow(&suf[b], bins[b].box.mn, bins[b].box.mx);
    }
    int best = -1;
    float best_cost = 0.0f;
    uint32_t total = precnt[64];
    for (int i = first; i < last; i++) {
        float lc = tree_sah_cost(&nd->outer, &pre[i], precnt[i]);
        float rc = tree_sah_cost(&nd->outer, &suf[i], total - precnt[i]);
        float cost = lc + rc;
        if (best < 0 || cost < best_cost) {
            best_cost = cost;
            best = i;
        }
    }
    *cost_out = best_cost;
    *left_count = best >= 0 ? precnt[best] : 0;
    return best;
}

/* InterimNode::split (bvh.rs:265-322).  Returns 1 when split. */
static int split_node(builder* B, int32_t ni) {
    inode nd = B->nodes[ni];
    const uint32_t n = nd.count;
    if (n <= 1) return 0;
    binacc bins[64];
    float best_cost = 0.0f;
    int best_axis = -1, best_uniform = 0;
    uint32_t best_left = 0;
    uint32_t* bin_of = B->bin_of + nd.begin;
    uint32_t* perm = B->perm + nd.begin;
    uint32_t* perm_tmp = (uint32_t*)(B->tmp + nd.begin); /* scratch, large enough */
    for (int axis = 0; axis < 3; axis++) {
        int uniform = bin_axis(B, &nd, axis, bins, bin_of, perm, perm_tmp);
        int nonempty = 0;
        for (int b = 0; b < 64; b++) nonempty += bins[b].count != 0;
        if (nonempty > 1) { /* are_bins_valid (bvh.rs:211-215) */
            float cost;
            uint32_t lcount;
            cheapest_split(&nd, bins, &cost, &lcount);
            if (best_axis < 0 || cost < best_cost) {
                best_cost = cost;
                best_axis = axis;
                best_uniform = uniform;
                best_left = lcount;
            }
        }
    }
    if (best_axis < 0) {
        B->error = -2; /* assert!(!is_first), bvh.rs:303 */
        return 0;
    }
    float no_split_cost = (float)n * 120.0f; /* tris_cost */
    if (no_split_cost < best_cost) return 0;

    /* Re-bin the winning axis and reorder the segment as the reference
     * collects it: bins in order, triangles within a bin in push order. */
    int uniform = bin_axis(B, &nd, best_axis, bins, bin_of, perm, perm_tmp);
    (void)best_uniform;
    (void)uniform;
    tref* R = B->refs + nd.begin;
    tref* T = B->tmp + nd.begin;
    uint32_t off[65];
    off[0] = 0;
    for (int b = 0; b < 64; b++) off[b + 1] = off[b] + bins[b].count;
    if (uniform) {
        for (uint32_t s = 0; s < n; s++) T[s] = R[perm[s]];
    } else {
        uint32_t cur[64];
        for (int b = 0; b < 64; b++) cur[b] = off[b];
        for (uint32_t k = 0; k < n; k++) T[cur[bin_of[k]]++] = R[k];
    }
    memcpy(R, T, n * sizeof(tref));
    int32_t l = new_inode(B, B->refs, nd.begin, best_left);
    int32_t r = new_inode(B, B->refs, nd.begin + best_left, n - best_left);
    B->nodes[ni].child[0] = l;
    B->nodes[ni].child[1] = r;
    return 1;
}

/* split_recursive (bvh.rs:325-356), depth first, iterative. */
static void split_recursive(builder* B, int32_t root) {
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * 4096);
    int cap = 4096, sp = 0;
    stack[sp++] = root;
    while (sp > 0 && !B->error) {
        int32_t ni = stack[--sp];
        if (!split_node(B, ni)) continue;
        inode* nd = &B->nodes[ni];
        int32_t c0 = nd->child[0], c1 = nd->child[1];
        /* smaller surface area first (bvh.rs:337-345) */
        if (aabb_area(B->nodes[c0].outer.mn, B->nodes[c0].outer.mx) >
            aabb_area(B->nodes[c1].outer.mn, B->nodes[c1].outer.mx)) {
            nd->child[0] = c1;
            nd->child[1] = c0;
        }
        if (sp + 2 > cap) {
            cap *= 2;
            stack = (int32_t*)realloc(stack, sizeof(int32_t) * cap);
        }
        stack[sp++] = B->nodes[ni].child[1];
        stack[sp++] = B->nodes[ni].child[0];
    }
    free(stack);
}

/* summed_area_ratio (bvh.rs:380-386), with Rust's left-to-right f32 sums. */
static float summed_area_ratio(builder* B, int32_t ni) {
    inode* nd = &B->nodes[ni];
    if (nd->child[0] < 0) return 0.0f + 0.0f;
    float cc = 0.0f;
    cc += summed_area_ratio(B, nd->child[0]);
    cc += summed_area_ratio(B, nd->child[1]);
    float self_area = aabb_area(nd->outer.mn, nd->outer.mx);
    float cs = 0.0f;
    inode* a = &B->nodes[nd->child[0]];
    inode* b = &B->nodes[nd->child[1]];
    cs += aabb_area(a->outer.mn, a->outer.mx) / self_area;
    cs += aabb_area(b->outer.mn, b->outer.mx) / self_area;
    return cs + cc;
}

/* crystallize (bvh.rs:389-428): pairs allocated in depth-first order. */
static void crystallize(builder* B, int32_t ni, or_node* nodes, uint32_t* nn, uint32_t* order,
                        uint32_t* ntri, uint32_t into) {
    inode* nd = &B->nodes[ni];
    for (int k = 0; k < 3; k++) {
        nodes[into].min[k] = nd->outer.mn[k];
        nodes[into].max[k] = nd->outer.mx[k];
    }
    if (nd->child[0] >= 0) {
        uint32_t ci = *nn;
        *nn += 2;
        crystallize(B, nd->child[0], nodes, nn, order, ntri, ci);
        crystallize(B, nd->child[1], nodes, nn, order, ntri, ci + 1);
        nodes[into].index = ci;
        nodes[into].len = 0;
    } else {
        nodes[into].index = *ntri;
        nodes[into].len = nd->count;
        for (uint32_t k = 0; k < nd->count; k++) order[(*ntri)++] = B->refs[nd->begin + k].idx;
    }
}

int or_bvh_build(const float* v, uint32_t ntris, or_node* nodes_out, uint32_t* nnodes_out,
                 uint32_t* order_out, float* avg_area_ratio, float* avg_tris_per_leaf) {
    return or_bvh_build_bins(v, ntris, 64, nodes_out, nnodes_out, order_out, avg_area_ratio,
                             avg_tris_per_leaf);
}

int or_bvh_build_bins(const float* v, uint32_t ntris, uint32_t nbins, or_node* nodes_out,
                      uint32_t* nnodes_out, uint32_t* order_out, float* avg_area_ratio,
                      float* avg_tris_per_leaf) {
    if (ntris == 0) return -1;
    if (nbins != 16 && nbins != 32 && nbins != 64) return -4;
    builder B;
    memset(&B, 0, sizeof(B));
    B.nbins = nbins;
    B.refs = (tref*)malloc(sizeof(tref) * ntris);
    B.tmp = (tref*)malloc(sizeof(tref) * ntris);
    B.bin_of = (uint32_t*)malloc(sizeof(uint32_t) * ntris);
    B.perm = (uint32_t*)malloc(sizeof(uint32_t) * ntris);
    B.keys = (float*)malloc(sizeof(float) * ntris);
    const float third = 1.0f / 3.0f; /* 3.0f32.recip() */
    for (uint32_t i = 0; i < ntris; i++) {
        const float* t = v + 9 * (size_t)i;
        tref* r = &B.refs[i];
        for (int k = 0; k < 3; k++) {
            float a = t[k], b = t[3 + k], c = t[6 + k];
            if (a != a || b != b || c != c) B.error = -3;
            /* Aabb::enclose_points (aabb.rs:56-71) */
            r->bmin[k] = tmin_f(tmin_f(a, b), c);
            r->bmax[k] = tmax_f(tmax_f(a, b), c);
            /* Triangle::barycenter (triangle.rs:55-57) */
            r->c[k] = ((a + b) + c) * third;
        }
        r->idx = i;
    }
    int status = 0;
    if (!B.error) {
        int32_t root = new_inode(&B, B.refs, 0, ntris);
        split_recursive(&B, root);
        if (B.error) {
            status = B.error;
        } else if (B.nodes[root].child[0] < 0) {
            status = -1; /* assert_eq!(2, root.children.len()), bvh.rs:512 */
        } else {
            uint32_t nn = 2, nt = 0;
            crystallize(&B, B.nodes[root].child[0], nodes_out, &nn, order_out, &nt, 0);
            crystallize(&B, B.nodes[root].child[1], nodes_out, &nn, order_out, &nt, 1);
            *nnodes_out = nn;
            /* statistics (bvh.rs:536-544) */
            uint32_t leaves = 0;
            for (uint32_t k = 0; k < B.nnodes; k++) leaves += B.nodes[k].child[0] < 0;
            float ratio = summed_area_ratio(&B, root);
            if (avg_area_ratio) *avg_area_ratio = ratio / (float)B.nnodes;
            if (avg_tris_per_leaf) *avg_tris_per_leaf = (float)ntris / (float)leaves;
        }
    } else {
        status = B.error;
    }
    free(B.refs);
    free(B.tmp);
    free(B.bin_of);
    free(B.perm);
    free(B.keys);
    free(B.nodes);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Scene (scene.rs)                                                         */
/* ------------------------------------------------------------------------ */

struct or_scene {
    uint32_t ntris;
    float* v;       /* BVH order, 9 per triangle */
    float* uv;      /* BVH order, 6 per triangle */
    uint32_t* mat;  /* BVH order */
    uint32_t* src;  /* BVH order -> source index */
    or_node* nodes;
    uint32_t nnodes;
    uint32_t* direct; /* Scene::direct_sample (scene.rs:114-129) */
    uint32_t ndirect;
    v3 cam_pos, cam_dpos;
    quat cam_q, cam_dq;
    float screen_distance;
    int emission_override;
    v3 emission;
    int approx;
};

or_scene* or_scene_create(const float* v, const float* uv, const uint32_t* mat, uint32_t ntris,
                          const or_node* nodes, uint32_t nnodes, const uint32_t* order,
                          int approx, int* status) {
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->ntris = ntris;
    s->approx = approx;
    uint32_t* ord = (uint32_t*)malloc(sizeof(uint32_t) * (ntris ? ntris : 1));
    if (nodes) {
        s->nnodes = nnodes;
        s->nodes = (or_node*)malloc(sizeof(or_node) * nnodes);
        memcpy(s->nodes, nodes, sizeof(or_node) * nnodes);
        memcpy(ord, order, sizeof(uint32_t) * ntris);
    } else {
        s->nodes = (or_node*)malloc(sizeof(or_node) * 2 * (size_t)(ntris ? ntris : 1));
        int st = or_bvh_build(v, ntris, s->nodes, &s->nnodes, ord, NULL, NULL);
        if (st != 0) {
            if (status) *status = st;
            free(ord);
            or_scene_destroy(s);
            return NULL;
        }
    }
    s->v = (float*)malloc(sizeof(float) * 9 * (size_t)ntris);
    s->uv = (float*)calloc(6 * (size_t)ntris, sizeof(float));
    s->mat = (uint32_t*)malloc(sizeof(uint32_t) * ntris);
    s->src = ord;
    s->direct = (uint32_t*)malloc(sizeof(uint32_t) * (ntris ? ntris : 1));
    for (uint32_t k = 0; k < ntris; k++) {
        uint32_t i = ord[k];
        memcpy(s->v + 9 * (size_t)k, v + 9 * (size_t)i, 36);
        if (uv) memcpy(s->uv + 6 * (size_t)k, uv + 6 * (size_t)i, 24);
        s->mat[k] = mat[i];
        if ((mat[i] & 0x40000000u) == 0x40000000u) s->direct[s->ndirect++] = k;
    }
    /* Camera::new (scene.rs:33-42) */
    s->cam_q.a = 1.0f;
    s->screen_distance = 1.0f / sinf(3.14159274f / 5.0f);
    if (status) *status = 0;
    return s;
}

void or_scene_destroy(or_scene* s) {
    if (!s) return;
    free(s->v);
    free(s->uv);
    free(s->mat);
    free(s->src);
    free(s->nodes);
    free(s->direct);
    free(s);
}

void or_scene_set_camera(or_scene* s, const float pos[3], const float dpos[3], const float q[4],
                         const float dq[4], float screen_distance) {
    s->cam_pos = ld3(pos);
    s->cam_dpos = ld3(dpos);
    s->cam_q.a = q[0];
    s->cam_q.b = q[1];
    s->cam_q.c = q[2];
    s->cam_q.d = q[3];
    s->cam_dq.a = dq[0];
    s->cam_dq.b = dq[1];
    s->cam_dq.c = dq[2];
    s->cam_dq.d = dq[3];
    s->screen_distance = screen_distance;
}

void or_scene_set_emission_override(or_scene* s, int enable, const float rgb[3]) {
    s->emission_override = enable;
    if (enable) s->emission = ld3(rgb);
}
void or_scene_set_approx(or_scene* s, int approx) { s->approx = approx; }
void or_scene_get_view(const or_scene* s, or_scene_view* out) {
    out->v = s->v;
    out->uv = s->uv;
    out->mat = s->mat;
    out->nodes = s->nodes;
    out->direct = s->direct;
    out->src = s->src;
    out->ntris = s->ntris;
    out->nnodes = s->nnodes;
    out->ndirect = s->ndirect;
    const float pos[3] = {s->cam_pos.x, s->cam_pos.y, s->cam_pos.z};
    const float dpos[3] = {s->cam_dpos.x, s->cam_dpos.y, s->cam_dpos.z};
    const float q[4] = {s->cam_q.a, s->cam_q.b, s->cam_q.c, s->cam_q.d};
    const float dq[4] = {s->cam_dq.a, s->cam_dq.b, s->cam_dq.c, s->cam_dq.d};
    memcpy(out->cam_pos, pos, 12);
    memcpy(out->cam_dpos, dpos, 12);
    memcpy(out->cam_q, q, 16);
    memcpy(out->cam_dq, dq, 16);
    out->screen_distance = s->screen_distance;
    out->emission_override = s->emission_override;
    out->emission[0] = s->emission.x;
    out->emission[1] = s->emission.y;
    out->emission[2] = s->emission.z;
    out->approx = s->approx;
}
uint32_t or_scene_num_nodes(const or_scene* s) { return s->nnodes; }
uint32_t or_scene_num_direct(const or_scene* s) { return s->ndirect; }
void or_scene_copy_bvh(const or_scene* s, or_node* nodes, uint32_t* order) {
    if (nodes) memcpy(nodes, s->nodes, sizeof(or_node) * s->nnodes);
    if (order) memcpy(order, s->src, sizeof(uint32_t) * s->ntris);
}

/* ------------------------------------------------------------------------ */
/* Traversal (bvh.rs:589-657)                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    v3 o, d;
    uint32_t active; /* sign bit 1 = inactive (ray.rs:21-33) */
} ray1;

typedef struct {
    v3 pos, normal;
    float dist;
    uint32_t mat;
    float tu, tv;
    uint32_t tri; /* BVH order index or OR_NO_HIT */
    float bu, bv; /* barycentrics (u weights v2, v weights v1) */
} isect1;

/* Scene::intersect_nearest's far-away sky hit (scene.rs:220-230) */
static inline isect1 far_away(const ray1* r) {
    isect1 is;
    is.pos = muladd3(r->d, 1.0e5f, r->o);
    is.normal = r->d;
    is.dist = 1.0e5f;
    is.mat = 0xC0000000u; /* SMaterial::sky (material.rs:85-90) */
    is.tu = is.tv = 0.0f;
    is.tri = OR_NO_HIT;
    is.bu = is.bv = 0.0f;
    return is;
}

typedef struct {
    uint32_t node;
    float tmin;
} stk_entry;

/* Single-ray traversal: the reference packet traversal with the ray in every
 * lane (MRay::broadcast), so `should_try_before` over all lanes reduces to
 * this ray's sign test.  Deferred outputs (normal, position, tex coords) of
 * Triangle::intersect are computed once for the winning triangle; they are
 * pure functions of (triangle, ray, t, u, v) so the values are identical. */
static void traverse1(const or_scene* s, const ray1* r, isect1* is, uint32_t* numi_aabb,
                      uint32_t* numi_tri, int approx) {
    stk_entry stack[256];
    int sp = 0;
    uint32_t na = 2, nt = 0;
    v3 inv = mk3(or_rcp(r->d.x, approx), or_rcp(r->d.y, approx), or_rcp(r->d.z, approx));
    const int active = (r->active >> 31) == 0;
    float best_t = is->dist;
    uint32_t best = OR_NO_HIT;
    float bu = 0, bv = 0, bw = 0;

#define PUSH_PAIR(base)                                                                  \
    do {                                                                                 \
        const or_node* c0 = &s->nodes[(base)];                                           \
        const or_node* c1 = &s->nodes[(base) + 1];                                       \
        aabbx i0 = aabb_isect(c0->min, c0->max, r->o, inv);                              \
        aabbx i1 = aabb_isect(c1->min, c1->max, r->o, inv);                              \
        int h0 = active && (i0.mask >> 31) && sgn(i0.tmax) == 0;                         \
        int h1 = active && (i1.mask >> 31) && sgn(i1.tmax) == 0;                         \
        if (sgn(i0.tmin - i1.tmin) == 0) {                                               \
            if (h0) { stack[sp].node = (base); stack[sp].tmin = i0.tmin; sp++; }         \
            if (h1) { stack[sp].node = (base) + 1; stack[sp].tmin = i1.tmin; sp++; }     \
        } else {                                                                         \
            if (h1) { stack[sp].node = (base) + 1; stack[sp].tmin = i1.tmin; sp++; }     \
            if (h0) { stack[sp].node = (base); stack[sp].tmin = i0.tmin; sp++; }         \
        }                                                                                \
    } while (0)

    PUSH_PAIR(0u);
    while (sp > 0) {
        stk_entry e = stack[--sp];
        /* is_further_away_than (aabb.rs:161-170) */
        if (geq(e.tmin, best_t) >> 31) continue;
        const or_node* nd = &s->nodes[e.node];
        if (nd->len == 0) {
            na += 2;
            PUSH_PAIR(nd->index);
        } else {
            for (uint32_t k = nd->index; k < nd->index + nd->len; k++) {
                float t, u, v, w;
                uint32_t m = tri_core(s->v + 9 * (size_t)k, r->o, r->d, r->active, best_t, &t,
                                      &u, &v, &w);
                nt++;
                if ((m >> 31) == 0) {
                    best_t = t;
                    best = k;
                    bu = u;
                    bv = v;
                    bw = w;
                }
            }
        }
    }
#undef PUSH_PAIR
    if (best != OR_NO_HIT) {
        const float* tv = s->v + 9 * (size_t)best;
        v3 v0 = ld3(tv), v1 = ld3(tv + 3), v2 = ld3(tv + 6);
        v3 n = cross3(sub3(v0, v2), sub3(v1, v0));
        const float* uv = s->uv + 6 * (size_t)best;
        is->pos = muladd3(r->d, best_t, r->o);
        is->normal = normalized3(n, approx);
        is->dist = best_t;
        is->mat = s->mat[best];
        is->tu = fmaf(uv[0], bw, fmaf(uv[2], bv, uv[4] * bu));
        is->tv = fmaf(uv[1], bw, fmaf(uv[3], bv, uv[5] * bu));
        is->tri = best;
        is->bu = bu;
        is->bv = bv;
    }
    if (numi_aabb) *numi_aabb = na;
    if (numi_tri) *numi_tri = nt;
}

typedef struct {
    const or_scene* s;
    const float* rays;
    size_t n;
    uint32_t* tri_id;
    float* tuv;
    uint32_t* na;
    uint32_t* nt;
    float* isect;
    size_t begin, end;
} isect_job;

static void* isect_worker(void* arg) {
    isect_job* j = (isect_job*)arg;
    const or_scene* s = j->s;
    for (size_t i = j->begin; i < j->end; i++) {
        ray1 r;
        r.o = ld3(j->rays + 6 * i);
        r.d = ld3(j->rays + 6 * i + 3);
        r.active = 0;
        isect1 is = far_away(&r);
        uint32_t na, nt;
        traverse1(s, &r, &is, &na, &nt, s->approx);
        if (j->tri_id) j->tri_id[i] = is.tri == OR_NO_HIT ? OR_NO_HIT : s->src[is.tri];
        if (j->tuv) {
            j->tuv[3 * i] = is.dist;
            j->tuv[3 * i + 1] = is.bu;
            j->tuv[3 * i + 2] = is.bv;
        }
        if (j->na) j->na[i] = na;
        if (j->nt) j->nt[i] = nt;
        if (j->isect) {
            float* o = j->isect + 12 * i;
            o[0] = is.pos.x;
            o[1] = is.pos.y;
            o[2] = is.pos.z;
            o[3] = is.normal.x;
            o[4] = is.normal.y;
            o[5] = is.normal.z;
            o[6] = is.dist;
            o[7] = u2f(is.mat);
            o[8] = is.tu;
            o[9] = is.tv;
            o[10] = 0;
            o[11] = 0;
        }
    }
    return NULL;
}

void or_intersect_nearest(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                          float* tuv, uint32_t* numi_aabb, uint32_t* numi_tri, float* isect,
                          int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    isect_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        isect_job* j = &jobs[t];
        j->s = s;
        j->rays = rays;
        j->n = n;
        j->tri_id = tri_id;
        j->tuv = tuv;
        j->na = numi_aabb;
        j->nt = numi_tri;
        j->isect = isect;
        j->begin = n * t / nthreads;
        j->end = n * (t + 1) / nthreads;
        if (nthreads > 1)
            pthread_create(&th[t], NULL, isect_worker, j);
        else
            isect_worker(j);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* The reference packet traversal (bvh.rs:589-657) on 8 lanes, verbatim:
 * pushes when any active lane hits, orders children by all 8 lanes, skips
 * when every hitting active lane's tmin is beyond its current distance. */
void or_intersect_nearest_packet(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                                 float* tuv, uint32_t* numi_aabb, uint32_t* numi_tri) {
    typedef struct {
        aabbx x[8];
        uint32_t node;
    } pentry;
    pentry* stack = (pentry*)malloc(sizeof(pentry) * 512);
    for (size_t p = 0; p + 8 <= n; p += 8) {
        v3 o[8], d[8], inv[8];
        float best_t[8], bu[8], bv[8];
        uint32_t best[8];
        for (int l = 0; l < 8; l++) {
            o[l] = ld3(rays + 6 * (p + l));
            d[l] = ld3(rays + 6 * (p + l) + 3);
            inv[l] = mk3(or_rcp(d[l].x, s->approx), or_rcp(d[l].y, s->approx),
                         or_rcp(d[l].z, s->approx));
            best_t[l] = 1.0e5f;
            best[l] = OR_NO_HIT;
            bu[l] = bv[l] = 0.0f;
        }
        int sp = 0;
        uint32_t na = 2, nt = 0;
        uint32_t pairs[1] = {0};
        (void)pairs;
        uint32_t base = 0;
        for (;;) {
            aabbx a0[8], a1[8];
            int any0 = 0, any1 = 0, before = 1;
            for (int l = 0; l < 8; l++) {
                a0[l] = aabb_isect(s->nodes[base].min, s->nodes[base].max, o[l], inv[l]);
                a1[l] = aabb_isect(s->nodes[base + 1].min, s->nodes[base + 1].max, o[l], inv[l]);
                any0 |= (a0[l].mask >> 31) && sgn(a0[l].tmax) == 0;
                any1 |= (a1[l].mask >> 31) && sgn(a1[l].tmax) == 0;
                before &= sgn(a0[l].tmin - a1[l].tmin) == 0;
            }
            if (before) {
                if (any0) { memcpy(stack[sp].x, a0, sizeof(a0)); stack[sp++].node = base; }
                if (any1) { memcpy(stack[sp].x, a1, sizeof(a1)); stack[sp++].node = base + 1; }
            } else {
                if (any1) { memcpy(stack[sp].x, a1, sizeof(a1)); stack[sp++].node = base + 1; }
                if (any0) { memcpy(stack[sp].x, a0, sizeof(a0)); stack[sp++].node = base; }
            }
            int found = 0;
            while (sp > 0) {
                pentry* e = &stack[--sp];
                int further = 1;
                for (int l = 0; l < 8; l++)
                    if ((e->x[l].mask >> 31) && !(geq(e->x[l].tmin, best_t[l]) >> 31)) further = 0;
                if (further) continue;
                const or_node* nd = &s->nodes[e->node];
                if (nd->len == 0) {
                    na += 2;
                    base = nd->index;
                    found = 1;
                    break;
                }
                for (uint32_t k = nd->index; k < nd->index + nd->len; k++) {
                    nt++;
                    for (int l = 0; l < 8; l++) {
                        float t, u, v, w;
                        uint32_t m = tri_core(s->v + 9 * (size_t)k, o[l], d[l], 0u, best_t[l], &t,
                                              &u, &v, &w);
                        if ((m >> 31) == 0) {
                            best_t[l] = t;
                            best[l] = k;
                            bu[l] = u;
                            bv[l] = v;
                        }
                    }
                }
            }
            if (!found) break;
        }
        for (int l = 0; l < 8; l++) {
            size_t i = p + l;
            if (tri_id) tri_id[i] = best[l] == OR_NO_HIT ? OR_NO_HIT : s->src[best[l]];
            if (tuv) {
                tuv[3 * i] = best_t[l];
                tuv[3 * i + 1] = bu[l];
                tuv[3 * i + 2] = bv[l];
            }
            if (numi_aabb) numi_aabb[i] = na;
            if (numi_tri) numi_tri[i] = nt;
        }
    }
    free(stack);
}

void or_intersect_brute_f64(const or_scene* s, const float* rays, size_t n, uint32_t* tri_id,
                            double* tuv) {
    for (size_t i = 0; i < n; i++) {
        double o[3], d[3];
        for (int k = 0; k < 3; k++) {
            o[k] = rays[6 * i + k];
            d[k] = rays[6 * i + 3 + k];
        }
        double best_t = 1.0e5;
        uint32_t best = OR_NO_HIT;
        double bu = 0, bv = 0;
        for (uint32_t k = 0; k < s->ntris; k++) {
            const float* tv = s->v + 9 * (size_t)k;
            double v0[3], v1[3], v2[3], e1[3], e2[3], nn[3], fr[3], c[3];
            for (int q = 0; q < 3; q++) {
                v0[q] = tv[q];
                v1[q] = tv[3 + q];
                v2[q] = tv[6 + q];
                e1[q] = v0[q] - v2[q];
                e2[q] = v1[q] - v0[q];
                fr[q] = v0[q] - o[q];
            }
            nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
            nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
            nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
            double den = d[0] * nn[0] + d[1] * nn[1] + d[2] * nn[2];
            if (den == 0.0) continue;
            double t = (fr[0] * nn[0] + fr[1] * nn[1] + fr[2] * nn[2]) / den;
            c[0] = d[1] * fr[2] - d[2] * fr[1];
            c[1] = d[2] * fr[0] - d[0] * fr[2];
            c[2] = d[0] * fr[1] - d[1] * fr[0];
            double u = (c[0] * e2[0] + c[1] * e2[1] + c[2] * e2[2]) / den;
            double v = (c[0] * e1[0] + c[1] * e1[1] + c[2] * e1[2]) / den;
            double w = 1.0 - u - v;
            if (t >= 0 && u >= 0 && v >= 0 && w >= 0 && t < best_t) {
                best_t = t;
                best = k;
                bu = u;
                bv = v;
            }
        }
        tri_id[i] = best == OR_NO_HIT ? OR_NO_HIT : s->src[best];
        if (tuv) {
            tuv[3 * i] = best_t;
            tuv[3 * i + 1] = bu;
            tuv[3 * i + 2] = bv;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Shading (material.rs, scene.rs)                                          */
/* ------------------------------------------------------------------------ */

/* MMaterial::get_color (material.rs:151-168) */
static inline v3 mat_color(uint32_t m) {
    int32_t mi = (int32_t)m;
    float r = (float)(mi & 0xff);
    float g = (float)((mi >> 8) & 0xff);
    float b = (float)((mi >> 16) & 0xff);
    const float k = 1.0f / 255.0f;
    return mk3(r * k, g * k, b * k);
}
/* get_glossiness (material.rs:171-178) */
static inline int32_t mat_gloss(uint32_t m) { return ((int32_t)m >> 26) & 7; }
/* get_texture (material.rs:181-188) */
static inline int32_t mat_texture(uint32_t m) { return ((int32_t)m >> 24) & 3; }
/* has_texture (material.rs:192-202): sign bit = bit 25 | bit 24 */
static inline uint32_t mat_has_texture(uint32_t m) { return (m << 6) | (m << 7); }

/* sky_intensity (material.rs:205-215) */
static inline v3 sky_intensity(v3 dir) {
    v3 up = mk3(0.0f, 0.0f, 1.0f);
    float d = fmaf(dot3(dir, up), 0.5f, 0.5f);
    float r = d;
    float g = d * d;
    float b = d * (d * d);
    return muladd3(mk3(r, g, b), 2.0f, mk3(0.5f, 0.5f, 0.5f));
}

/* microfacet_normal_dist (material.rs:475-513) */
static inline float normal_dist(v3 h, v3 n, int32_t gloss) {
    const float pi = 3.14159274f;
    float c1 = dot3(h, n);
    float c2 = c1 * c1;
    float c4 = c2 * c2;
    float c8 = c4 * c4;
    float c16 = c8 * c8;
    switch (gloss) {
    case 0: return 0.5f / pi;
    case 1: return xmax(c1, 0.0f) * (1.0f / pi);
    case 2: return c2 * (1.5f / pi);
    case 3: return c4 * (2.5f / pi);
    case 4: return c8 * (4.5f / pi);
    default: return c16 * (8.5f / pi); /* index 5; 6 is out of range in the reference */
    }
}

/* microfacet_fresnel (material.rs:453-469) */
static inline v3 fresnel(v3 incoming, v3 h, v3 color, float* raw) {
    v3 r0 = color;
    v3 r1 = sub3(mk3(1.0f, 1.0f, 1.0f), r0);
    float ct = 1.0f - fabs_bits(dot3(h, incoming));
    float ct2 = ct * ct;
    float ct3 = ct2 * ct;
    float ct5 = fabs_bits(ct2 * ct3);
    *raw = ct5;
    return muladd3(r1, ct5, r0);
}

/* microfacet_brdf (material.rs:415-450) */
static inline v3 microfacet_brdf(uint32_t material, v3 in_dir, v3 out_dir, v3 normal,
                                 int ignore_fresnel, int approx, float* fr_raw) {
    v3 color = mat_color(material);
    int32_t gloss = mat_gloss(material);
    v3 h = normalized3(sub3(in_dir, out_dir), approx);
    float d = normal_dist(h, normal, gloss);
    v3 fc = fresnel(in_dir, h, color, fr_raw);
    if (ignore_fresnel) fc = pick3(fc, mk3(1.0f, 1.0f, 1.0f), mat_has_texture(material));
    return scale3(fc, d);
}

/* Scene::get_direct_sample for one lane (scene.rs:143-202).  normal and area
 * (may be NULL) are computed as the reference computes them; its caller
 * continue_path_direct_sample (material.rs:253-265) uses only the position. */
static inline void direct_sample(const or_scene* s, uint32_t bits, float u, float v, v3* pos,
                                 v3* normal, float* area) {
    uint32_t idx = bits >> 29;
    const float* tv = s->v + 9 * (size_t)s->direct[idx];
    v3 v0 = ld3(tv), v1 = ld3(tv + 3), v2 = ld3(tv + 6);
    v3 e1 = sub3(v0, v2);
    v3 e2 = sub3(v1, v0);
    if (normal || area) {
        v3 nd = cross3(e1, e2);
        float rn = or_rsqrt(dot3(nd, nd), s->approx);
        if (normal) *normal = scale3(nd, rn);
        if (area) *area = 0.5f * or_rcp(rn, s->approx);
    }
    float sgn_s = (1.0f - u) - v;
    uint32_t m = f2u(sgn_s);
    float uu = pickf(u, 1.0f - u, m);
    float vv = pickf(v, 1.0f - v, m);
    *pos = muladd3(e2, vv, negmuladd3(e1, uu, v0));
}

/* pd_brdf (material.rs:245-250) */
static inline float pd_brdf(v3 normal, v3 dir) {
    return xmax(dot3(normal, dir), 0.0f) * (1.0f / 3.14159274f);
}

/* pd_direct_sample (material.rs:281-320) */
static inline float pd_direct(const or_scene* s, v3 o, v3 d, int approx) {
    float total = 0.0f;
    for (uint32_t k = 0; k < s->ndirect; k++) {
        direct_isect si = tri_direct(s->v + 9 * (size_t)s->direct[k], o, d, approx);
        float dist_sqr = si.distance * si.distance;
        float dot_em = fabs_bits(dot3(si.normal, d)) + 0.0001f;
        float pd = dist_sqr * or_rcp(si.area * dot_em, approx);
        total = pickf(total + pd, total, si.mask);
    }
    return total * (1.0f / (float)s->ndirect);
}

typedef struct {
    uint32_t d_angle, d_unit, d_u32, d_u, d_v, d_rr;
} path_draws;

/* continue_path for one lane (material.rs:327-411).  Returns the new ray,
 * the colour modulation and the raw Fresnel factor. */
static void continue_path(const or_scene* s, const ray1* ray, const isect1* is,
                          const path_draws* dr, int ignore_fresnel, ray1* out, v3* color_mod,
                          float* fresnel_raw) {
    const int approx = s->approx;
    uint32_t active = ray->active | is->mat;
    /* continue_path_brdf (material.rs:228-242) */
    v3 dir_z = hemisphere_from(u32_angle(dr->d_angle), u32_unit(dr->d_unit));
    v3 dir_b = rotate_hemisphere(dir_z, is->normal, approx);
    v3 org_b = muladd3(dir_b, 1.0e-4f, is->pos);
    /* continue_path_direct_sample (material.rs:253-265) */
    v3 p;
    direct_sample(s, dr->d_u32, u32_unit(dr->d_u), u32_unit(dr->d_v), &p, NULL, NULL);
    v3 dir_d = normalized3(sub3(p, is->pos), approx);
    v3 org_d = muladd3(dir_d, 1.0e-4f, is->pos);
    float rr = fmaf(is->normal.z, 0.8f, u32_biunit(dr->d_rr));
    uint32_t rrm = f2u(rr);
    v3 n_org = pick3(org_b, org_d, rrm);
    v3 n_dir = pick3(dir_b, dir_d, rrm);
    float pd_b = pd_brdf(is->normal, n_dir);
    float pd_d = pd_direct(s, n_org, n_dir, approx);
    float weight_denom = (pd_b + pd_d) + 0.01f;
    float p_brdf = fmaf(is->normal.z, 0.4f, 0.5f);
    float p_direct = fmaf(-is->normal.z, 0.4f, 0.5f);
    float pp = pickf(p_brdf, p_direct, rrm);
    float modulation = or_rcp(weight_denom * pp, approx);
    float cos_theta = xmax(dot3(is->normal, n_dir), 0.0f);
    v3 brdf = microfacet_brdf(is->mat, n_dir, ray->d, is->normal, ignore_fresnel, approx,
                              fresnel_raw);
    v3 cm = scale3(brdf, modulation * cos_theta);
    cm = mk3(xmin(cm.x, 2.0f), xmin(cm.y, 2.0f), xmin(cm.z, 2.0f));
    out->o = pick3(n_org, ray->o, active);
    out->d = pick3(n_dir, ray->d, active);
    out->active = active;
    *color_mod = pick3(cm, mk3(1.0f, 1.0f, 1.0f), active);
}

/* ---- single-lane shading entry points (checked against independent FP64
 *      restatements and closed forms by the tests) ---- */

float or_microfacet_normal_dist(const float h[3], const float n[3], int32_t gloss) {
    return normal_dist(ld3(h), ld3(n), gloss);
}

void or_microfacet_fresnel(const float incoming[3], const float h[3], const float color[3],
                           float out[3], float* raw) {
    v3 f = fresnel(ld3(incoming), ld3(h), ld3(color), raw);
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
}

void or_microfacet_brdf(uint32_t material, const float in_dir[3], const float out_dir[3],
                        const float normal[3], int ignore_fresnel, int approx, float out[3],
                        float* raw) {
    v3 b = microfacet_brdf(material, ld3(in_dir), ld3(out_dir), ld3(normal), ignore_fresnel,
                           approx, raw);
    out[0] = b.x;
    out[1] = b.y;
    out[2] = b.z;
}

float or_pd_brdf(const float normal[3], const float dir[3]) {
    return pd_brdf(ld3(normal), ld3(dir));
}

void or_sky_intensity_lane(const float dir[3], float out[3]) {
    v3 c = sky_intensity(ld3(dir));
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

int or_pd_direct(const or_scene* s, const float* rays, size_t n, float* out) {
    if (s->ndirect == 0) return -1;
    for (size_t i = 0; i < n; i++)
        out[i] = pd_direct(s, ld3(rays + 6 * i), ld3(rays + 6 * i + 3), s->approx);
    return 0;
}

int or_direct_sample(const or_scene* s, const uint32_t* draws3, size_t n, float* pos,
                     float* normal, float* area) {
    if (s->ndirect != 8) return -1; /* the index bits >> 29 address exactly 8 (scene.rs:147) */
    for (size_t i = 0; i < n; i++) {
        const uint32_t* d = draws3 + 3 * i;
        v3 p, nm;
        float a;
        direct_sample(s, d[0], u32_unit(d[1]), u32_unit(d[2]), &p, &nm, &a);
        pos[3 * i] = p.x;
        pos[3 * i + 1] = p.y;
        pos[3 * i + 2] = p.z;
        if (normal) {
            normal[3 * i] = nm.x;
            normal[3 * i + 1] = nm.y;
            normal[3 * i + 2] = nm.z;
        }
        if (area) area[i] = a;
    }
    return 0;
}

int or_continue_path(const or_scene* s, size_t n, const float* ray7, const float* isect7,
                     const uint32_t* draws6, int ignore_fresnel, float* out_ray7,
                     float* color_mod3, float* fresnel_raw) {
    if (s->ndirect != 8) return -1;
    for (size_t i = 0; i < n; i++) {
        const float* r = ray7 + 7 * i;
        const float* is = isect7 + 7 * i;
        const uint32_t* d = draws6 + 6 * i;
        ray1 ray = {ld3(r), ld3(r + 3), f2u(r[6])};
        isect1 hit;
        memset(&hit, 0, sizeof(hit));
        hit.pos = ld3(is);
        hit.normal = ld3(is + 3);
        hit.mat = f2u(is[6]);
        path_draws dr = {d[0], d[1], d[2], d[3], d[4], d[5]};
        ray1 nr;
        v3 cm;
        float fr;
        continue_path(s, &ray, &hit, &dr, ignore_fresnel, &nr, &cm, &fr);
        float* o = out_ray7 + 7 * i;
        o[0] = nr.o.x;
        o[1] = nr.o.y;
        o[2] = nr.o.z;
        o[3] = nr.d.x;
        o[4] = nr.d.y;
        o[5] = nr.d.z;
        o[6] = u2f(nr.active);
        color_mod3[3 * i] = cm.x;
        color_mod3[3 * i + 1] = cm.y;
        color_mod3[3 * i + 2] = cm.z;
        if (fresnel_raw) fresnel_raw[i] = fr;
    }
    return 0;
}

/* Camera::get_ray (scene.rs:81-99) */
static inline ray1 camera_ray(const or_scene* s, float x, float y, float t) {
    ray1 r;
    r.o = muladd3(s->cam_dpos, t, s->cam_pos);
    quat q = quat_interpolate(s->cam_q, s->cam_dq, t);
    float dist = -s->screen_distance;
    v3 src = normalized3(mk3(x, y, dist), s->approx);
    r.d = quat_rotate(src, q);
    r.active = 0;
    return r;
}

/* ------------------------------------------------------------------------ */
/* Renderer (renderer.rs)                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    v3 color[8];
    int32_t tex_index[8];
    float tu[8], tv[8], fresnel[8];
} pixel_data8;

typedef struct {
    uint64_t queries, sum_aabb, sum_tri;
} work_counts;

/* render_pixels for one 8-lane group (renderer.rs:413-468) */
static void render_pixels(const or_scene* s, const float x[8], const float y[8], rng8* rng,
                          pixel_data8* out, work_counts* wc) {
    float t[8];
    rng8_unit(rng, t);
    ray1 ray[8];
    v3 color[8];
    uint32_t hit_emissive[8];
    int term_tri[8];
    for (int l = 0; l < 8; l++) {
        ray[l] = camera_ray(s, x[l], y[l], t[l]);
        color[l] = mk3(1.0f, 1.0f, 1.0f);
        hit_emissive[l] = 0;
        term_tri[l] = 0;
        out->tex_index[l] = 0;
        out->tu[l] = out->tv[l] = out->fresnel[l] = 0.0f;
    }
    /* In PIXEL mode every pixel is its own group: its loop ends when its own
     * ray hits an emissive surface.  The radiance is the same as with the
     * 8-lane break (a deactivated lane keeps its ray and gets white
     * modulations); only the first-bounce gbuffer of a pixel whose camera ray
     * hits a light differs (zeros instead of the draws of its neighbours). */
    int done[8] = {0};
    const int per_pixel = rng->mode == OR_RNG_PIXEL;
    const int max_bounces = 5;
    for (int i = 0; i < max_bounces; i++) {
        isect1 is[8];
        uint32_t all_em = 1;
        for (int l = 0; l < 8; l++) {
            if (done[l]) {
                is[l] = far_away(&ray[l]);
                continue;
            }
            is[l] = far_away(&ray[l]);
            if ((ray[l].active >> 31) == 0) {
                uint32_t na, nt;
                traverse1(s, &ray[l], &is[l], &na, &nt, s->approx);
                if (wc) {
                    wc->queries++;
                    wc->sum_aabb += na;
                    wc->sum_tri += nt;
                }
                if (is[l].mat >> 31) term_tri[l] = is[l].tri != OR_NO_HIT;
            }
            hit_emissive[l] = is[l].mat;
            all_em &= is[l].mat >> 31;
        }
        if (all_em) break;
        uint32_t d[6][8];
        for (int k = 0; k < 6; k++) rng8_next(rng, d[k]);
        for (int l = 0; l < 8; l++) {
            if (per_pixel && (done[l] || (is[l].mat >> 31))) {
                done[l] = 1;
                continue;
            }
            path_draws dr = {d[0][l], d[1][l], d[2][l], d[3][l], d[4][l], d[5][l]};
            ray1 nr;
            v3 cm;
            float fr;
            continue_path(s, &ray[l], &is[l], &dr, i == 0, &nr, &cm, &fr);
            ray[l] = nr;
            color[l] = mulc3(color[l], cm);
            if (i == 0) {
                out->tex_index[l] = mat_texture(is[l].mat);
                out->tu[l] = is[l].tu;
                out->tv[l] = is[l].tv;
                out->fresnel[l] = fr;
            }
        }
    }
    for (int l = 0; l < 8; l++) {
        v3 em = sky_intensity(ray[l].d);
        if (s->emission_override && term_tri[l]) em = s->emission;
        v3 c = mulc3(color[l], em);
        out->color[l] = pick3(mk3(0.0f, 0.0f, 0.0f), c, hit_emissive[l]);
    }
}

/* get_pixel_coords_16x4 + render_block_16x4 (renderer.rs:171-206,292-300) */
static void render_block(const or_scene* s, uint32_t width, uint32_t height, uint32_t x,
                         uint32_t y, uint32_t frame, rng8* patch_rng, int rng_mode,
                         pixel_data8 out[8], work_counts* wc) {
    const float scale = 2.0f / (float)width;
    const float sm[4] = {2.0f * scale, 4.0f * scale, 8.0f * scale, 12.0f * scale};
    const float off_x[8] = {0, 1, 2, 3, 0, 1, 2, 3};
    const float off_y[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    float bx[8], by[8];
    const float cx = (float)x - (float)width * 0.5f;
    const float cy = (float)y - (float)height * 0.5f;
    for (int l = 0; l < 8; l++) {
        bx[l] = scale * (off_x[l] + cx);
        by[l] = scale * (off_y[l] + cy);
    }
    float xs[8][8], ys[8][8];
    for (int g = 0; g < 8; g++)
        for (int l = 0; l < 8; l++) {
            xs[g][l] = (g / 2 == 0) ? bx[l] : bx[l] + sm[g / 2];
            ys[g][l] = (g % 2 == 0) ? by[l] : by[l] + sm[0];
        }
    rng8 group_rng[8];
    if (rng_mode == OR_RNG_REF_PATCH) {
        float u[8];
        for (int g = 0; g < 8; g++) {
            rng8_unit(patch_rng, u);
            for (int l = 0; l < 8; l++) xs[g][l] = fmaf(u[l], scale, xs[g][l]);
        }
        for (int g = 0; g < 8; g++) {
            rng8_unit(patch_rng, u);
            for (int l = 0; l < 8; l++) ys[g][l] = fmaf(u[l], scale, ys[g][l]);
        }
    } else {
        for (int g = 0; g < 8; g++) {
            group_rng[g].mode = OR_RNG_PIXEL;
            for (int l = 0; l < 8; l++) {
                uint32_t px = x + 4 * (g / 2) + (l % 4);
                uint32_t py = y + 2 * (g % 2) + (l / 4);
                uint64_t st = rng_pixel_seed(py * width + px, frame);
                float ux = u32_unit(rng_pixel_next(&st));
                float uy = u32_unit(rng_pixel_next(&st));
                xs[g][l] = fmaf(ux, scale, xs[g][l]);
                ys[g][l] = fmaf(uy, scale, ys[g][l]);
                group_rng[g].pix[l] = st;
            }
        }
    }
    for (int g = 0; g < 8; g++)
        render_pixels(s, xs[g], ys[g], rng_mode == OR_RNG_REF_PATCH ? patch_rng : &group_rng[g],
                      &out[g], wc);
}

/* Pixel of (group, lane) inside a 16x4 block (renderer.rs:154-170). */
static inline void block_pixel(int g, int l, uint32_t* dx, uint32_t* dy) {
    *dx = 4 * (g / 2) + (l % 4);
    *dy = 2 * (g % 2) + (l / 4);
}

/* store_pixels_color_16x4 (renderer.rs:239-256) */
static inline uint32_t pack_color(v3 c) {
    v3 x2 = scale3(c, 2.0f);
    float r = xmin(1.0f, x2.x) * 255.0f;
    float g = xmin(1.0f, x2.y) * 255.0f;
    float b = xmin(1.0f, x2.z) * 255.0f;
    uint32_t ri = (uint32_t)or_cvt_f32_i32(r);
    uint32_t gi = (uint32_t)or_cvt_f32_i32(g) << 8;
    uint32_t bi = (uint32_t)or_cvt_f32_i32(b) << 16;
    return (ri | gi) | bi;
}
/* store_pixels_gbuffer_16x4 (renderer.rs:260-286) */
static inline uint32_t pack_gbuffer(int32_t tex_index, float tu, float tv, float fr) {
    uint32_t r = (uint32_t)or_cvt_f32_i32(tu * 255.0f) & 0xffu;
    uint32_t g = ((uint32_t)or_cvt_f32_i32(tv * 255.0f) & 0xffu) << 8;
    uint32_t b = (uint32_t)or_cvt_f32_i32(fr * 255.0f) << 16;
    uint32_t a = (uint32_t)tex_index << 24;
    return (r | g) | (b | a);
}

static void store_block(uint32_t width, uint32_t* bitmap, uint32_t* gbuffer, uint32_t x, uint32_t y,
                        const pixel_data8 d[8], int color) {
    for (int g = 0; g < 8; g++)
        for (int l = 0; l < 8; l++) {
            uint32_t dx, dy;
            block_pixel(g, l, &dx, &dy);
            size_t idx = (size_t)(y + dy) * width + (x + dx);
            if (color && bitmap) bitmap[idx] = pack_color(d[g].color[l]);
            if (gbuffer)
                gbuffer[idx] = pack_gbuffer(d[g].tex_index[l], d[g].tu[l], d[g].tv[l],
                                            d[g].fresnel[l]);
        }
}

static int check_render_args(const or_scene* s, uint32_t width, uint32_t height, uint32_t pw,
                             uint32_t x, uint32_t y) {
    if ((width & 15) || (height & 15) || (pw & 15) || pw == 0) return -10;
    if (x + pw > width || y + pw > height) return -11;
    if (s->ndirect < 8) return -12;
    return 0;
}

static int patch_impl(const or_scene* s, uint32_t width, uint32_t height, uint32_t* bitmap,
                      float* hdr, uint32_t* gbuffer, uint32_t pw, uint32_t x, uint32_t y,
                      uint32_t frame, int rng_mode, work_counts* wc) {
    int st = check_render_args(s, width, height, pw, x, y);
    if (st) return st;
    uint32_t w = pw / 16, h = pw / 4;
    rng8 rng;
    memset(&rng, 0, sizeof(rng));
    rng.mode = rng_mode;
    if (rng_mode == OR_RNG_REF_PATCH) rng_ref_seed(rng.ref, x, y, frame);
    for (uint32_t i = 0; i < w; i++)
        for (uint32_t j = 0; j < h; j++) {
            uint32_t xb = x + i * 16, yb = y + j * 4;
            pixel_data8 data[8];
            render_block(s, width, height, xb, yb, frame, &rng, rng_mode, data, wc);
            if (hdr) {
                size_t index = (size_t)(y / 4 + j) * (width / 16) + (x / 16 + i);
                float* blk = hdr + index * 192;
                for (int g = 0; g < 8; g++)
                    for (int l = 0; l < 8; l++) {
                        blk[g * 24 + 0 * 8 + l] += data[g].color[l].x;
                        blk[g * 24 + 1 * 8 + l] += data[g].color[l].y;
                        blk[g * 24 + 2 * 8 + l] += data[g].color[l].z;
                    }
                store_block(width, NULL, gbuffer, xb, yb, data, 0);
            } else {
                store_block(width, bitmap, gbuffer, xb, yb, data, 1);
            }
        }
    return 0;
}

/* PIXEL mode over a rectangle of whole 16x4 blocks.  Every pixel draws from
 * its own stream, so the result does not depend on how the frame is cut
 * into patches; this lets PIXEL frames have heights that are multiples of 4
 * (an extension: the reference's RenderBuffer asks 16, renderer.rs:45-47). */
static void pixel_rect_impl(const or_scene* s, uint32_t width, uint32_t height, uint32_t* bitmap,
                            float* hdr, uint32_t* gbuffer, uint32_t x0, uint32_t y0, uint32_t w,
                            uint32_t h, uint32_t frame, work_counts* wc) {
    rng8 unused;
    memset(&unused, 0, sizeof(unused));
    unused.mode = OR_RNG_PIXEL;
    for (uint32_t xb = x0; xb < x0 + w; xb += 16)
        for (uint32_t yb = y0; yb < y0 + h; yb += 4) {
            pixel_data8 data[8];
            render_block(s, width, height, xb, yb, frame, &unused, OR_RNG_PIXEL, data, wc);
            if (hdr) {
                float* blk = hdr + ((size_t)(yb / 4) * (width / 16) + xb / 16) * 192;
                for (int g = 0; g < 8; g++)
                    for (int l = 0; l < 8; l++) {
                        blk[g * 24 + 0 * 8 + l] += data[g].color[l].x;
                        blk[g * 24 + 1 * 8 + l] += data[g].color[l].y;
                        blk[g * 24 + 2 * 8 + l] += data[g].color[l].z;
                    }
                store_block(width, NULL, gbuffer, xb, yb, data, 0);
            } else {
                store_block(width, bitmap, gbuffer, xb, yb, data, 1);
            }
        }
}

int or_render_patch_u8(const or_scene* s, uint32_t width, uint32_t height, uint32_t* bitmap,
                       uint32_t* gbuffer, uint32_t patch_width, uint32_t x, uint32_t y,
                       uint32_t frame, int rng_mode) {
    return patch_impl(s, width, height, bitmap, NULL, gbuffer, patch_width, x, y, frame, rng_mode,
                      NULL);
}

int or_accumulate_patch_f32(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                            uint32_t* gbuffer, uint32_t patch_width, uint32_t x, uint32_t y,
                            uint32_t frame, int rng_mode) {
    return patch_impl(s, width, height, NULL, hdr, gbuffer, patch_width, x, y, frame, rng_mode,
                      NULL);
}

int or_buffer_f32_into_bitmap(uint32_t width, uint32_t height, const float* hdr,
                              uint32_t* bitmap, uint32_t num_samples) {
    if ((width & 15) || (height & 3)) return -10;
    uint32_t w = width / 16, h = height / 4;
    float factor = 1.0f / (float)num_samples;
    for (uint32_t j = 0; j < h; j++)
        for (uint32_t i = 0; i < w; i++) {
            const float* blk = hdr + ((size_t)j * w + i) * 192;
            pixel_data8 d[8];
            for (int g = 0; g < 8; g++)
                for (int l = 0; l < 8; l++)
                    d[g].color[l] = mk3(blk[g * 24 + l] * factor, blk[g * 24 + 8 + l] * factor,
                                        blk[g * 24 + 16 + l] * factor);
            store_block(width, bitmap, NULL, i * 16, j * 4, d, 1);
        }
    return 0;
}

void or_hdr_blocks_to_rows(uint32_t width, uint32_t height, const float* hdr, float* rows) {
    uint32_t w = width / 16, h = height / 4;
    for (uint32_t j = 0; j < h; j++)
        for (uint32_t i = 0; i < w; i++) {
            const float* blk = hdr + ((size_t)j * w + i) * 192;
            for (int g = 0; g < 8; g++)
                for (int l = 0; l < 8; l++) {
                    uint32_t dx, dy;
                    block_pixel(g, l, &dx, &dy);
                    size_t p = (size_t)(j * 4 + dy) * width + (i * 16 + dx);
                    rows[3 * p] = blk[g * 24 + l];
                    rows[3 * p + 1] = blk[g * 24 + 8 + l];
                    rows[3 * p + 2] = blk[g * 24 + 16 + l];
                }
        }
}

/* Patch loop of main.rs:174-214 over a worker pool. */
typedef struct {
    const or_scene* s;
    uint32_t width, height;
    uint32_t* bitmap;
    float* hdr;
    uint32_t* gbuffer;
    uint32_t first_frame, n_frames;
    int rng_mode;
    uint32_t npx, npy;
    volatile uint32_t next;
    pthread_mutex_t lock;
    int status;
    work_counts wc;
} frame_job;

static void patch_geometry(const frame_job* j, uint32_t p, uint32_t* x, uint32_t* y, uint32_t* pw) {
    /* columns outer, rows inner (main.rs:180-181) */
    uint32_t i = p / j->npy, jj = p % j->npy;
    *x = i * 32;
    *y = jj * 32;
    uint32_t wrem = j->width - *x, hrem = j->height - *y;
    *pw = (wrem >= 32 && hrem >= 32) ? 32 : 16;
}

static void* frame_worker(void* arg) {
    frame_job* j = (frame_job*)arg;
    work_counts local = {0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&j->lock);
        uint32_t p = j->next++;
        pthread_mutex_unlock(&j->lock);
        if (p >= j->npx * j->npy) break;
        uint32_t x, y, pw;
        patch_geometry(j, p, &x, &y, &pw);
        /* edge patches narrower than 32 are tiled with 16-wide patches */
        uint32_t xe = x + 32 <= j->width ? x + 32 : j->width;
        uint32_t ye = y + 32 <= j->height ? y + 32 : j->height;
        for (uint32_t f = 0; f < j->n_frames; f++) {
            uint32_t frame = j->first_frame + f;
            if (j->rng_mode == OR_RNG_PIXEL) {
                pixel_rect_impl(j->s, j->width, j->height, j->bitmap, j->hdr, j->gbuffer, x, y,
                                xe - x, ye - y, frame, &local);
                continue;
            }
            for (uint32_t yy = y; yy < ye; yy += pw)
                for (uint32_t xx = x; xx < xe; xx += pw) {
                    int st = patch_impl(j->s, j->width, j->height, j->bitmap, j->hdr, j->gbuffer,
                                        pw, xx, yy, frame, j->rng_mode, &local);
                    if (st) j->status = st;
                }
        }
    }
    pthread_mutex_lock(&j->lock);
    j->wc.queries += local.queries;
    j->wc.sum_aabb += local.sum_aabb;
    j->wc.sum_tri += local.sum_tri;
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

static int run_frames(frame_job* j, int nthreads) {
    /* REF_PATCH frames are whole 16x16 patches; PIXEL frames whole 16x4 blocks */
    if ((j->width & 15) || (j->height & (j->rng_mode == OR_RNG_PIXEL ? 3u : 15u))) return -10;
    if (j->s->ndirect < 8) return -12;
    j->npx = (j->width + 31) / 32;
    j->npy = (j->height + 31) / 32;
    j->next = 0;
    j->status = 0;
    memset(&j->wc, 0, sizeof(j->wc));
    pthread_mutex_init(&j->lock, NULL);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, frame_worker, j);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j->lock);
    return j->status;
}

int or_accumulate_frames(const or_scene* s, uint32_t width, uint32_t height, float* hdr,
                         uint32_t* gbuffer, uint32_t first_frame, uint32_t n_frames, int rng_mode,
                         int nthreads) {
    frame_job j;
    memset(&j, 0, sizeof(j));
    j.s = s;
    j.width = width;
    j.height = height;
    j.hdr = hdr;
    j.gbuffer = gbuffer;
    j.first_frame = first_frame;
    j.n_frames = n_frames;
    j.rng_mode = rng_mode;
    return run_frames(&j, nthreads);
}

int or_render_frame_u8(const or_scene* s, uint32_t width, uint32_t height, uint32_t* bitmap,
                       uint32_t* gbuffer, uint32_t frame, int rng_mode, int nthreads) {
    frame_job j;
    memset(&j, 0, sizeof(j));
    j.s = s;
    j.width = width;
    j.height = height;
    j.bitmap = bitmap;
    j.gbuffer = gbuffer;
    j.first_frame = frame;
    j.n_frames = 1;
    j.rng_mode = rng_mode;
    return run_frames(&j, nthreads);
}

int or_frame_work(const or_scene* s, uint32_t width, uint32_t height, uint32_t frame, int rng_mode,
                  int nthreads, uint64_t* n_queries, uint64_t* sum_aabb, uint64_t* sum_tri) {
    frame_job j;
    memset(&j, 0, sizeof(j));
    j.s = s;
    j.width = width;
    j.height = height;
    size_t np = (size_t)width * height;
    j.hdr = (float*)calloc(np * 3 + 192, sizeof(float));
    j.first_frame = frame;
    j.n_frames = 1;
    j.rng_mode = rng_mode;
    int st = run_frames(&j, nthreads);
    free(j.hdr);
    *n_queries = j.wc.queries;
    *sum_aabb = j.wc.sum_aabb;
    *sum_tri = j.wc.sum_tri;
    return st;
}

/* A bounded CPU-baseline sample: renders the listed 32x32 patches (indices
 * into the column-major patch order of main.rs:174-203) of one frame with
 * nthreads workers and returns the work done. */
typedef struct {
    const or_scene* s;
    uint32_t width, height, frame;
    int rng_mode;
    const uint32_t* ids;
    uint32_t nids;
    float* hdr;
    volatile uint32_t next;
    pthread_mutex_t lock;
    work_counts wc;
    int status;
} sample_job;

static void* sample_worker(void* arg) {
    sample_job* j = (sample_job*)arg;
    work_counts local = {0, 0, 0};
    uint32_t npy = (j->height + 31) / 32;
    for (;;) {
        pthread_mutex_lock(&j->lock);
        uint32_t k = j->next++;
        pthread_mutex_unlock(&j->lock);
        if (k >= j->nids) break;
        uint32_t p = j->ids[k];
        uint32_t x = (p / npy) * 32, y = (p % npy) * 32;
        uint32_t pw = (x + 32 <= j->width && y + 32 <= j->height) ? 32 : 16;
        uint32_t xe = x + 32 <= j->width ? x + 32 : j->width;
        uint32_t ye = y + 32 <= j->height ? y + 32 : j->height;
        for (uint32_t yy = y; yy < ye; yy += pw)
            for (uint32_t xx = x; xx < xe; xx += pw) {
                int st = patch_impl(j->s, j->width, j->height, NULL, j->hdr, NULL, pw, xx, yy,
                                    j->frame, j->rng_mode, &local);
                if (st) j->status = st;
            }
    }
    pthread_mutex_lock(&j->lock);
    j->wc.queries += local.queries;
    j->wc.sum_aabb += local.sum_aabb;
    j->wc.sum_tri += local.sum_tri;
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

int or_sample_patches(const or_scene* s, uint32_t width, uint32_t height, uint32_t frame,
                      int rng_mode, int nthreads, const uint32_t* patch_ids, uint32_t n_ids,
                      uint64_t* out3) {
    if ((width & 15) || (height & 15)) return -10;
    if (s->ndirect < 8) return -12;
    sample_job j;
    memset(&j, 0, sizeof(j));
    j.s = s;
    j.width = width;
    j.height = height;
    j.frame = frame;
    j.rng_mode = rng_mode;
    j.ids = patch_ids;
    j.nids = n_ids;
    j.hdr = (float*)calloc((size_t)width * height * 3, sizeof(float));
    pthread_mutex_init(&j.lock, NULL);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, sample_worker, &j);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j.lock);
    free(j.hdr);
    out3[0] = j.wc.queries;
    out3[1] = j.wc.sum_aabb;
    out3[2] = j.wc.sum_tri;
    return j.status;
}
