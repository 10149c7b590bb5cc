// device_math.hpp -- gfx950 device helpers that reproduce the reference's
// AVX arithmetic bit for bit (ruuda/convector src/simd.rs, src/vector3.rs).
//
// Compiled with -ffp-contract=off: a product is fused only where the
// reference calls mul_add / mul_sub / neg_mul_add (simd.rs:81-94), written
// here as fmaf.  Division and sqrt are the correctly rounded HIP defaults.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvk {

constexpr uint32_t kNoHit = 0xffffffffu;
constexpr uint32_t kMatSky = 0xC0000000u;  // SMaterial::sky (material.rs:85-90)
constexpr uint32_t kNegNaN = 0xFFC00000u;  // x86 "real indefinite" default NaN

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t sgn(float f) { return __float_as_uint(f) >> 31; }

// x86 minps / maxps (simd.rs:318-326): the second operand on NaN or +-0 ties.
__device__ __forceinline__ float xmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float xmax(float a, float b) { return a > b ? a : b; }
// cmpps NLT_UQ (simd.rs:329-333) as a sign-bit mask.
__device__ __forceinline__ uint32_t geq(float a, float b) { return (a < b) ? 0u : 0xffffffffu; }
// blendvps (simd.rs:376-378): b when the mask's sign bit is set.
__device__ __forceinline__ float pickf(float a, float b, uint32_t m) { return (m >> 31) ? b : a; }
// Mf32 negation is 0 - x (simd.rs:576-588).
__device__ __forceinline__ float neg0(float x) { return 0.0f - x; }
__device__ __forceinline__ float absb(float x) { return u2f(f2u(x) & 0x7fffffffu); }
// A NaN produced by an invalid operation is negative on x86.
__device__ __forceinline__ float x86nan(float x) { return x != x ? u2f(kNegNaN) : x; }
// Sign bit of the x86 difference a - b, NaN propagation included.
__device__ __forceinline__ uint32_t sgn_x86_sub(float a, float b) {
  float r = a - b;
  if (r == r) return sgn(r);
  if (a != a) return sgn(a);
  if (b != b) return sgn(b);
  return 1u;
}

// CV_APPROX_PORTABLE: recip_fast and rsqrt as correctly rounded operations.
__device__ __forceinline__ float rcp(float x) { return 1.0f / x; }
__device__ __forceinline__ float rsqrt(float x) { return 1.0f / sqrtf(x); }

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 scale3(V3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 mulc3(V3 a, V3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 neg3(V3 a) { return mk3(neg0(a.x), neg0(a.y), neg0(a.z)); }
// MVector3::dot = dot_naive (vector3.rs:201-217)
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// MVector3::cross = cross_fma (vector3.rs:185-199)
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return mk3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
             fmaf(a.x, b.y, -(a.y * b.x)));
}
// MVector3::mul_add / neg_mul_add (vector3.rs:254-269)
__device__ __forceinline__ V3 muladd3(V3 a, float f, V3 o) {
  return mk3(fmaf(a.x, f, o.x), fmaf(a.y, f, o.y), fmaf(a.z, f, o.z));
}
__device__ __forceinline__ V3 negmuladd3(V3 a, float f, V3 o) {
  return mk3(fmaf(-a.x, f, o.x), fmaf(-a.y, f, o.y), fmaf(-a.z, f, o.z));
}
// MVector3::normalized (vector3.rs:290-306)
__device__ __forceinline__ V3 normalized3(V3 a) { return scale3(a, rsqrt(dot3(a, a))); }
__device__ __forceinline__ V3 pick3(V3 a, V3 b, uint32_t m) { return (m >> 31) ? b : a; }

// Layout of one BVH node on the device: the reference's 32-byte BvhNode
// (bvh.rs:20-30) as two float4: (min.xyz, max.x), (max.y, max.z, index, len).
struct DevNode {
  float4 a;
  float4 b;
};

// Triangle geometry in BVH order, three float4 per triangle:
// (v0.xyz, n.x), (e1.xyz, n.y), (e2.xyz, n.z) with e1 = v0 - v2,
// e2 = v1 - v0, n = e1 x e2 -- exactly the values Triangle::intersect
// computes per test (triangle.rs:65-80).
struct DevTri {
  float4 p;
  float4 q;
  float4 r;
};

// Light triangle data for direct sampling and its pdf (scene.rs:143-202,
// triangle.rs:152-185), precomputed with the same operations.
struct DevLight {
  float v0[3];
  float e1[3];
  float e2[3];
  float n[3];
  float normal[3];  // n * rnorm
  float area;       // 0.5 * rcp(rnorm)
};

// Offset of a pixel in the 16x4-block HDR layout (renderer.rs:154-206,356-358).
__device__ __forceinline__ size_t hdr_offset(uint32_t px, uint32_t py, uint32_t width) {
  uint32_t dx = px & 15u, dy = py & 3u;
  uint32_t g = (dx >> 2) * 2 + (dy >> 1);
  uint32_t l = (dx & 3u) + 4 * (dy & 1u);
  size_t blk = (size_t)(py >> 2) * (width >> 4) + (px >> 4);
  return blk * 192 + g * 24 + l;
}

struct Hit {
  uint32_t tri;
  float t, u, v;
};

}  // namespace cvk
