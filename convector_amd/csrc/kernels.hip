// kernels.hip -- hand-written gfx950 kernels of the per-ray path:
// BVH traversal with triangle intersection (extend), camera rays
// (generate), one-sample-MIS shading (shade) and the output conversions.
//
// One ray per lane of a wave64.  Paths live in HBM as structure-of-arrays
// queues; every bounce is an extend launch and a shade launch that appends
// the surviving paths to the next queue with one atomic per block.  Extend
// is persistent: its waves take rays from a device cursor, so no host round
// trip is needed between launches.  A step renders one frame of every job
// in a device job table; the paths of all jobs share the queues.
#include "device_math.hpp"
#include "internal.hpp"

namespace cvk {

// ---------------------------------------------------------------------------
// Traversal (bvh.rs:589-657, aabb.rs:105-183, triangle.rs:59-148)
// ---------------------------------------------------------------------------

// Aabb::intersect with the reciprocals hoisted (aabb.rs:105-141), with the
// minps / maxps selection rules of x86 (the second operand on NaN and on
// +-0 ties).
template <bool EXACT_NAN>
__device__ __forceinline__ void slab(float4 a, float4 b, V3 o, V3 inv, float& tmin, float& tmax) {
  float tx1 = (a.x - o.x) * inv.x, tx2 = (a.w - o.x) * inv.x;
  float ty1 = (a.y - o.y) * inv.y, ty2 = (b.x - o.y) * inv.y;
  float tz1 = (a.z - o.z) * inv.z, tz2 = (b.y - o.z) * inv.z;
  if (EXACT_NAN) {  // 0 * inf gives the negative default NaN on x86
    tx1 = x86nan(tx1);
    tx2 = x86nan(tx2);
    ty1 = x86nan(ty1);
    ty2 = x86nan(ty2);
    tz1 = x86nan(tz1);
    tz2 = x86nan(tz2);
  }
  float txmin = xmin(tx1, tx2), txmax = xmax(tx1, tx2);
  float tymin = xmin(ty1, ty2), tymax = xmax(ty1, ty2);
  float tzmin = xmin(tz1, tz2), tzmax = xmax(tz1, tz2);
  tmin = xmax(txmin, xmax(tymin, tzmin));
  tmax = xmin(txmax, xmin(tymax, tzmax));
}

// The same slab test with the hardware v_min / v_max (v_min3 / v_max3).
// Valid only for NaN-free slab distances (finite origin and reciprocal):
// then every min / max chain has the same value as the x86 chain and can
// differ from it only in the sign of a zero result.
__device__ __forceinline__ void slab_hw(float4 a, float4 b, V3 o, V3 inv, float& tmin,
                                        float& tmax) {
  float tx1 = (a.x - o.x) * inv.x, tx2 = (a.w - o.x) * inv.x;
  float ty1 = (a.y - o.y) * inv.y, ty2 = (b.x - o.y) * inv.y;
  float tz1 = (a.z - o.z) * inv.z, tz2 = (b.y - o.z) * inv.z;
  tmin = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
  tmax = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2));
}

struct TravState {
  float best_t;
  uint32_t best;
  float bu, bv;
  uint32_t na, nt;
};

// Stack entries are 64-bit words: x = node.index | node.len << 27 (PACKED)
// or the node id (PACKED == false, index/len read at pop time) in the low
// half, the tmin bits in the high half.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
typedef __attribute__((address_space(1))) unsigned long long glb_u64;
typedef unsigned long long entry_t;

__device__ __forceinline__ entry_t pack_entry(uint32_t x, float tmin) {
  return (entry_t)x | ((entry_t)f2u(tmin) << 32);
}
__device__ __forceinline__ uint32_t entry_x(entry_t e) { return (uint32_t)e; }
__device__ __forceinline__ float entry_tmin(entry_t e) { return u2f((uint32_t)(e >> 32)); }

// Per-lane stack: S entries in LDS (stride 64 lanes, bank-conflict free),
// deeper entries in a global spill area (stride = all persistent lanes).
// The pointers are address-space qualified so that LDS accesses compile to
// ds_read_b64 / ds_write_b64 and spill accesses to global loads / stores.
// The top entry (sp - 1) is held in a register and entries [0, sp - 1) in
// memory: a pop returns the register at once and reloads the next top,
// which is not needed before the next pop, so no step waits on the stack
// read; a push stores the previous top.
template <int S, bool PACKED>
struct Stack {
  lds_u64* lds;
  glb_u64* spill;
  uint32_t stride;
  int sp;
  entry_t top;
  __device__ __forceinline__ Stack(uint2* l, uint2* g, uint32_t str, int p)
      : lds((lds_u64*)l), spill((glb_u64*)g), stride(str), sp(p), top(0) {}
  // Pushes e when p is set.  The previous top goes to slot sp - 1, which
  // lies above the memory part of the stack, so when p is not set (or the
  // stack is empty, slot 0) the LDS write is harmless and is done anyway.
  // The LDS-only case is tested for the whole wave, so the common path has
  // no divergent branch.
  __device__ __forceinline__ void push_if(bool p, entry_t e) {
    const int k = sp > 0 ? sp - 1 : 0;
    if (__builtin_expect(!__any(k >= S), 1)) {
      lds[k * 64] = top;
    } else if (k < S) {
      lds[k * 64] = top;
    } else if (p) {
      spill[(size_t)(k - S) * stride] = top;
    }
    if (p) {
      top = e;
      sp++;
    }
  }
  __device__ __forceinline__ entry_t pop() {
    const entry_t e = top;
    sp--;
    const int k = sp > 0 ? sp - 1 : 0;  // the new top (slot 0 when empty: unused)
    if (__builtin_expect(!__any(k >= S), 1))
      top = lds[k * 64];
    else if (k < S)
      top = lds[k * 64];
    else
      top = spill[(size_t)(k - S) * stride];
    return e;
  }
};

// Both children of an inner node (bvh.rs:604-631).  The reference pushes the
// hit children, the one that should be tried later first, and the next loop
// iteration pops the last push.  Here the last push stays in registers as
// `cur` (have = true) and only the other child goes to the stack, which
// visits the nodes in exactly the reference's order.
template <int S, bool PACKED>
__device__ __forceinline__ void visit_pair(const DevNode* __restrict__ nodes, uint32_t base, V3 o,
                                           V3 inv, bool exact, Stack<S, PACKED>& st,
                                           entry_t& cur, bool& have) {
  const float4* p = reinterpret_cast<const float4*>(nodes + base);
  float4 a0 = p[0], b0 = p[1], a1 = p[2], b1 = p[3];
  float tmin0, tmax0, tmin1, tmax1;
  slab_hw(a0, b0, o, inv, tmin0, tmax0);
  slab_hw(a1, b1, o, inv, tmin1, tmax1);
  // Zero results are where the signs of the hardware min / max may differ
  // from minps / maxps; the sign of tmax decides the hit and the signs of two
  // zero tmins decide the order (tmax0 == 0 | tmax1 == 0 | (tmin0 == 0 &
  // tmin1 == 0), as one min3).  Those, and waves with a lane whose slab
  // distances can be NaN (wave-uniform `exact`), redo the x86 chains.
  const float z = fminf(fminf(fabsf(tmax0), fabsf(tmax1)), fmaxf(fabsf(tmin0), fabsf(tmin1)));
  if (__builtin_expect(exact || z == 0.0f, 0)) {
    if (exact) {
      slab<true>(a0, b0, o, inv, tmin0, tmax0);
      slab<true>(a1, b1, o, inv, tmin1, tmax1);
    } else {
      slab<false>(a0, b0, o, inv, tmin0, tmax0);
      slab<false>(a1, b1, o, inv, tmin1, tmax1);
    }
  }
  // sign bit of tmin0 - tmin1 as x86 computes it: inf - inf (and NaN inputs)
  // give the negative default NaN.  Without NaN inputs the class test (mask
  // 0x3f: NaN, -inf, -normal, -denormal, -0) is that sign bit.
  const bool swap = exact ? sgn_x86_sub(tmin0, tmin1) != 0
                          : __builtin_amdgcn_classf(tmin0 - tmin1, 0x3f);
  // any_masked: mask (tmax NLT tmin) and tmax's sign bit 0 (aabb.rs:155-158)
  const bool h0 = !(tmax0 < tmin0) && sgn(tmax0) == 0;
  const bool h1 = !(tmax1 < tmin1) && sgn(tmax1) == 0;
  uint32_t x0, x1;
  if (PACKED) {
    x0 = f2u(b0.z) | (f2u(b0.w) << 27);
    x1 = f2u(b1.z) | (f2u(b1.w) << 27);
  } else {
    x0 = base;
    x1 = base + 1;
  }
  // should_try_before (bvh.rs:616-622): without swap child 0 is pushed first
  // and child 1 is popped first.  The first push goes to the stack when both
  // children are hit; the entry popped next (the second push if it was hit,
  // else the first) stays in registers: child 0 when h0 and (swap or !h1).
  // (when neither child is hit `cur` is not used, so c0 may be false then)
  const bool c0 = h0 & (swap | !h1);
  cur = pack_entry(c0 ? x0 : x1, c0 ? tmin0 : tmin1);
  have = h0 | h1;
  st.push_if(h0 && h1, swap ? pack_entry(x1, tmin1) : pack_entry(x0, tmin0));
}

// Triangle::intersect, keeping only what decides the pick (triangle.rs:59-127).
__device__ __forceinline__ void test_tri(const DevTri* __restrict__ tris, uint32_t k, V3 o, V3 d,
                                         TravState& ts) {
  const float4* tp = reinterpret_cast<const float4*>(tris + k);
  float4 P = tp[0], Q = tp[1], R = tp[2];
  V3 v0 = mk3(P.x, P.y, P.z);
  V3 n = mk3(P.w, Q.w, R.w);
  V3 e1 = mk3(Q.x, Q.y, Q.z);
  V3 e2 = mk3(R.x, R.y, R.z);
  V3 fr = sub3(v0, o);
  float denom = 1.0f / dot3(d, n);
  float t = dot3(fr, n) * denom;
  V3 c = cross3(d, fr);
  float u = dot3(c, e2) * denom;
  float v = dot3(c, e1) * denom;
  float w = x86nan((1.0f - u) - v);
  uint32_t mp = (f2u(t) | f2u(u)) | (f2u(v) | f2u(w));
  if ((mp >> 31) == 0 && t < ts.best_t) {
    ts.best_t = t;
    ts.best = k;
    ts.bu = u;
    ts.bv = v;
  }
}

// One iteration of the traversal loop (bvh.rs:597-651): take the next
// entry (the one kept in registers, else the top of the stack), skip it when
// it is further away than the best hit (is_further_away_than,
// aabb.rs:161-170), then visit an inner node's children or test a leaf's
// triangles.  The ray is finished when have == false and the stack is empty.
// The root is the entry (index 0, len 0, tmin -inf), so that its pair is
// visited first and counted like every other pair.
//
// With `waiting` (lane refill), a lane that reaches a leaf outside a leaf
// round keeps the leaf's entry and waits (*waiting); it does nothing until
// the wave's next leaf round, when it takes the same entry again (the same
// cull test, the same best hit) and tests the triangles.  Each lane thus
// performs the same operations in the same order, only at later steps, while
// the wave runs its triangle tests for many lanes at once instead of for the
// few that reach a leaf in any one step.
template <bool COUNT, int S, bool PACKED>
__device__ __forceinline__ void trav_step(const DevNode* __restrict__ nodes,
                                          const DevTri* __restrict__ tris, V3 o, V3 d, V3 inv,
                                          bool exact, Stack<S, PACKED>& st, TravState& ts,
                                          entry_t& cur, bool& have, bool leaf_round = true,
                                          bool* waiting = nullptr) {
  if (waiting && *waiting && !leaf_round) return;
  entry_t e = cur;
  if (!have) e = st.pop();
  have = false;
  if (waiting) *waiting = false;
  if (!(entry_tmin(e) < ts.best_t)) return;
  const uint32_t x = entry_x(e);
  uint32_t idx, len;
  if (PACKED) {
    idx = x & 0x07ffffffu;
    len = x >> 27;
  } else if (x == 0xffffffffu) {
    idx = 0;
    len = 0;
  } else {
    float4 nb = reinterpret_cast<const float4*>(nodes + x)[1];
    idx = f2u(nb.z);
    len = f2u(nb.w);
  }
  if (len == 0) {
    if (COUNT) ts.na += 2;
    visit_pair(nodes, idx, o, inv, exact, st, cur, have);
  } else {
    if (waiting && !leaf_round) {
      cur = e;
      have = true;
      *waiting = true;
      return;
    }
    for (uint32_t k = idx; k < idx + len; k++) {
      test_tri(tris, k, o, d, ts);
      if (COUNT) ts.nt++;
    }
  }
}

template <bool PACKED>
__device__ __forceinline__ entry_t root_entry() {
  return pack_entry(PACKED ? 0u : 0xffffffffu, -INFINITY);
}

// A lane needs the exact NaN rules when a slab distance can be NaN: a
// non-finite origin or reciprocal (0 * inf, inf - inf).
__device__ __forceinline__ bool needs_exact_nan(V3 o, V3 inv) {
  return !(isfinite(inv.x) && isfinite(inv.y) && isfinite(inv.z) && isfinite(o.x) &&
           isfinite(o.y) && isfinite(o.z));
}

// Nearest hit of one ray seeded with the far-away sky hit at 1e5
// (scene.rs:220-230).
template <bool COUNT, int S, bool PACKED>
__device__ __forceinline__ TravState traverse(const DevNode* __restrict__ nodes,
                                              const DevTri* __restrict__ tris, V3 o, V3 d,
                                              uint2* lds, uint2* spill, uint32_t stride,
                                              bool valid) {
  TravState ts;
  ts.best_t = 1.0e5f;
  ts.best = kNoHit;
  ts.bu = ts.bv = 0.0f;
  ts.na = 0;
  ts.nt = 0;
  V3 inv = mk3(rcp(d.x), rcp(d.y), rcp(d.z));
  Stack<S, PACKED> st(lds, spill, stride, 0);
  const bool exact = needs_exact_nan(o, inv);
  if (!valid) return ts;
  entry_t cur = root_entry<PACKED>();
  bool have = true;
  const bool wave_exact = __any(exact);
  while (have || st.sp > 0) trav_step<COUNT>(nodes, tris, o, d, inv, wave_exact, st, ts, cur, have);
  return ts;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-wide work fetch: lane 0 takes `count` entries from the cursor.  The
// result is wave-uniform and held in a scalar register.
__device__ __forceinline__ uint32_t fetch(uint32_t* cursor, uint32_t count) {
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(cursor, count);
  return __builtin_amdgcn_readfirstlane(__shfl(base, 0));
}

// One index below `limit` from a cursor that only grows, or `limit` when it
// has reached it: a share seen used up costs a load, not an atomic.
__device__ __forceinline__ uint32_t fetch_below(uint32_t* cursor, uint32_t limit) {
  uint32_t k = limit;
  if (lane_id() == 0 &&
      __hip_atomic_load(cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < limit)
    k = atomicAdd(cursor, 1u);
  return __shfl(k, 0);
}

// ---------------------------------------------------------------------------
// cv_bvh_intersect_nearest
// ---------------------------------------------------------------------------

// The n / 64 batches of 64 rays are split into kIsectGroups contiguous
// shares, each with its own cursor on its own cache line: one cursor for
// every batch serialises on that address (about 12 ns an atomic, 3 ms for
// 2^24 rays whatever the scene).  A wave starts on share blockIdx %
// kIsectGroups and moves on to the next share when its own is used up;
// shares only ever get used up, so a wave that has found kIsectGroups of
// them in a row used up has seen every share used up and ends.
template <bool COUNT, int S, bool PACKED>
__global__ __launch_bounds__(64) void k_intersect(IntersectArgs a) {
  __shared__ uint2 lds[S * 64];
  const uint32_t lane = lane_id();
  const uint32_t gtid = blockIdx.x * 64 + lane;
  const uint32_t stride = gridDim.x * 64;
  const uint64_t nb = (a.n + 63u) / 64u;
  uint32_t g = blockIdx.x % kIsectGroups;
  for (uint32_t left = kIsectGroups; left;) {
    const uint32_t lo = (uint32_t)(nb * g / kIsectGroups);
    const uint32_t len = (uint32_t)(nb * (g + 1) / kIsectGroups) - lo;
    const uint32_t k = fetch_below(a.cursor + g * kIsectCursorStride, len);
    if (k >= len) {
      g = g + 1 == kIsectGroups ? 0u : g + 1;
      --left;
      continue;
    }
    uint32_t i = (lo + k) * 64u + lane;
    bool valid = i < a.n;
    V3 o = mk3(0, 0, 0), d = mk3(1, 0, 0);
    if (valid) {
      const float* r = a.rays + 6 * (size_t)i;
      o = mk3(r[0], r[1], r[2]);
      d = mk3(r[3], r[4], r[5]);
    }
    TravState ts = traverse<COUNT, S, PACKED>(a.nodes, a.tris, o, d, lds + lane,
                                              a.spill + gtid, stride, valid);
    if (!valid) continue;
    uint32_t src = ts.best == kNoHit ? kNoHit : a.tri_src[ts.best];
    if (a.hits) {
      reinterpret_cast<uint4*>(a.hits)[i] =
          make_uint4(src, f2u(ts.best_t), f2u(ts.bu), f2u(ts.bv));
    }
    if (a.isects) {
      // The deferred MIntersection fields (triangle.rs:129-142, scene.rs:221-228).
      V3 pos, nrm;
      uint32_t mat;
      float tu = 0.0f, tv = 0.0f;
      if (ts.best == kNoHit) {
        pos = muladd3(d, 1.0e5f, o);
        nrm = d;
        mat = kMatSky;
      } else {
        const float4* tp = reinterpret_cast<const float4*>(a.tris + ts.best);
        V3 n = mk3(tp[0].w, tp[1].w, tp[2].w);
        pos = muladd3(d, ts.best_t, o);
        nrm = normalized3(n);
        mat = a.tri_mat[ts.best];
        const float* uv = a.tri_uv + 6 * (size_t)ts.best;
        float w = (1.0f - ts.bu) - ts.bv;
        tu = fmaf(uv[0], w, fmaf(uv[2], ts.bv, uv[4] * ts.bu));
        tv = fmaf(uv[1], w, fmaf(uv[3], ts.bv, uv[5] * ts.bu));
      }
      uint32_t* out = a.isects + 16 * (size_t)i;
      out[0] = f2u(pos.x);
      out[1] = f2u(pos.y);
      out[2] = f2u(pos.z);
      out[3] = f2u(nrm.x);
      out[4] = f2u(nrm.y);
      out[5] = f2u(nrm.z);
      out[6] = f2u(ts.best_t);
      out[7] = mat;
      out[8] = f2u(tu);
      out[9] = f2u(tv);
      out[10] = src;
      out[11] = f2u(ts.bu);
      out[12] = f2u(ts.bv);
      out[13] = ts.na;
      out[14] = ts.nt;
      out[15] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Path state helpers
// ---------------------------------------------------------------------------

// PIXEL stream: splitmix64 premix, xorshift64 (13, 7, 17), high 32 bits.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t pixel_seed(uint32_t pixel, uint32_t frame) {
  uint64_t s = splitmix64((uint64_t)pixel ^ ((uint64_t)frame << 32));
  return s ? s : 0x9E3779B97F4A7C15ull;
}
__device__ __forceinline__ uint32_t rng_next(uint64_t& s) {
  uint64_t x = s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  s = x;
  return (uint32_t)(x >> 32);
}
// Rng conversions (random.rs:104-136)
__device__ __forceinline__ float u32_unit(uint32_t u) {
  return fmaf((float)(int32_t)u, 0.5f / -2147483648.0f, 0.5f);
}
__device__ __forceinline__ float u32_biunit(uint32_t u) {
  return (float)(int32_t)u * (1.0f / -2147483648.0f);
}
__device__ __forceinline__ float u32_angle(uint32_t u) {
  return (float)(int32_t)u * (3.14159274f / -2147483648.0f);
}

// Polynomial sine and cosine (simd.rs:190-233).
__device__ __forceinline__ float sin_poly(float x) {
  float x2 = x * x;
  float x3 = x * x2;
  float x5 = x3 * x2;
  return fmaf(x5, 0.00533738081738135917f, fmaf(x3, -0.152175915294f, x * 0.982006669405f));
}
__device__ __forceinline__ float cos_poly(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  float x6 = x2 * x4;
  return fmaf(x6, -0.000966231179636657107f,
              fmaf(x4, 0.0392596924214f, fmaf(x2, -0.496000299455f, 1.0f)));
}

// MVector3::rotate_hemisphere (vector3.rs:219-251)
__device__ __forceinline__ V3 rotate_hemisphere(V3 v, V3 normal) {
  uint32_t flip = f2u(normal.z);
  V3 n = pick3(normal, neg3(normal), flip);
  float rz = rcp(1.0f + n.z);
  float c = (n.x * n.y) * rz;
  float x = fmaf(v.x, fmaf(n.y, n.y * rz, n.z), -fmaf(v.y, c, -(v.z * n.x)));
  float y = fmaf(-v.x, c, fmaf(v.y, fmaf(n.x, n.x * rz, n.z), v.z * n.y));
  float z = fmaf(-v.x, n.x, fmaf(-v.y, n.y, v.z * n.z));
  V3 r = mk3(x, y, z);
  return pick3(r, neg3(r), flip);
}

// MQuaternion::interpolate + quaternion::rotate (quaternion.rs:62-153)
__device__ __forceinline__ V3 camera_dir(const CameraParams& cam, float x, float y, float t) {
  float a = fmaf(cam.dq[0], t, cam.q[0]);
  float b = fmaf(cam.dq[1], t, cam.q[1]);
  float c = fmaf(cam.dq[2], t, cam.q[2]);
  float d = fmaf(cam.dq[3], t, cam.q[3]);
  float ns = fmaf(a, a, b * b) + fmaf(c, c, d * d);
  float rnorm = 1.0f / sqrtf(ns);
  float qa = a * rnorm, qb = b * rnorm, qc = c * rnorm, qd = d * rnorm;
  V3 v = normalized3(mk3(x, y, -cam.screen_distance));
  float pa = (qb * v.x + qc * v.y) + qd * v.z;
  float pb = (((qb * v.x - qc * v.y) + qd * v.z) + (qa - qb) * v.x) + (qc - qd) * (v.y + v.z);
  float pc = (qd * v.x + qa * v.y) - qb * v.z;
  float pd = (((qd * v.x - qa * v.y) - qb * v.z) - (qc + qd) * v.x) + (qa + qb) * (v.y + v.z);
  float rb = (((((pa + pb) * (qa + qb) - (pc - pd) * (qc + qd)) - pa * qa) - pb * qb) + pc * qc) -
             pd * qd;
  float rc = ((pc * qa - pd * qb) + pa * qc) + pb * qd;
  float rd = (((((pc + pd) * (qa + qb) + (pa - pb) * (qc + qd)) - pc * qa) - pd * qb) - pa * qc) +
             pb * qd;
  return mk3(rb, rc, rd);
}

// sky_intensity (material.rs:205-215)
__device__ __forceinline__ V3 sky_intensity(V3 dir) {
  float dd = fmaf(dot3(dir, mk3(0.0f, 0.0f, 1.0f)), 0.5f, 0.5f);
  return muladd3(mk3(dd, dd * dd, dd * (dd * dd)), 2.0f, mk3(0.5f, 0.5f, 0.5f));
}

// cvtps2dq: round to nearest even, NaN/out of range -> 0x80000000.
__device__ __forceinline__ uint32_t cvt_rne(float x) {
  if (!(x >= -2147483648.0f && x < 2147483648.0f)) return 0x80000000u;
  return (uint32_t)(int32_t)rintf(x);
}

// store_pixels_color_16x4 (renderer.rs:239-256)
__device__ __forceinline__ uint32_t pack_color(V3 c) {
  V3 x2 = scale3(c, 2.0f);
  uint32_t r = cvt_rne(xmin(1.0f, x2.x) * 255.0f);
  uint32_t g = cvt_rne(xmin(1.0f, x2.y) * 255.0f) << 8;
  uint32_t b = cvt_rne(xmin(1.0f, x2.z) * 255.0f) << 16;
  return (r | g) | b;
}
// store_pixels_gbuffer_16x4 (renderer.rs:260-286)
__device__ __forceinline__ uint32_t pack_gbuffer(uint32_t tex_index, float tu, float tv, float fr) {
  uint32_t r = cvt_rne(tu * 255.0f) & 0xffu;
  uint32_t g = (cvt_rne(tv * 255.0f) & 0xffu) << 8;
  uint32_t b = cvt_rne(fr * 255.0f) << 16;
  uint32_t a = tex_index << 24;
  return (r | g) | (b | a);
}


// ---------------------------------------------------------------------------
// generate: camera rays (renderer.rs:171-206,413-416, scene.rs:81-99)
// ---------------------------------------------------------------------------

// Job of path p: the last job whose path_base <= p (jobs sorted by
// path_base, the first at 0).
__device__ __forceinline__ uint32_t job_of(const JobDesc* jobs, uint32_t njobs, uint32_t p) {
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (jobs[mid].path_base <= p)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_generate(GenerateArgs a) {
  const uint32_t npaths = a.ctrl->step_paths, njobs = a.ctrl->step_jobs;
  if (blockIdx.x == 0 && threadIdx.x < kMaxBounceSlots) {
    a.ctrl->count[threadIdx.x] = threadIdx.x == 0 ? npaths : 0u;
    a.ctrl->ext_cursor[threadIdx.x] = 0u;
    a.ctrl->shade_cursor[threadIdx.x] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x < kMaxBounceSlots * kParts)
    a.ctrl->part_cursor[threadIdx.x / kParts][(threadIdx.x % kParts) * kPartStride] = 0u;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npaths; p += gridDim.x * blockDim.x) {
    const uint32_t j = job_of(a.jobs, njobs, p);
    const JobDesc& J = a.jobs[j];
    const uint32_t k = p - J.path_base;
    // The rectangle in bands of 8 rows, each walked column by column, so
    // that the 64 paths of a wave are an 8 x 8 tile of camera rays rather
    // than a 64-pixel row strip; the last band holds the rect_h % 8 rows
    // left over.  Images do not depend on the order: every result is
    // keyed by the path's pixel.
    const uint32_t full = J.rect_h & ~7u;
    uint32_t tx, ty;
    if (k < full * J.rect_w) {
      const uint32_t b = k / (8u * J.rect_w), r = k - b * (8u * J.rect_w);
      tx = r >> 3;
      ty = b * 8u + (r & 7u);
    } else {
      const uint32_t hb = J.rect_h - full, r = k - full * J.rect_w;
      tx = r / hb;
      ty = full + r % hb;
    }
    uint32_t px = J.rect_x + tx;
    uint32_t py = J.rect_y + ty;
    uint32_t pixel = py * J.width + px;
    uint32_t xb = px & ~15u, yb = py & ~3u, dx = px & 15u, dy = py & 3u;
    uint32_t g = (dx >> 2) * 2 + (dy >> 1);
    uint32_t l = (dx & 3u) + 4 * (dy & 1u);
    const float scale = 2.0f / (float)J.width;
    const float cx = (float)xb - (float)J.width * 0.5f;
    const float cy = (float)yb - (float)J.height * 0.5f;
    float bx = scale * ((float)(l & 3u) + cx);
    float by = scale * ((float)(l >> 2) + cy);
    float xs = bx, ys = by;
    uint32_t gx = g >> 1;
    if (gx == 1) xs = bx + 4.0f * scale;
    if (gx == 2) xs = bx + 8.0f * scale;
    if (gx == 3) xs = bx + 12.0f * scale;
    if (g & 1u) ys = by + 2.0f * scale;
    uint64_t st = pixel_seed(pixel, J.frame);
    float x = fmaf(u32_unit(rng_next(st)), scale, xs);
    float y = fmaf(u32_unit(rng_next(st)), scale, ys);
    float t = u32_unit(rng_next(st));
    V3 o = muladd3(mk3(J.cam.dpos[0], J.cam.dpos[1], J.cam.dpos[2]), t,
                   mk3(J.cam.pos[0], J.cam.pos[1], J.cam.pos[2]));
    V3 d = camera_dir(J.cam, x, y, t);
    a.q.r0[p] = make_float4(o.x, o.y, o.z, u2f(pixel));
    a.q.r1[p] = make_float4(d.x, d.y, d.z, u2f(j));
    a.q.col[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    a.q.rng[p] = st;
    if (J.gbuffer) J.gbuffer[pixel] = 0u;
    if (J.bitmap) J.bitmap[pixel] = 0u;
  }
}

// ---------------------------------------------------------------------------
// extend: nearest hit of every queued ray
// ---------------------------------------------------------------------------

// A queued ray: the extend queue's float4 pair or, for batched queries, the
// caller's rays as 6 floats (QUERY: known at compile time; otherwise a
// wave-uniform branch).
template <bool QUERY = false>
__device__ __forceinline__ void load_ray(const ExtendArgs& a, uint32_t i, V3& o, V3& d) {
  if (QUERY || a.rays6) {
    const float* r = a.rays6 + 6 * (size_t)i;
    o = mk3(r[0], r[1], r[2]);
    d = mk3(r[3], r[4], r[5]);
  } else {
    const float4 r0 = a.r0[i], r1 = a.r1[i];
    o = mk3(r0.x, r0.y, r0.z);
    d = mk3(r1.x, r1.y, r1.z);
  }
}
// The hit's triangle: its BVH position, or its source index for queries.
template <bool QUERY = false>
__device__ __forceinline__ uint32_t hit_id(const ExtendArgs& a, uint32_t best) {
  return (QUERY || a.tri_src) && best != kNoHit ? a.tri_src[best] : best;
}

template <bool COUNT, int S, bool PACKED>
__global__ __launch_bounds__(64) void k_extend(ExtendArgs a) {
  __shared__ uint2 lds[S * 64];
  const uint32_t lane = lane_id();
  const uint32_t gtid = blockIdx.x * 64 + lane;
  const uint32_t stride = gridDim.x * 64;
  const uint32_t n = a.ctrl->count[a.bounce];
  if (blockIdx.x == 0 && lane == 0) atomicAdd(&a.ctrl->total_rays, (unsigned long long)n);
  uint64_t na = 0, nt = 0;
  const uint32_t chunk = 64u * a.batches;
  for (;;) {
    uint32_t cbase = fetch(&a.ctrl->ext_cursor[a.bounce], chunk);
    if (cbase >= n) break;
    for (uint32_t b = 0; b < a.batches; b++) {
    uint32_t base = cbase + 64u * b;
    if (base >= n) break;
    bool valid = base + lane < n;
    const uint32_t i = valid && a.order ? a.order[base + lane] : base + lane;
    V3 o = mk3(0, 0, 0), d = mk3(1, 0, 0);
    if (valid) load_ray(a, i, o, d);
    TravState ts = traverse<COUNT, S, PACKED>(a.nodes, a.tris, o, d, lds + lane,
                                              a.spill + gtid, stride, valid);
    if (valid) {
      a.hits[i] = make_uint4(hit_id(a, ts.best), f2u(ts.best_t), f2u(ts.bu), f2u(ts.bv));
      if (COUNT) {
        na += ts.na;
        nt += ts.nt;
      }
    }
    }
  }
  if (COUNT) {
    for (int off = 32; off > 0; off >>= 1) {
      na += __shfl_down(na, off);
      nt += __shfl_down(nt, off);
    }
    if (lane == 0) {
      atomicAdd(&a.ctrl->sum_aabb, (unsigned long long)na);
      atomicAdd(&a.ctrl->sum_tri, (unsigned long long)nt);
    }
  }
}

// ---------------------------------------------------------------------------
// Wide BVH (4 or 8 children per node): the reference tree with inner levels
// collapsed (each wide node holds the leaves and the inner nodes up to W
// that its binary subtree expands to, largest surface area first).  The
// leaves and every triangle test are the reference's, and a node's children
// are visited nearest first, so the nearest hit is the reference's except
// where two triangles hit at the same t or a box test's last bit decides: a
// tolerance mode (its counters are the wide tree's).
// ---------------------------------------------------------------------------

// Child references: an inner wide node's index, or kWideLeaf | count << 26 |
// first triangle (count 1..31); kWideLeaf alone marks an empty slot.
constexpr uint32_t kWideLeaf = 0x80000000u;

// A node is W / 4 quads of children in float4s: min.x of children 4q..4q+3,
// then max.x, min.y, max.y, min.z, max.z for every quad, then the child
// references (uint4 per quad), padded to whole 64-byte halves of a 128-byte
// line: 8 float4 (128 B) for W = 4, 16 (256 B) for W = 8.
template <int W>
struct Wide {
  static constexpr int kQuads = W / 4;
  static constexpr int kFloat4 = W == 4 ? 8 : 16;
};

// The reciprocal and origin-times-reciprocal of a direction for the
// multiply-add slab test, zero components replaced by 1e-20 of their sign
// (no infinities: no 0 * inf in the test).
__device__ __forceinline__ void wide_ray(V3 d, V3 o, V3& inv, V3& oi) {
  const float e = 1.0e-20f;
  const float dx = fabsf(d.x) < e ? copysignf(e, d.x) : d.x;
  const float dy = fabsf(d.y) < e ? copysignf(e, d.y) : d.y;
  const float dz = fabsf(d.z) < e ? copysignf(e, d.z) : d.z;
  inv = mk3(rcp(dx), rcp(dy), rcp(dz));
  oi = mk3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
}

__device__ __forceinline__ void wide_cas(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
  const bool sw = kb < ka;
  const float k = sw ? kb : ka;
  const uint32_t r = sw ? rb : ra;
  kb = sw ? ka : kb;
  rb = sw ? ra : rb;
  ka = k;
  ra = r;
}

// One iteration of the wide traversal: take the next entry, cull it against
// the best hit, then test an inner node's W child boxes and keep the nearest
// hit child in registers while the others are pushed farthest first, or
// test a leaf's triangles (waiting for the wave's leaf round as trav_step).
template <bool COUNT, int S, int W>
__device__ __forceinline__ void trav_step_wide(const float4* __restrict__ wn,
                                               const DevTri* __restrict__ tris, V3 o, V3 d,
                                               V3 inv, V3 oi, Stack<S, true>& st, TravState& ts,
                                               entry_t& cur, bool& have, bool leaf_round,
                                               bool* waiting) {
  if (waiting && *waiting && !leaf_round) return;
  entry_t e = cur;
  if (!have) e = st.pop();
  have = false;
  if (waiting) *waiting = false;
  if (!(entry_tmin(e) < ts.best_t)) return;
  const uint32_t x = entry_x(e);
  if (!(x & kWideLeaf)) {
    const float4* p = wn + (size_t)x * Wide<W>::kFloat4;
    constexpr int Q = Wide<W>::kQuads;
    float key[W];
    uint32_t ref[W];
    uint32_t nhit = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const float4 lx = p[0 * Q + q], hx = p[1 * Q + q], ly = p[2 * Q + q], hy = p[3 * Q + q],
                   lz = p[4 * Q + q], hz = p[5 * Q + q];
      const uint4 cr = reinterpret_cast<const uint4*>(p)[6 * Q + q];
      const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
      const float lya[4] = {ly.x, ly.y, ly.z, ly.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w};
      const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
      const uint32_t cra[4] = {cr.x, cr.y, cr.z, cr.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const float tx1 = fmaf(lxa[j], inv.x, -oi.x), tx2 = fmaf(hxa[j], inv.x, -oi.x);
        const float ty1 = fmaf(lya[j], inv.y, -oi.y), ty2 = fmaf(hya[j], inv.y, -oi.y);
        const float tz1 = fmaf(lza[j], inv.z, -oi.z), tz2 = fmaf(hza[j], inv.z, -oi.z);
        const float tmin = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
        const float tmax = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2));
        const bool hit = tmax >= tmin && tmax >= 0.0f && tmin < ts.best_t && cra[j] != kWideLeaf;
        key[4 * q + j] = hit ? tmin : INFINITY;
        ref[4 * q + j] = cra[j];
        nhit += hit ? 1u : 0u;
        if (COUNT) ts.na += cra[j] != kWideLeaf ? 1u : 0u;
      }
    }
    // ascending by entry distance (misses last): the sorting networks of
    // 4 and 8 keys (Batcher's odd-even merge sort)
    if (W == 4) {
      wide_cas(key[0], ref[0], key[1], ref[1]);
      wide_cas(key[2], ref[2], key[3], ref[3]);
      wide_cas(key[0], ref[0], key[2], ref[2]);
      wide_cas(key[1], ref[1], key[3], ref[3]);
      wide_cas(key[1], ref[1], key[2], ref[2]);
    } else {
      constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6},
                                  {5, 7}, {1, 2}, {5, 6}, {0, 4}, {1, 5}, {2, 6}, {3, 7},
                                  {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
      for (int c = 0; c < 19; c++)
        wide_cas(key[net[c][0]], ref[net[c][0]], key[net[c][1]], ref[net[c][1]]);
    }
    have = nhit > 0;
    cur = pack_entry(ref[0], key[0]);
#pragma unroll
    for (int j = W - 1; j >= 1; j--) st.push_if((uint32_t)j < nhit, pack_entry(ref[j], key[j]));
  } else {
    if (waiting && !leaf_round) {
      cur = e;
      have = true;
      *waiting = true;
      return;
    }
    const uint32_t first = x & 0x03ffffffu, count = (x >> 26) & 31u;
    for (uint32_t k = first; k < first + count; k++) {
      test_tri(tris, k, o, d, ts);
      if (COUNT) ts.nt++;
    }
  }
}

// Traversal with lane refill: a lane whose ray is finished takes a new ray
// from the queue while the other lanes continue, so a wave stays full until
// the queue is empty.  A wave refills when at least `refill` lanes are idle
// (one atomic for all of them).  Every ray is traversed exactly as by
// traverse(), one trav_step per loop iteration.
// Workgroups of 1, 2 or 4 independent waves (the launch's choice); each
// wave owns S * 64 entries of the dynamic LDS.
template <bool COUNT, int S, bool PACKED, bool MULTI, bool QUERY = false, int W = 2>
__global__ __launch_bounds__(256) void k_extend_refill(ExtendArgs a) {
  extern __shared__ uint2 lds_waves[];
  const uint32_t lane = lane_id();
  uint2* lds = lds_waves + (threadIdx.x >> 6) * (S * 64);
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t n = a.ctrl->count[a.bounce];
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.ctrl->total_rays, (unsigned long long)n);
  const uint64_t lt = (1ull << lane) - 1ull;
  // Queue partitions (parts == kParts): the queue is cut into chunks of
  // 2^glog rays, chunk c belonging to partition c % parts.  A wave takes
  // rays from the partition of its XCD first, so that the rays in flight on
  // one XCD, which has an L2 of its own, are neighbours in the queue (an
  // image band, one job of a batch); when that partition is used up, from
  // the others in turn.  Local position l of partition x is queue position
  // ((l >> glog) * parts + x) << glog | (l & (2^glog - 1)).  parts == 1: one
  // cursor, local position = queue position.
  const uint32_t parts = a.parts == kParts ? kParts : 1u;
  uint32_t glog = 31u - (uint32_t)__builtin_clz(max(n / parts, 1u));
  glog = min(max(glog, 12u), 18u);
  const uint32_t gmask = (1u << glog) - 1u;
  auto part_size = [&](uint32_t x) -> uint32_t {
    const uint32_t full = n >> glog;
    const uint32_t c = full > x ? (full - x + parts - 1) / parts : 0u;
    return (c << glog) + ((full & (parts - 1)) == x ? (n & gmask) : 0u);
  };
  auto part_cursor = [&](uint32_t x) -> uint32_t* {
    return parts == 1 ? &a.ctrl->ext_cursor[a.bounce]
                      : &a.ctrl->part_cursor[a.bounce][x * kPartStride];
  };
  uint32_t part = 0;
  if (parts > 1)
    part = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (parts - 1);  // HW_REG_XCC_ID
  uint32_t visited = 0;
  Stack<S, PACKED> st(lds + lane, a.spill + gtid, stride, 0);
  TravState ts;
  V3 o = mk3(0, 0, 0), d = mk3(1, 0, 0), inv = mk3(1, 0, 0), oi = mk3(0, 0, 0);
  const DevNode* nodes = a.nodes;
  const DevTri* tris = a.tris;
  uint32_t ray = 0;
  entry_t cur = 0;
  bool have = false;
  bool active = false, exact = false, waiting = false;
  // The wave's reserved rays [pool, pool_end) of partition pool_part: a
  // cursor hands out chunks of max(idle lanes, 64 * batches) rays, so most
  // refills take no atomic and a low refill threshold does not contend on
  // a cursor.  `drained`: the cursors have passed the last ray of every
  // partition.  (Wave-uniform values.)
  uint32_t pool = 0, pool_end = 0, pool_part = 0;
  bool drained = false;
  const uint32_t chunk = 64u * (a.batches ? a.batches : 1u);
  uint64_t na = 0, nt = 0;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if ((pool < pool_end || !drained) && (nidle >= a.refill || nidle == 64)) {
      const uint32_t from_pool = min(nidle, pool_end - pool);
      const uint32_t base0 = pool, part0 = pool_part;
      pool += from_pool;
      const uint32_t need = nidle - from_pool;
      uint32_t base1 = 0, end1 = 0;
      if (need && !drained) {
        const uint32_t sz = max(need, chunk);
        // at most one fetch per partition finds it used up
        for (;;) {
          const uint32_t size = part_size(part);
          base1 = size ? fetch(part_cursor(part), sz) : 0u;
          const bool got = base1 < size;
          if (got) {
            end1 = min(base1 + sz, size);
            pool_part = part;
          }
          if (!got || base1 + sz >= size) {
            // this partition is used up: the next fetch goes to the next one
            if (++visited == parts) drained = true;
            part = (part + 1) & (parts - 1);
          }
          if (got || drained) break;
        }
        if (end1 == 0) base1 = 0;
        pool = min(base1 + need, end1);
        pool_end = end1;
      }
      if (!active) {
        const uint32_t rank = __popcll(idle & lt);
        const bool old = rank < from_pool;
        const uint32_t l = old ? base0 + rank : base1 + (rank - from_pool);
        if (old || l < end1) {
          const uint32_t x = old ? part0 : pool_part;
          const uint32_t k = parts == 1 ? l : (((l >> glog) * parts + x) << glog) | (l & gmask);
          // hits-only queries are never sorted
          const uint32_t i = !QUERY && a.order ? a.order[k] : k;
          if (MULTI) {
            const float4 r0 = a.r0[i], r1 = a.r1[i];
            o = mk3(r0.x, r0.y, r0.z);
            d = mk3(r1.x, r1.y, r1.z);
            const JobDesc& J = a.jobs[f2u(r1.w)];
            nodes = J.nodes;
            tris = J.tris;
          } else if (QUERY) {
            load_ray<true>(a, i, o, d);
          } else {
            const float4 r0 = a.r0[i], r1 = a.r1[i];
            o = mk3(r0.x, r0.y, r0.z);
            d = mk3(r1.x, r1.y, r1.z);
          }
          if (W == 2) {
            inv = mk3(rcp(d.x), rcp(d.y), rcp(d.z));
            exact = needs_exact_nan(o, inv);
          } else {
            wide_ray(d, o, inv, oi);
          }
          ts.best_t = 1.0e5f;
          ts.best = kNoHit;
          ts.bu = ts.bv = 0.0f;
          ts.na = 0;
          ts.nt = 0;
          st.sp = 0;
          cur = W == 2 ? root_entry<PACKED>() : pack_entry(0u, -INFINITY);
          have = true;
          waiting = false;
          ray = i;
          active = true;
        }
      }
    }
    if (!__any(active)) {
      if (drained && pool >= pool_end) break;
      continue;
    }
    const bool wave_exact = __any(active && exact);
    // several traversal steps per refill check: the check's votes cost
    // about as much as a step's mask logic, and a lane that finishes early
    // idles for at most interval - 1 steps
#pragma unroll 1
    for (uint32_t rep = 0; rep < a.interval; rep++) {
      // a leaf round when leaf_batch / 64 of the active lanes wait at
      // leaves (a share, so that a wave with few rays left does not hold
      // its leaves back), or every active lane does (so that the wave
      // always progresses)
      bool leaf_round = true;
      if (a.leaf_batch) {
        const uint64_t wm = __ballot(active && waiting), am = __ballot(active);
        leaf_round = (uint32_t)__popcll(wm) * 64u >= a.leaf_batch * (uint32_t)__popcll(am) ||
                     wm == am;
      }
      if (active) {
        if constexpr (W == 2)
          trav_step<COUNT>(nodes, tris, o, d, inv, wave_exact, st, ts, cur, have, leaf_round,
                           &waiting);
        else
          trav_step_wide<COUNT, S, W>(reinterpret_cast<const float4*>(nodes), tris, o, d, inv, oi,
                                      st, ts, cur, have, leaf_round, &waiting);
        if (!have && st.sp == 0) {
          a.hits[ray] = make_uint4(QUERY ? hit_id<true>(a, ts.best) : ts.best, f2u(ts.best_t),
                                   f2u(ts.bu), f2u(ts.bv));
          if (COUNT) {
            na += ts.na;
            nt += ts.nt;
          }
          active = false;
        }
      }
    }
  }
  if (COUNT) {
    for (int off = 32; off > 0; off >>= 1) {
      na += __shfl_down(na, off);
      nt += __shfl_down(nt, off);
    }
    if (lane == 0) {
      atomicAdd(&a.ctrl->sum_aabb, (unsigned long long)na);
      atomicAdd(&a.ctrl->sum_tri, (unsigned long long)nt);
    }
  }
}

// ---------------------------------------------------------------------------
// shade: continue_path (material.rs:327-411) and the loop of render_pixels
// (renderer.rs:413-468), one pixel per lane.
// ---------------------------------------------------------------------------

struct ShadeOut {
  V3 o, d, cm;
  float fresnel;
};

// The six draws of one continue_path call in the reference's order
// (material.rs:342-354): hemisphere angle and unit, light index bits and
// the two unit coordinates of the light sample, the biunit of the choice.
struct PathDraws {
  uint32_t angle, unit, bits, u, v, rr;
};

__device__ __forceinline__ ShadeOut continue_path(const JobDesc& J, V3 ray_d, V3 pos, V3 normal,
                                                  uint32_t mat, const PathDraws& dr,
                                                  bool ignore_fresnel) {
  const uint32_t d_angle = dr.angle, d_unit = dr.unit, d_bits = dr.bits, d_u = dr.u, d_v = dr.v;
  const uint32_t d_rr = dr.rr;
  // continue_path_brdf (material.rs:228-242), sample_hemisphere_vector (random.rs:140-159)
  float phi = u32_angle(d_angle), r_sqr = u32_unit(d_unit);
  float rr_ = sqrtf(r_sqr);
  V3 dir_z = mk3(sin_poly(phi) * rr_, cos_poly(phi) * rr_, sqrtf(1.0f - r_sqr));
  V3 dir_b = rotate_hemisphere(dir_z, normal);
  V3 org_b = muladd3(dir_b, 1.0e-4f, pos);
  // continue_path_direct_sample + Scene::get_direct_sample (scene.rs:143-202)
  const DevLight& L = J.lights[d_bits >> 29];
  float su = u32_unit(d_u), sv = u32_unit(d_v);
  uint32_t sm = f2u((1.0f - su) - sv);
  float uu = pickf(su, 1.0f - su, sm);
  float vv = pickf(sv, 1.0f - sv, sm);
  V3 e1 = mk3(L.e1[0], L.e1[1], L.e1[2]);
  V3 e2 = mk3(L.e2[0], L.e2[1], L.e2[2]);
  V3 lv0 = mk3(L.v0[0], L.v0[1], L.v0[2]);
  V3 p = muladd3(e2, vv, negmuladd3(e1, uu, lv0));
  V3 dir_d = normalized3(sub3(p, pos));
  V3 org_d = muladd3(dir_d, 1.0e-4f, pos);
  float rr = fmaf(normal.z, 0.8f, u32_biunit(d_rr));
  uint32_t rrm = f2u(rr);
  V3 no = pick3(org_b, org_d, rrm);
  V3 nd = pick3(dir_b, dir_d, rrm);
  // pd_brdf (material.rs:245-250)
  float pd_b = xmax(dot3(normal, nd), 0.0f) * (1.0f / 3.14159274f);
  // pd_direct_sample (material.rs:281-320) over every light in BVH order
  float total = 0.0f;
  const uint32_t nlights = J.nlights;
  for (uint32_t k = 0; k < nlights; k++) {
    const DevLight& M = J.lights[k];
    V3 m0 = mk3(M.v0[0], M.v0[1], M.v0[2]);
    V3 mn = mk3(M.n[0], M.n[1], M.n[2]);
    V3 me1 = mk3(M.e1[0], M.e1[1], M.e1[2]);
    V3 me2 = mk3(M.e2[0], M.e2[1], M.e2[2]);
    V3 fr = sub3(m0, no);
    float denom = rcp(dot3(nd, mn));
    float t = x86nan(dot3(fr, mn) * denom);
    V3 c = cross3(nd, fr);
    float u = dot3(c, me2) * denom;
    float v = dot3(c, me1) * denom;
    uint32_t mask_uv = geq(u + v, 1.0f);
    uint32_t mask = (f2u(u) | f2u(v)) | (f2u(t) | mask_uv);
    float dist_sqr = t * t;
    float dot_em = absb(dot3(mk3(M.normal[0], M.normal[1], M.normal[2]), nd)) + 0.0001f;
    float pd = dist_sqr * rcp(M.area * dot_em);
    total = pickf(total + pd, total, mask);
  }
  float pd_d = total * (1.0f / (float)nlights);
  float weight_denom = (pd_b + pd_d) + 0.01f;
  float p_brdf = fmaf(normal.z, 0.4f, 0.5f);
  float p_direct = fmaf(-normal.z, 0.4f, 0.5f);
  float pp = pickf(p_brdf, p_direct, rrm);
  float modulation = rcp(weight_denom * pp);
  float cos_theta = xmax(dot3(normal, nd), 0.0f);
  // microfacet_brdf (material.rs:415-450): ray_in = the new ray, ray_out = the old
  int32_t mi = (int32_t)mat;
  const float k255 = 1.0f / 255.0f;
  V3 color = mk3((float)(mi & 0xff) * k255, (float)((mi >> 8) & 0xff) * k255,
                 (float)((mi >> 16) & 0xff) * k255);
  int32_t gloss = (mi >> 26) & 7;
  V3 h = normalized3(sub3(nd, ray_d));
  // microfacet_normal_dist (material.rs:475-513)
  const float pi = 3.14159274f;
  float c1 = dot3(h, normal);
  float c2 = c1 * c1;
  float c4 = c2 * c2;
  float c8 = c4 * c4;
  float c16 = c8 * c8;
  float dterm;
  switch (gloss) {
    case 0: dterm = 0.5f / pi; break;
    case 1: dterm = xmax(c1, 0.0f) * (1.0f / pi); break;
    case 2: dterm = c2 * (1.5f / pi); break;
    case 3: dterm = c4 * (2.5f / pi); break;
    case 4: dterm = c8 * (4.5f / pi); break;
    default: dterm = c16 * (8.5f / pi); break;
  }
  // microfacet_fresnel (material.rs:453-469)
  V3 r1 = sub3(mk3(1.0f, 1.0f, 1.0f), color);
  float ct = 1.0f - absb(dot3(h, nd));
  float ct2 = ct * ct;
  float ct3 = ct2 * ct;
  float ct5 = absb(ct2 * ct3);
  V3 fc = muladd3(r1, ct5, color);
  if (ignore_fresnel && (((uint32_t)mi << 6) | ((uint32_t)mi << 7)) >> 31) fc = mk3(1.0f, 1.0f, 1.0f);
  V3 brdf = scale3(fc, dterm);
  V3 cm = scale3(brdf, modulation * cos_theta);
  ShadeOut r;
  r.cm = mk3(xmin(cm.x, 2.0f), xmin(cm.y, 2.0f), xmin(cm.z, 2.0f));
  r.o = no;
  r.d = nd;
  r.fresnel = ct5;
  return r;
}

// continue_path with the draws taken from a PIXEL stream.
__device__ __forceinline__ ShadeOut continue_path(const JobDesc& J, V3 ray_d, V3 pos, V3 normal,
                                                  uint32_t mat, uint64_t& st,
                                                  bool ignore_fresnel) {
  PathDraws dr;
  dr.angle = rng_next(st);
  dr.unit = rng_next(st);
  dr.bits = rng_next(st);
  dr.u = rng_next(st);
  dr.v = rng_next(st);
  dr.rr = rng_next(st);
  return continue_path(J, ray_d, pos, normal, mat, dr, ignore_fresnel);
}

// One path per thread; blocks beyond the queue exit at once.  Survivors are
// appended with one atomic per block: ballots per wave, a prefix over the
// block's waves in LDS.  MULTI == false: every path belongs to job 0.
template <bool MULTI, bool SHARED>
__global__ __launch_bounds__(kShadeBlock) void k_shade(ShadeArgs a) {
  __shared__ uint32_t wcount[kShadeBlock / 64];
  __shared__ uint32_t block_base;
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t n = a.ctrl->count[a.bounce];
  if (blockIdx.x * kShadeBlock >= n) return;
  const bool last = a.bounce + 1 >= a.max_bounces;
  const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
  const bool valid = i < n;
  bool keep = false;
  float4 r0 = make_float4(0, 0, 0, 0), r1 = make_float4(0, 0, 0, 0);
  uint64_t st = 0;
  ShadeOut so;
  if (valid) {
    r0 = a.in.r0[i];
    r1 = a.in.r1[i];
    float4 col = a.in.col[i];
    st = a.in.rng[i];
    uint4 hit = a.hits[i];
    // J: the path's job (image, outputs); S: its scene, read from the first
    // job when every job of the step shares one (wave-uniform scalar loads)
    const JobDesc& J = a.jobs[MULTI ? f2u(r1.w) : 0u];
    const JobDesc& S = SHARED ? a.jobs[0] : J;
    V3 o = mk3(r0.x, r0.y, r0.z), d = mk3(r1.x, r1.y, r1.z);
    V3 color = mk3(col.x, col.y, col.z);
    uint32_t pixel = f2u(r0.w);
    const uint32_t width = J.width;
    uint32_t px = pixel % width, py = pixel / width;
    float t = u2f(hit.y);
    V3 pos, normal;
    uint32_t mat;
    float tu = 0.0f, tv = 0.0f;
    uint32_t* gbuffer = a.bounce == 0 ? J.gbuffer : nullptr;
    if (hit.x == kNoHit) {
      pos = muladd3(d, 1.0e5f, o);
      normal = d;
      mat = kMatSky;
    } else {
      const float4* tp = reinterpret_cast<const float4*>(S.tris + hit.x);
      V3 nn = mk3(tp[0].w, tp[1].w, tp[2].w);
      pos = muladd3(d, t, o);
      normal = normalized3(nn);
      mat = S.tri_mat[hit.x];
      if (gbuffer) {
        const float* uv = S.tri_uv + 6 * (size_t)hit.x;
        float bu = u2f(hit.z), bv = u2f(hit.w);
        float w = (1.0f - bu) - bv;
        tu = fmaf(uv[0], w, fmaf(uv[2], bv, uv[4] * bu));
        tv = fmaf(uv[1], w, fmaf(uv[3], bv, uv[5] * bu));
      }
    }
    if (a.bounce == 0 && J.features) {
      // first-hit features of the denoiser: albedo, normal, distance, count
      float* f = J.features + (size_t)pixel * 8;
      const float k255 = 1.0f / 255.0f;
      V3 alb = (mat >> 31) ? mk3(1.0f, 1.0f, 1.0f)
                           : mk3((float)(mat & 0xffu) * k255, (float)((mat >> 8) & 0xffu) * k255,
                                 (float)((mat >> 16) & 0xffu) * k255);
      f[0] += alb.x;
      f[1] += alb.y;
      f[2] += alb.z;
      f[3] += normal.x;
      f[4] += normal.y;
      f[5] += normal.z;
      f[6] += hit.x == kNoHit ? 1.0e5f : t;
      f[7] += 1.0f;
    }
    if (mat >> 31) {
      // The path found an emitter: colour x emission (renderer.rs:452-460).
      V3 em = (S.has_override && hit.x != kNoHit)
                  ? mk3(S.emission[0], S.emission[1], S.emission[2])
                  : sky_intensity(d);
      V3 c = mulc3(color, em);
      if (J.hdr) {
        float* h = J.hdr + hdr_offset(px, py, width);
        h[0] += c.x;
        h[8] += c.y;
        h[16] += c.z;
      }
      if (J.bitmap) J.bitmap[pixel] = pack_color(c);
    } else if (!last || gbuffer) {
      so = continue_path(S, d, pos, normal, mat, st, a.bounce == 0);
      if (gbuffer) {
        uint32_t tex_index = ((int32_t)mat >> 24) & 3;
        gbuffer[pixel] = pack_gbuffer(tex_index, tu, tv, so.fresnel);
      }
      if (!last) {
        keep = true;
        so.cm = mulc3(color, so.cm);
      }
    }
  }
  const uint64_t ballot = __ballot(keep);
  if (lane == 0) wcount[wave] = __popcll(ballot);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < kShadeBlock / 64; w++) tot += wcount[w];
    block_base = tot ? atomicAdd(&a.ctrl->count[a.bounce + 1], tot) : 0u;
  }
  __syncthreads();
  if (keep) {
    uint32_t j = block_base;
    for (uint32_t w = 0; w < wave; w++) j += wcount[w];
    j += __popcll(ballot & ((1ull << lane) - 1ull));
    a.out.r0[j] = make_float4(so.o.x, so.o.y, so.o.z, r0.w);
    a.out.r1[j] = make_float4(so.d.x, so.d.y, so.d.z, r1.w);
    a.out.col[j] = make_float4(so.cm.x, so.cm.y, so.cm.z, 0.0f);
    a.out.rng[j] = st;
  }
}

// buffer_f32_into_render_buffer (renderer.rs:378-410)
__global__ __launch_bounds__(256) void k_hdr_to_bitmap(const float* hdr, uint32_t* bitmap,
                                                       uint32_t width, uint32_t height,
                                                       uint32_t num_samples) {
  const float factor = 1.0f / (float)num_samples;
  uint32_t npix = width * height;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    uint32_t px = p % width, py = p / width;
    const float* h = hdr + hdr_offset(px, py, width);
    bitmap[p] = pack_color(mk3(h[0] * factor, h[8] * factor, h[16] * factor));
  }
}

// Per-triangle device geometry (DevTri) and light data (DevLight) with the
// reference's operations (triangle.rs:65-80,152-165).
__global__ __launch_bounds__(256) void k_prepare_tris(const float* v9, const uint32_t* order,
                                                      uint32_t ntris, DevTri* out) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ntris; k += gridDim.x * blockDim.x) {
    const float* t = v9 + 9 * (size_t)order[k];
    V3 v0 = mk3(t[0], t[1], t[2]), v1 = mk3(t[3], t[4], t[5]), v2 = mk3(t[6], t[7], t[8]);
    V3 e1 = sub3(v0, v2);
    V3 e2 = sub3(v1, v0);
    V3 n = cross3(e1, e2);
    out[k].p = make_float4(v0.x, v0.y, v0.z, n.x);
    out[k].q = make_float4(e1.x, e1.y, e1.z, n.y);
    out[k].r = make_float4(e2.x, e2.y, e2.z, n.z);
  }
}

__global__ void k_prepare_lights(const DevTri* tris, const uint32_t* direct, uint32_t n,
                                 DevLight* out) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const DevTri& T = tris[direct[k]];
  V3 v0 = mk3(T.p.x, T.p.y, T.p.z);
  V3 e1 = mk3(T.q.x, T.q.y, T.q.z);
  V3 e2 = mk3(T.r.x, T.r.y, T.r.z);
  V3 nn = mk3(T.p.w, T.q.w, T.r.w);
  float rnorm = rsqrt(dot3(nn, nn));
  float area = 0.5f * rcp(rnorm);
  V3 nrm = scale3(nn, rnorm);
  DevLight L;
  L.v0[0] = v0.x; L.v0[1] = v0.y; L.v0[2] = v0.z;
  L.e1[0] = e1.x; L.e1[1] = e1.y; L.e1[2] = e1.z;
  L.e2[0] = e2.x; L.e2[1] = e2.y; L.e2[2] = e2.z;
  L.n[0] = nn.x; L.n[1] = nn.y; L.n[2] = nn.z;
  L.normal[0] = nrm.x; L.normal[1] = nrm.y; L.normal[2] = nrm.z;
  L.area = area;
  out[k] = L;
}

// Frames rendered in the same launches accumulate into separate buffers;
// this adds buffers 1..count to the HDR image in frame order and clears
// them, so that every pixel receives the additions of a sequential
// accumulation, in the same order.  One thread per pixel of the rectangle.
__global__ void k_fold_hdr(float* __restrict__ hdr, float* __restrict__ slots, size_t slot_stride,
                           uint32_t count, uint32_t width, uint32_t x0, uint32_t y0, uint32_t w,
                           uint32_t npix) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    const uint32_t px = x0 + p % w, py = y0 + p / w;
    const size_t off = hdr_offset(px, py, width);
    float r = hdr[off], g = hdr[off + 8], b = hdr[off + 16];
    // 8 buffers at a time: their loads first (issued together rather than
    // each behind the previous buffer's clearing stores), then the sums in
    // frame order, then the clearing
    for (uint32_t k0 = 0; k0 < count; k0 += 8) {
      float v[8][3];
#pragma unroll
      for (uint32_t j = 0; j < 8; j++) {
        if (k0 + j >= count) break;
        const float* s = slots + (k0 + j) * slot_stride + off;
        v[j][0] = s[0];
        v[j][1] = s[8];
        v[j][2] = s[16];
      }
#pragma unroll
      for (uint32_t j = 0; j < 8; j++) {
        if (k0 + j >= count) break;
        r = r + v[j][0];
        g = g + v[j][1];
        b = b + v[j][2];
      }
#pragma unroll
      for (uint32_t j = 0; j < 8; j++) {
        if (k0 + j >= count) break;
        float* s = slots + (k0 + j) * slot_stride + off;
        s[0] = 0.0f;
        s[8] = 0.0f;
        s[16] = 0.0f;
      }
    }
    hdr[off] = r;
    hdr[off + 8] = g;
    hdr[off + 16] = b;
  }
}

// The same for the denoiser's first-hit features (8 floats per pixel,
// row-major over n pixels): features += slots[0] + ... in frame order, the
// slots cleared.  One thread per float.
__global__ void k_fold_features(float* __restrict__ feat, float* __restrict__ slots,
                                size_t slot_stride, uint32_t count, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    float v = feat[i];
    // 8 buffers at a time, loads first (as in k_fold_hdr)
    for (uint32_t k0 = 0; k0 < count; k0 += 8) {
      float t[8];
#pragma unroll
      for (uint32_t j = 0; j < 8; j++)
        if (k0 + j < count) t[j] = slots[(k0 + j) * slot_stride + i];
#pragma unroll
      for (uint32_t j = 0; j < 8; j++)
        if (k0 + j < count) v = v + t[j];
#pragma unroll
      for (uint32_t j = 0; j < 8; j++)
        if (k0 + j < count) slots[(k0 + j) * slot_stride + i] = 0.0f;
    }
    feat[i] = v;
  }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Ray sort: a counting sort of a bounce's paths by key into a traversal
// order.  The key of a path is its job index (multi-job steps), the octant
// of its direction and the Morton code of its origin quantised to
// 2^axis_bits cells per axis of the job's scene bounds (outside: clamped).
// Paths of one key keep no particular order (atomics), which changes only
// which lane traverses which ray.
// ---------------------------------------------------------------------------

// The low 8 bits of v moved to every third bit (bit k to bit 3k).
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0xffu;
  v = (v | (v << 8)) & 0x0000f00fu;
  v = (v | (v << 4)) & 0x000c30c3u;
  v = (v | (v << 2)) & 0x00249249u;
  return v;
}

__device__ __forceinline__ uint32_t ray_key(const SortArgs& a, float4 r0, float4 r1) {
  const uint32_t job = a.multi_job ? f2u(r1.w) : 0u;
  const JobDesc& J = a.jobs[job];
  const float cells = (float)(1u << a.axis_bits);
  const float p[3] = {r0.x, r0.y, r0.z};
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    // NaN and points below the bounds go to cell 0 (fmaxf keeps the number)
    const float u = fminf(fmaxf((p[k] - J.lo[k]) * J.inv_ext[k] * cells, 0.0f), cells - 1.0f);
    m |= spread3((uint32_t)u) << k;
  }
  const uint32_t oct =
      (r1.x < 0.0f ? 1u : 0u) | (r1.y < 0.0f ? 2u : 0u) | (r1.z < 0.0f ? 4u : 0u);
  const uint32_t mb = 3 * a.axis_bits;
  return (job << (mb + 3)) | (oct << mb) | m;
}

__global__ __launch_bounds__(256) void k_sort_clear(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0u;
}

__global__ __launch_bounds__(256) void k_sort_keys(SortArgs a) {
  const uint32_t n = a.ctrl->count[a.bounce];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t key = ray_key(a, a.r0[i], a.r1[i]);
    a.keys[i] = key;
    atomicAdd(&a.buckets[key], 1u);
  }
}

// In-place exclusive prefix sums of b[0, nb) within blocks of
// kSortScanBlock words (1024 threads, 4 words each); each block's total
// goes to sums[blockIdx.x].
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t* b, uint32_t nb, uint32_t* sums) {
  __shared__ uint32_t wsum[16];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t base = blockIdx.x * kSortScanBlock + threadIdx.x * 4;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = base + k < nb ? b[base + k] : 0u;
  const uint32_t t = v[0] + v[1] + v[2] + v[3];
  uint32_t inc = t;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < 16; w++) {
      const uint32_t x = wsum[w];
      wsum[w] = run;
      run += x;
    }
    sums[blockIdx.x] = run;
  }
  __syncthreads();
  uint32_t run = wsum[wave] + inc - t;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < nb) b[base + k] = run;
    run += v[k];
  }
}

// order[position] = path: the position is the key's start (its block's
// prefix plus the prefix within the block) plus the path's rank among the
// key's paths.
__global__ __launch_bounds__(256) void k_sort_scatter(SortArgs a) {
  const uint32_t n = a.ctrl->count[a.bounce];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t key = a.keys[i];
    const uint32_t pos = a.block_sums[key / kSortScanBlock] + atomicAdd(&a.buckets[key], 1u);
    a.order[pos] = i;
  }
}

}  // namespace cvk

// ---------------------------------------------------------------------------
// Host launch wrappers
// ---------------------------------------------------------------------------

namespace cvk {

// ---------------------------------------------------------------------------
// REF_PATCH: the reference's own random streams (random.rs:33-136,
// renderer.rs:292-468), one 8-lane group at a time
// ---------------------------------------------------------------------------

// Rng::with_seed (random.rs:43-70) and Rng::next with the Mu64 -> 8 x u32
// view (random.rs:73-101): every lane of a group holds the whole 4 x u64
// state and advances it identically; lane l reads the low (l even) or high
// (l odd) half of u64 l / 2.
struct RefRng {
  uint64_t s0, s1, s2, s3;
};

__device__ __forceinline__ RefRng ref_rng_seed(uint32_t x, uint32_t y, uint32_t i) {
  uint64_t seed = (uint64_t)x * 12276630456901467871ull + (uint64_t)y * 7661526868048087387ull +
                  (uint64_t)i * 2268244495640532043ull;
  seed = seed + seed % 9358246936573323101ull;
  RefRng r;
  r.s0 = seed * 14491630826648200009ull;
  r.s1 = seed * 13149596372461506851ull;
  r.s2 = seed * 6119410235796056053ull;
  r.s3 = seed * 14990141545859273719ull;
  return r;
}

__device__ __forceinline__ uint32_t ref_rng_next(RefRng& r, uint32_t l) {
  const uint32_t k = l >> 1;
  const uint64_t v = k == 0 ? r.s0 : k == 1 ? r.s1 : k == 2 ? r.s2 : r.s3;
  r.s0 *= 3ull * 1073243692214514217ull;
  r.s1 *= 5ull * 3335100457702756523ull;
  r.s2 *= 7ull * 8789056573444181ull;
  r.s3 *= 11ull * 781436371140792079ull;
  return (l & 1u) ? (uint32_t)(v >> 32) : (uint32_t)v;
}

// One 8-lane group per patch: lanes 8k .. 8k+7 of a wave render patch
// blockIdx.x * 8 + k (and every 8 * gridDim.x-th after it), frames in
// order, blocks of 16 x 4 pixels column by column (renderer.rs:306-362),
// the block's 16 coordinate draws, then its 8 groups one after the other,
// each drawing t and six draws per bounce from the patch's stream until all
// 8 of its lanes have hit emitters (renderer.rs:413-468).  A lane's own
// traversal is the single-ray traversal of the other modes.  The 8 lanes
// of a group take the same branches at every step (their loop counts are
// the group's), so the all-emissive vote sees all 8.
template <int S, bool PACKED>
__global__ __launch_bounds__(64) void k_ref_patch(RefPatchArgs a) {
  __shared__ uint2 lds[S * 64];
  const uint32_t lane = lane_id(), sub = lane >> 3, l = lane & 7u;
  const uint32_t gtid = blockIdx.x * 64 + lane;
  const uint32_t stride = gridDim.x * 64;
  const JobDesc& J = a.job;
  const float scale = 2.0f / (float)J.width;
  const V3 cam_pos = mk3(J.cam.pos[0], J.cam.pos[1], J.cam.pos[2]);
  const V3 cam_dpos = mk3(J.cam.dpos[0], J.cam.dpos[1], J.cam.dpos[2]);
  for (uint32_t pidx = blockIdx.x * 8 + sub; pidx < a.npatches; pidx += gridDim.x * 8) {
    const uint4 P = a.patches[pidx];
    const uint32_t x = P.x, y = P.y, pw = P.z;
    for (uint32_t f = 0; f < a.nframes; f++) {
      RefRng rng = ref_rng_seed(x, y, a.first_frame + f);
      for (uint32_t bi = 0; bi < pw / 16; bi++)
        for (uint32_t bj = 0; bj < pw / 4; bj++) {
          const uint32_t xb = x + 16 * bi, yb = y + 4 * bj;
          const float cx = (float)xb - (float)J.width * 0.5f;
          const float cy = (float)yb - (float)J.height * 0.5f;
          const float bx = scale * ((float)(l & 3u) + cx);
          const float by = scale * ((float)(l >> 2) + cy);
          float xs[8], ys[8];
#pragma unroll
          for (int g = 0; g < 8; g++) {
            const int gx = g >> 1;
            xs[g] = gx == 0 ? bx : gx == 1 ? bx + 4.0f * scale : gx == 2 ? bx + 8.0f * scale
                                                                          : bx + 12.0f * scale;
            ys[g] = (g & 1) ? by + 2.0f * scale : by;
          }
#pragma unroll
          for (int g = 0; g < 8; g++) xs[g] = fmaf(u32_unit(ref_rng_next(rng, l)), scale, xs[g]);
#pragma unroll
          for (int g = 0; g < 8; g++) ys[g] = fmaf(u32_unit(ref_rng_next(rng, l)), scale, ys[g]);
          for (int g = 0; g < 8; g++) {
            // render_pixels for group g (renderer.rs:413-468)
            const float t = u32_unit(ref_rng_next(rng, l));
            V3 o = muladd3(cam_dpos, t, cam_pos);
            V3 d = camera_dir(J.cam, xs[g], ys[g], t);
            uint32_t active = 0;  // sign bit 1: inactive (ray.rs:21-33)
            V3 color = mk3(1.0f, 1.0f, 1.0f);
            uint32_t hit_em = 0;
            bool term_tri = false;
            uint32_t tex = 0;
            float gtu = 0.0f, gtv = 0.0f, gfr = 0.0f;
            for (uint32_t b = 0; b < a.max_bounces; b++) {
              const bool act = (active >> 31) == 0;
              TravState ts = traverse<false, S, PACKED>(J.nodes, J.tris, o, d, lds + lane,
                                                        a.spill + gtid, stride, act);
              V3 pos, normal;
              uint32_t mat;
              float tu = 0.0f, tv = 0.0f;
              if (!act || ts.best == kNoHit) {
                // Scene::intersect_nearest's far-away sky hit (scene.rs:220-230)
                pos = muladd3(d, 1.0e5f, o);
                normal = d;
                mat = kMatSky;
              } else {
                const float4* tp = reinterpret_cast<const float4*>(J.tris + ts.best);
                pos = muladd3(d, ts.best_t, o);
                normal = normalized3(mk3(tp[0].w, tp[1].w, tp[2].w));
                mat = J.tri_mat[ts.best];
                const float* uv = J.tri_uv + 6 * (size_t)ts.best;
                const float w = (1.0f - ts.bu) - ts.bv;
                tu = fmaf(uv[0], w, fmaf(uv[2], ts.bv, uv[4] * ts.bu));
                tv = fmaf(uv[1], w, fmaf(uv[3], ts.bv, uv[5] * ts.bu));
              }
              if (act && (mat >> 31)) term_tri = ts.best != kNoHit;
              hit_em = mat;
              // break when all 8 lanes of the group hit emitters (renderer.rs:433)
              const uint64_t em = __ballot((mat >> 31) != 0);
              if (((em >> (sub * 8)) & 0xffull) == 0xffull) break;
              PathDraws dr;
              dr.angle = ref_rng_next(rng, l);
              dr.unit = ref_rng_next(rng, l);
              dr.bits = ref_rng_next(rng, l);
              dr.u = ref_rng_next(rng, l);
              dr.v = ref_rng_next(rng, l);
              dr.rr = ref_rng_next(rng, l);
              const ShadeOut so = continue_path(J, d, pos, normal, mat, dr, b == 0);
              // an emissive hit (or an inactive ray) keeps its ray and gets
              // a white modulation (material.rs:338,401-408)
              active |= mat;
              if ((active >> 31) == 0) {
                o = so.o;
                d = so.d;
                color = mulc3(color, so.cm);
              }
              if (b == 0) {
                tex = ((int32_t)mat >> 24) & 3;
                gtu = tu;
                gtv = tv;
                gfr = so.fresnel;
              }
            }
            // renderer.rs:452-460: colour x emission of the last ray, zero
            // unless the last intersection was emissive
            V3 emis = (J.has_override && term_tri)
                          ? mk3(J.emission[0], J.emission[1], J.emission[2])
                          : sky_intensity(d);
            V3 c = mulc3(color, emis);
            if (!(hit_em >> 31)) c = mk3(0.0f, 0.0f, 0.0f);
            const uint32_t px = xb + 4 * ((uint32_t)g >> 1) + (l & 3u);
            const uint32_t py = yb + 2 * ((uint32_t)g & 1u) + (l >> 2);
            const size_t pixel = (size_t)py * J.width + px;
            if (J.hdr) {
              float* h = J.hdr + hdr_offset(px, py, J.width);
              h[0] += c.x;
              h[8] += c.y;
              h[16] += c.z;
            }
            if (J.bitmap) J.bitmap[pixel] = pack_color(c);
            if (J.gbuffer) J.gbuffer[pixel] = pack_gbuffer(tex, gtu, gtv, gfr);
          }
        }
    }
  }
}

#define CVK_STACK_VARIANTS(X) X(8) X(16) X(24) X(32)

template <bool COUNT, int S>
static hipError_t launch_intersect_s(const IntersectArgs& a, bool packed, uint32_t blocks,
                                     hipStream_t s) {
  if (packed)
    hipLaunchKernelGGL((k_intersect<COUNT, S, true>), dim3(blocks), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((k_intersect<COUNT, S, false>), dim3(blocks), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_intersect(const IntersectArgs& a, bool count, bool packed, int stack,
                            uint32_t blocks, hipStream_t s) {
#define CASE(S)                                                                \
  if (stack == S)                                                              \
    return count ? launch_intersect_s<true, S>(a, packed, blocks, s)           \
                 : launch_intersect_s<false, S>(a, packed, blocks, s);
  CVK_STACK_VARIANTS(CASE)
#undef CASE
  return hipErrorInvalidValue;
}

// Wide trees (4 or 8 children): lane refill only (whole-batch requests run
// with refill at 64 lanes), stack entries are child references.
template <bool COUNT, int S, int W>
static hipError_t launch_extend_wide(const ExtendArgs& a, bool multi, uint32_t blocks,
                                     uint32_t block_waves, hipStream_t s) {
  const uint32_t bw = block_waves == 2 || block_waves == 4 ? block_waves : 1u;
  const dim3 grid((blocks + bw - 1) / bw), block(64 * bw);
  const size_t shmem = (size_t)S * 64 * sizeof(uint2) * bw;
  ExtendArgs b = a;
  if (!b.refill) b.refill = 64;
  if (multi)
    hipLaunchKernelGGL((k_extend_refill<COUNT, S, true, true, false, W>), grid, block, shmem, s, b);
  else if (b.rays6 && !COUNT)
    hipLaunchKernelGGL((k_extend_refill<false, S, true, false, true, W>), grid, block, shmem, s,
                       b);
  else
    hipLaunchKernelGGL((k_extend_refill<COUNT, S, true, false, false, W>), grid, block, shmem, s,
                       b);
  return hipGetLastError();
}

template <bool COUNT, int S>
static hipError_t launch_extend_s(const ExtendArgs& a, bool packed, bool multi, uint32_t blocks,
                                  uint32_t block_waves, uint32_t width, hipStream_t s) {
  if (width == 4) return launch_extend_wide<COUNT, S, 4>(a, multi, blocks, block_waves, s);
  if (width == 8) return launch_extend_wide<COUNT, S, 8>(a, multi, blocks, block_waves, s);
  const uint32_t bw = block_waves == 2 || block_waves == 4 ? block_waves : 1u;
  const dim3 grid((blocks + bw - 1) / bw), block(64 * bw);
  const size_t shmem = (size_t)S * 64 * sizeof(uint2) * bw;
  if (multi) {
    if (packed)
      hipLaunchKernelGGL((k_extend_refill<COUNT, S, true, true>), grid, block, shmem, s, a);
    else
      hipLaunchKernelGGL((k_extend_refill<COUNT, S, false, true>), grid, block, shmem, s, a);
  } else if (a.refill && a.rays6 && !COUNT) {
    // batched queries: the caller's rays and source ids, compiled apart so
    // that the render instantiations carry neither
    if (packed)
      hipLaunchKernelGGL((k_extend_refill<false, S, true, false, true>), grid, block, shmem, s, a);
    else
      hipLaunchKernelGGL((k_extend_refill<false, S, false, false, true>), grid, block, shmem, s,
                         a);
  } else if (a.refill) {
    if (packed)
      hipLaunchKernelGGL((k_extend_refill<COUNT, S, true, false>), grid, block, shmem, s, a);
    else
      hipLaunchKernelGGL((k_extend_refill<COUNT, S, false, false>), grid, block, shmem, s, a);
  } else if (packed) {
    hipLaunchKernelGGL((k_extend<COUNT, S, true>), dim3(blocks), dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_extend<COUNT, S, false>), dim3(blocks), dim3(64), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_extend(const ExtendArgs& a, bool count, bool packed, bool multi, int stack,
                         uint32_t blocks, uint32_t block_waves, hipStream_t s, uint32_t width) {
  if (multi && !a.refill) return hipErrorInvalidValue;
  if (width != 2 && width != 4 && width != 8) return hipErrorInvalidValue;
#define CASE(S)                                                                               \
  if (stack == S)                                                                             \
    return count ? launch_extend_s<true, S>(a, packed, multi, blocks, block_waves, width, s)  \
                 : launch_extend_s<false, S>(a, packed, multi, blocks, block_waves, width, s);
  CVK_STACK_VARIANTS(CASE)
#undef CASE
  return hipErrorInvalidValue;
}

hipError_t launch_generate(const GenerateArgs& a, uint32_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_generate, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sort(const SortArgs& a, uint32_t paths, hipStream_t s) {
  if (a.key_bits > 24 || a.key_bits < 3) return hipErrorInvalidValue;
  const uint32_t nb = 1u << a.key_bits;
  const uint32_t nblocks = (nb + kSortScanBlock - 1) / kSortScanBlock;  // <= 4096
  const uint32_t pgrid = std::max(1u, std::min((paths + 255) / 256, 8192u));
  hipLaunchKernelGGL(k_sort_clear, dim3(std::min((nb + 255) / 256, 8192u)), dim3(256), 0, s,
                     a.buckets, nb);
  hipLaunchKernelGGL(k_sort_keys, dim3(pgrid), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_sort_scan, dim3(nblocks), dim3(1024), 0, s, a.buckets, nb, a.block_sums);
  // the block totals' own prefix (one block; its total lands past them)
  hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, a.block_sums, nblocks,
                     a.block_sums + kSortScanBlock);
  hipLaunchKernelGGL(k_sort_scatter, dim3(pgrid), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_shade(const ShadeArgs& a, bool multi, bool shared_scene, uint32_t max_paths,
                        hipStream_t s) {
  uint32_t blocks = (max_paths + kShadeBlock - 1) / kShadeBlock;
  if (blocks == 0) return hipSuccess;
  if (multi && shared_scene)
    hipLaunchKernelGGL((k_shade<true, true>), dim3(blocks), dim3(kShadeBlock), 0, s, a);
  else if (multi)
    hipLaunchKernelGGL((k_shade<true, false>), dim3(blocks), dim3(kShadeBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_shade<false, false>), dim3(blocks), dim3(kShadeBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_hdr_to_bitmap(const float* hdr, uint32_t* bitmap, uint32_t width,
                                uint32_t height, uint32_t num_samples, hipStream_t s) {
  uint32_t npix = width * height;
  uint32_t blocks = (npix + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(k_hdr_to_bitmap, dim3(blocks), dim3(256), 0, s, hdr, bitmap, width, height,
                     num_samples);
  return hipGetLastError();
}

hipError_t launch_prepare_tris(const float* v9, const uint32_t* order, uint32_t ntris, DevTri* out,
                               hipStream_t s) {
  uint32_t blocks = (ntris + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_prepare_tris, dim3(blocks), dim3(256), 0, s, v9, order, ntris, out);
  return hipGetLastError();
}

hipError_t launch_prepare_lights(const DevTri* tris, const uint32_t* direct, uint32_t n,
                                 DevLight* out, hipStream_t s) {
  uint32_t blocks = (n + 63) / 64;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prepare_lights, dim3(blocks), dim3(64), 0, s, tris, direct, n, out);
  return hipGetLastError();
}

hipError_t launch_fold_hdr(float* hdr, float* slots, size_t slot_stride, uint32_t count,
                           uint32_t width, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                           hipStream_t s) {
  const uint32_t npix = w * h;
  if (count == 0 || npix == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fold_hdr, dim3(std::min<uint32_t>((npix + 255) / 256, 8192)), dim3(256), 0,
                     s, hdr, slots, slot_stride, count, width, x, y, w, npix);
  return hipGetLastError();
}

hipError_t launch_fold_features(float* feat, float* slots, size_t slot_stride, uint32_t count,
                                uint32_t npix, hipStream_t s) {
  const size_t n = (size_t)npix * 8;
  if (count == 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fold_features, dim3((uint32_t)std::min<size_t>((n + 255) / 256, 8192)),
                     dim3(256), 0, s, feat, slots, slot_stride, count, n);
  return hipGetLastError();
}

// cv_bvh_intersect_nearest without cv_isect output runs the lane-refill
// traversal of the extend step over the caller's rays: the control block of
// a one-bounce step.
// One word per lane: zeros, and n in count[0].
__global__ __launch_bounds__(64) void k_query_ctrl(Ctrl* ctrl, uint32_t n) {
  static_assert(sizeof(Ctrl) % 4 == 0, "Ctrl is whole words");
  constexpr uint32_t kWords = sizeof(Ctrl) / 4;
  constexpr uint32_t kCount0 = offsetof(Ctrl, count) / 4;
  uint32_t* w = reinterpret_cast<uint32_t*>(ctrl);
  for (uint32_t k = threadIdx.x; k < kWords; k += 64) w[k] = k == kCount0 ? n : 0u;
}

hipError_t launch_query_ctrl(Ctrl* ctrl, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_query_ctrl, dim3(1), dim3(64), 0, s, ctrl, n);
  return hipGetLastError();
}

template <int S>
static hipError_t launch_ref_patch_s(const RefPatchArgs& a, bool packed, uint32_t blocks,
                                     hipStream_t s) {
  if (packed)
    hipLaunchKernelGGL((k_ref_patch<S, true>), dim3(blocks), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((k_ref_patch<S, false>), dim3(blocks), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ref_patch(const RefPatchArgs& a, bool packed, int stack, uint32_t blocks,
                            hipStream_t s) {
  if (a.npatches == 0) return hipSuccess;
#define CASE(S) \
  if (stack == S) return launch_ref_patch_s<S>(a, packed, blocks, s);
  CVK_STACK_VARIANTS(CASE)
#undef CASE
  return hipErrorInvalidValue;
}

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace cvk
