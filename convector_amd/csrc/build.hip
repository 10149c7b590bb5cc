// build.hip -- the reference BVH build (bvh.rs:91-545) on gfx950.
//
// The tree is the one InterimNode::split_recursive builds: per node and per
// axis 64 centroid bins, replaced by bins of the stably sorted triangles when
// the node has at most 32 triangles or one bin holds more than n/8 of n > 64
// triangles; the cheapest split by the TreeSurfaceAreaHeuristic, ties to the
// first axis and the first position; no split when n*120 is cheaper; the
// child with the smaller surface area first.  crystallize's depth-first
// layout of sibling pairs is reproduced exactly.  The build also runs with
// 16 or 32 bins (bins.len() changed: at most bins/2 triangles or more than
// n/8 of n > bins in one bin take the sorted bins, in chunks of
// ceil(n/bins)); below, "32" stands for bins/2 and "64" for bins.
//
// Level-synchronous:
//   k_decide (one workgroup per node of more than 32 triangles) bins the
//     three axes in LDS and picks the split;
//   k_sort sorts the axes that fall back to sorted binning (stable LSD radix
//     sort, one workgroup per node and axis);
//   k_partition reorders the node's triangle references as the reference
//     collects them (stable by bin) and creates the children.
// A node of more than kWide triangles is split over chunks of kChunk
// references: k_bin_chunks bins each chunk into a partial that k_decide
// adds up (counts and ordered-integer min / max, so the bins are exact),
// k_decide also turns the chunks' counts into each chunk's first output
// position per bin, and k_part_chunks scatters the chunks and bounds their
// children, so that the top levels, where one node holds most triangles,
// use the whole GPU.
// Nodes of at most 32 triangles are finished by k_small, one wave per node,
// with the triangles in the lanes.  k_depth (which also lists the inner
// nodes by depth) and k_bottom_up (one launch per depth over that depth's
// list, deepest first) count leaves and the area ratio statistic; k_layout
// places every node where crystallize puts it.
//
// The TreeSAH cost uses powf/log2f; their values for every triangle count are
// tabulated on the host with the C library so that near-ties resolve as in the
// reference.
#include <math.h>

#include <vector>

#include "internal.hpp"

namespace cvk {
namespace bvh {

constexpr uint32_t kLeaf = 1, kParity = 2, kIsC1 = 4, kPending = 8, kSplit = 16;
// an id that a small subtree reserved and did not use (no node)
constexpr uint32_t kHole = 32;
constexpr uint32_t kBlock = 512;
constexpr uint32_t kChunk = 2048;   // references per chunk of a wide node
constexpr uint32_t kWide = 4096;    // nodes above this many triangles are chunked
constexpr uint32_t kNarrow = 2048;  // levels with no larger node use 64-thread blocks
constexpr uint32_t kBinWords = 3 * 64 * 7;  // counts, then outer bounds, of 3 x 64 bins

struct INode {
  float box[6];   // outer: min.xyz, max.xyz
  float cbox[6];  // barycenter bounds (large nodes only)
  uint32_t begin, count;
  int32_t parent;
  int32_t child[2];
  uint32_t flags;
  uint32_t leaves;
  float sar;
  uint32_t depth;  // set by k_depth
  uint32_t axis, split_bin, left_count, uniform_mask;
  uint32_t chunk_base, nchunks;  // chunks of a wide node in the level's chunk list
};

struct Ctx {
  uint32_t node_count;
  uint32_t next_large;
  uint32_t small_count;
  uint32_t error;
  uint32_t next_wide;     // wide nodes in the next level's list
  uint32_t total_chunks;  // chunks of the current level
  uint32_t next_max;      // largest triangle count in the next level's list
  uint32_t max_depth;
  uint32_t max_leaf;
  uint32_t root_box[6];   // okeys
  uint32_t root_cbox[6];  // okeys
};

struct Args {
  const float4* A;  // bmin.xyz, c.x
  const float4* B;  // bmax.xyz, c.y
  const float* C;   // c.z
  uint32_t* buf[2];
  uint32_t* sorted[3];
  uint32_t* tmp[3];
  INode* nodes;
  Ctx* ctx;
  const uint32_t* list;
  uint32_t nlist;
  uint32_t* next_list;
  uint32_t* small_list;
  const float* tpm;   // powf(0.8f, log2f(n))
  const float* tpm1;  // powf(0.8f, log2f(n) - 1.0f)
  uint32_t mode;
  uint32_t ntris;
  uint32_t nbins;      // centroid bins per axis: 64 (the reference), 32 or 16
  uint32_t small_max;  // nodes of at most nbins / 2 triangles go to k_small
  // chunked wide nodes: per chunk the node, the bins (kBinWords), the first
  // output position per bin of the split axis (64) and the children's
  // bounds (2 sides x 12 okeys)
  uint32_t* chunk_node;
  uint32_t* part;
  uint32_t* offs;
  uint32_t* side;
};

__device__ __forceinline__ uint32_t okey(float f) {
  uint32_t u = f2u(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float okey_dec(uint32_t k) {
  return u2f((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
// Total-order min/max (the reference's f32::min / f32::max except that the
// kept zero's sign is fixed).
__device__ __forceinline__ float tmin_f(float a, float b) { return okey(b) < okey(a) ? b : a; }
__device__ __forceinline__ float tmax_f(float a, float b) { return okey(b) > okey(a) ? b : a; }
// Aabb::area (aabb.rs:97-103)
__device__ __forceinline__ float area6(const float* b) {
  float sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
  float x = sy * sz;
  float y = sz * sx;
  float z = sx * sy;
  return 2.0f * ((x + y) + z);
}
// Sort key of a barycenter coordinate; -0.0 == +0.0 as in partial_cmp.
__device__ __forceinline__ uint32_t sort_key(float c) { return okey(c == 0.0f ? 0.0f : c); }

__device__ __forceinline__ float centroid(const Args& a, uint32_t r, uint32_t axis) {
  return axis == 0 ? a.A[r].w : axis == 1 ? a.B[r].w : a.C[r];
}

// TreeSurfaceAreaHeuristic::aabb_cost (bvh.rs:452-469), constants 40/120/0.8.
__device__ __forceinline__ float tree_sah(const Args& a, float parent_area, const float* b,
                                          uint32_t n) {
  const float p = 0.8f;
  float ac_ap = area6(b) / parent_area;
  float nf = (float)n;
  float aabb_term = 1.0f + (ac_ap * (2.0f * p - nf * a.tpm[n])) / (p - (2.0f * p) * p);
  float tri_term = (nf * a.tpm1[n]) * ac_ap;
  return aabb_term * 40.0f + tri_term * 120.0f;
}

// centroid bin of bin_triangles (bvh.rs:149-156) with nb bins; NaN and nb
// clamp to nb - 1.
__device__ __forceinline__ uint32_t centroid_bin(float c, float mn, float size, uint32_t nb) {
  const float fb = (float)nb;
  float f = floorf((fb * (c - mn)) / size);
  return (f >= 0.0f && f < fb) ? (uint32_t)f : nb - 1u;
}

__device__ __forceinline__ uint32_t load_coherent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_trefs(const float* v9, uint32_t n, float4* A, float4* B,
                                               float* C, uint32_t* buf0, Ctx* ctx) {
  __shared__ uint32_t sbox[12];
  if (threadIdx.x < 12) sbox[threadIdx.x] = (threadIdx.x % 6) < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  uint32_t lb[12];
  for (int k = 0; k < 12; k++) lb[k] = (k % 6) < 3 ? 0xffffffffu : 0u;
  const float third = 1.0f / 3.0f;  // 3.0f32.recip()
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float* t = v9 + 9 * (size_t)i;
    float mn[3], mx[3], c[3];
    bool nan = false;
    for (int k = 0; k < 3; k++) {
      float x = t[k], y = t[3 + k], z = t[6 + k];
      nan |= (x != x) || (y != y) || (z != z);
      mn[k] = tmin_f(tmin_f(x, y), z);  // Aabb::enclose_points (aabb.rs:56-71)
      mx[k] = tmax_f(tmax_f(x, y), z);
      c[k] = ((x + y) + z) * third;  // Triangle::barycenter (triangle.rs:55-57)
    }
    if (nan) atomicOr(&ctx->error, 1u);
    A[i] = make_float4(mn[0], mn[1], mn[2], c[0]);
    B[i] = make_float4(mx[0], mx[1], mx[2], c[1]);
    C[i] = c[2];
    buf0[i] = i;
    for (int k = 0; k < 3; k++) {
      lb[k] = min(lb[k], okey(mn[k]));
      lb[3 + k] = max(lb[3 + k], okey(mx[k]));
      lb[6 + k] = min(lb[6 + k], okey(c[k]));
      lb[9 + k] = max(lb[9 + k], okey(c[k]));
    }
  }
  for (int k = 0; k < 12; k++) {
    if ((k % 6) < 3)
      atomicMin(&sbox[k], lb[k]);
    else
      atomicMax(&sbox[k], lb[k]);
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    uint32_t k = threadIdx.x;
    if (k < 3) {
      atomicMin(&ctx->root_box[k], sbox[k]);
      atomicMin(&ctx->root_cbox[k], sbox[6 + k]);
    } else {
      atomicMax(&ctx->root_box[k], sbox[k]);
      atomicMax(&ctx->root_cbox[k], sbox[6 + k]);
    }
  }
}

__global__ void k_root(INode* nodes, Ctx* ctx, uint32_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  INode r;
  for (int k = 0; k < 6; k++) {
    r.box[k] = okey_dec(ctx->root_box[k]);
    r.cbox[k] = okey_dec(ctx->root_cbox[k]);
  }
  r.begin = 0;
  r.count = n;
  r.parent = -1;
  r.child[0] = r.child[1] = -1;
  r.flags = 0;
  r.leaves = 0;
  r.sar = 0.0f;
  r.depth = 0;
  r.axis = r.split_bin = r.left_count = r.uniform_mask = 0;
  r.chunk_base = r.nchunks = 0;
  nodes[0] = r;
  ctx->node_count = 1;
}

// Centroid bins of refs[0, n) on the axes not in umask (bin_triangles,
// bvh.rs:149-156): counts and outer bounds as ordered integers in LDS.
template <uint32_t BS>
__device__ __forceinline__ void bin_centroids(const Args& a, const uint32_t* refs, uint32_t n,
                                              const float* cmn, const float* csz, uint32_t umask,
                                              uint32_t (*cnt)[64], uint32_t (*bx)[64][6]) {
  for (uint32_t e = threadIdx.x; e < n; e += BS) {
    uint32_t r = refs[e];
    float4 A = a.A[r], B = a.B[r];
    float cz = a.C[r];
    float c[3] = {A.w, B.w, cz};
    uint32_t kmn[3] = {okey(A.x), okey(A.y), okey(A.z)};
    uint32_t kmx[3] = {okey(B.x), okey(B.y), okey(B.z)};
    for (uint32_t axis = 0; axis < 3; axis++) {
      if ((umask >> axis) & 1u) continue;
      uint32_t b = centroid_bin(c[axis], cmn[axis], csz[axis], a.nbins);
      atomicAdd(&cnt[axis][b], 1u);
      for (int j = 0; j < 3; j++) {
        atomicMin(&bx[axis][b][j], kmn[j]);
        atomicMax(&bx[axis][b][3 + j], kmx[j]);
      }
    }
  }
}

template <uint32_t BS>
__device__ __forceinline__ void clear_bins(uint32_t (*cnt)[64], uint32_t (*bx)[64][6]) {
  for (uint32_t k = threadIdx.x; k < 3 * 64; k += BS) {
    (&cnt[0][0])[k] = 0;
    for (int j = 0; j < 6; j++) (&bx[0][0][0])[k * 6 + j] = j < 3 ? 0xffffffffu : 0u;
  }
}

// Chunk list of the level's wide nodes (one thread; only levels that have
// wide nodes run it, and those lists are short).
// One workgroup: the list in tiles of 1024 nodes, a prefix sum of the chunk
// counts per tile; each node's thread writes its own chunk entries.
__global__ __launch_bounds__(1024) void k_wide_plan(Args a) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t s_base;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (uint32_t i0 = 0; i0 < a.nlist; i0 += 1024) {
    const uint32_t i = i0 + threadIdx.x;
    uint32_t id = 0, nc = 0;
    if (i < a.nlist) {
      id = a.list[i];
      const uint32_t n = a.nodes[id].count;
      nc = n > kWide ? (n + kChunk - 1) / kChunk : 0u;
    }
    uint32_t inc = nc;
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t prior = 0, tile = 0;
    for (uint32_t w = 0; w < 16; w++) {
      prior += w < wave ? wsum[w] : 0u;
      tile += wsum[w];
    }
    const uint32_t base = s_base + prior + inc - nc;
    if (i < a.nlist) {
      a.nodes[id].chunk_base = base;
      a.nodes[id].nchunks = nc;
      for (uint32_t c = 0; c < nc; c++) a.chunk_node[base + c] = id;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base += tile;
    __syncthreads();
  }
  if (threadIdx.x == 0) a.ctx->total_chunks = s_base;
}

// Centroid bins of one chunk of a wide node into its partial.
__global__ __launch_bounds__(kBlock) void k_bin_chunks(Args a) {
  const uint32_t c = blockIdx.x;
  if (c >= a.ctx->total_chunks) return;
  const INode* nd = &a.nodes[a.chunk_node[c]];
  const uint32_t lo = (c - nd->chunk_base) * kChunk;
  const uint32_t n = min(nd->count, lo + kChunk) - lo;
  const uint32_t* refs = a.buf[(nd->flags & kParity) ? 1 : 0] + nd->begin + lo;
  __shared__ uint32_t cnt[3][64];
  __shared__ uint32_t bx[3][64][6];
  clear_bins<kBlock>(cnt, bx);
  __syncthreads();
  float cmn[3], csz[3];
  for (int k = 0; k < 3; k++) {
    cmn[k] = nd->cbox[k];
    csz[k] = nd->cbox[3 + k] - cmn[k];
  }
  bin_centroids<kBlock>(a, refs, n, cmn, csz, 0u, cnt, bx);
  __syncthreads();
  uint32_t* out = a.part + (size_t)c * kBinWords;
  for (uint32_t k = threadIdx.x; k < 3 * 64; k += kBlock) out[k] = (&cnt[0][0])[k];
  for (uint32_t k = threadIdx.x; k < 3 * 64 * 6; k += kBlock) out[192 + k] = (&bx[0][0][0])[k];
}

// ---------------------------------------------------------------------------
// k_decide: bins + split choice for nodes of more than 32 triangles
// (InterimNode::split, bvh.rs:265-322).
// mode 0: centroid bins for all axes; a node whose binning falls back to
//         sorted bins on some axis is marked pending.
// mode 1: pending nodes only; sorted (uniform) bins from k_sort's output.
// ---------------------------------------------------------------------------

// BS threads per node: 512, or 64 on levels whose nodes all have at most
// kNarrow triangles (8 times as many nodes in flight per CU; the decision's
// serial scans dominate there).
template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_decide(Args a) {
  const uint32_t id = a.list[blockIdx.x];
  INode* nd = &a.nodes[id];
  const uint32_t flags = nd->flags;
  if (a.mode == 1 && !(flags & kPending)) return;
  const uint32_t n = nd->count, begin = nd->begin;
  const uint32_t* refs = a.buf[(flags & kParity) ? 1 : 0] + begin;
  const uint32_t umask = a.mode == 1 ? nd->uniform_mask : 0u;
  __shared__ uint32_t cnt[3][64];
  __shared__ uint32_t bx[3][64][6];
  __shared__ float cost[3][64];
  __shared__ uint32_t valid_axes;
  __shared__ int s_axis;
  __shared__ uint32_t s_bin;
  const uint32_t tid = threadIdx.x;
  const uint32_t nch = nd->nchunks, cbase = nd->chunk_base;
  clear_bins<BS>(cnt, bx);
  __syncthreads();
  float cmn[3], csz[3];
  for (int k = 0; k < 3; k++) {
    cmn[k] = nd->cbox[k];
    csz[k] = nd->cbox[3 + k] - cmn[k];
  }
  if (nch) {
    // a wide node: its chunks' partials added up (integer sums and ordered
    // min / max: any order gives the same bins; 16 loads are issued before
    // they are combined, so that the latency of the partials' reads, written
    // by other XCDs, is paid once per 16 chunks)
    const uint32_t* part = a.part + (size_t)cbase * kBinWords;
    for (uint32_t k = tid; k < kBinWords; k += BS) {
      const uint32_t axis = k < 192 ? k / 64 : (k - 192) / 384;
      if ((umask >> axis) & 1u) continue;
      const uint32_t* q = part + k;
      const int op = k < 192 ? 0 : ((k - 192) % 6) < 3 ? 1 : 2;  // sum, min, max
      const uint32_t id0 = op == 1 ? 0xffffffffu : 0u;
      uint32_t v0 = id0, v1 = id0, v2 = id0, v3 = id0;
      auto acc = [&](uint32_t v, uint32_t w) {
        return op == 0 ? v + w : op == 1 ? min(v, w) : max(v, w);
      };
      uint32_t c = 0;
      for (; c + 16 <= nch; c += 16) {
        uint32_t w[16];
#pragma unroll
        for (uint32_t j = 0; j < 16; j++) w[j] = q[(size_t)(c + j) * kBinWords];
#pragma unroll
        for (uint32_t j = 0; j < 16; j += 4) {
          v0 = acc(v0, w[j]);
          v1 = acc(v1, w[j + 1]);
          v2 = acc(v2, w[j + 2]);
          v3 = acc(v3, w[j + 3]);
        }
      }
      for (; c < nch; c++) v0 = acc(v0, q[(size_t)c * kBinWords]);
      const uint32_t v = acc(acc(v0, v1), acc(v2, v3));
      if (k < 192)
        (&cnt[0][0])[k] = v;
      else
        (&bx[0][0][0])[k - 192] = v;
    }
  } else {
    bin_centroids<BS>(a, refs, n, cmn, csz, umask, cnt, bx);
  }
  if (umask) {
    // bin_triangles_uniform (bvh.rs:176-201): chunks of ceil(n/bins) of the
    // stably sorted references.
    const uint32_t tpb = (n + a.nbins - 1) / a.nbins;
    for (uint32_t axis = 0; axis < 3; axis++) {
      if (!((umask >> axis) & 1u)) continue;
      const uint32_t* srt = a.sorted[axis] + begin;
      for (uint32_t s = tid; s < n; s += BS) {
        uint32_t r = srt[s];
        float4 A = a.A[r], B = a.B[r];
        uint32_t b = s / tpb;
        atomicAdd(&cnt[axis][b], 1u);
        atomicMin(&bx[axis][b][0], okey(A.x));
        atomicMin(&bx[axis][b][1], okey(A.y));
        atomicMin(&bx[axis][b][2], okey(A.z));
        atomicMax(&bx[axis][b][3], okey(B.x));
        atomicMax(&bx[axis][b][4], okey(B.y));
        atomicMax(&bx[axis][b][5], okey(B.z));
      }
    }
  }
  __syncthreads();
  if (a.mode == 0 && tid < 64) {
    // fallback test (bvh.rs:159-169): some bin above n/8 with n > bins
    uint32_t lane = tid;
    uint32_t fb = 0;
    for (uint32_t axis = 0; axis < 3; axis++) {
      uint32_t c = cnt[axis][lane];
      uint64_t over = __ballot(c > n / 8);
      if (over && n > a.nbins) fb |= 1u << axis;
    }
    if (fb) {
      if (lane == 0) {
        nd->uniform_mask = fb;
        nd->flags = flags | kPending;
      }
    }
    if (lane == 0) valid_axes = fb ? 0xffffffffu : 0u;  // reuse as "deferred" marker
  }
  __syncthreads();
  if (a.mode == 0 && valid_axes == 0xffffffffu) return;
  // split costs (find_cheapest_split, bvh.rs:218-261), one lane per position
  const float parent_area = area6(nd->box);
  if (tid < 64) {
    const uint32_t i = tid;
    uint32_t va = 0;
    for (uint32_t axis = 0; axis < 3; axis++) {
      uint64_t ne = __ballot(cnt[axis][i] != 0);
      uint32_t first = __ffsll((unsigned long long)ne);  // first non-empty + 1
      uint32_t last = 64 - __clzll((long long)ne);        // last non-empty + 1
      bool valid = __popcll(ne) > 1;
      if (valid) va |= 1u << axis;
      float cst = 0.0f;
      if (valid && i >= first && i < last) {
        float lb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        float rb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        uint32_t lc = 0, rc = 0;
        for (uint32_t j = 0; j < 64; j++) {
          uint32_t cj = cnt[axis][j];
          if (!cj) continue;
          float* tb = j < i ? lb : rb;
          for (int k = 0; k < 3; k++) {
            tb[k] = tmin_f(tb[k], okey_dec(bx[axis][j][k]));
            tb[3 + k] = tmax_f(tb[3 + k], okey_dec(bx[axis][j][3 + k]));
          }
          if (j < i)
            lc += cj;
          else
            rc += cj;
        }
        cst = tree_sah(a, parent_area, lb, lc) + tree_sah(a, parent_area, rb, rc);
      }
      cost[axis][i] = cst;
    }
    if (i == 0) valid_axes = va;
  }
  __syncthreads();
  if (tid == 0) {
    // the reference's scan order: axes X, Y, Z; positions ascending;
    // replace only on a strictly smaller cost
    int best_axis = -1;
    uint32_t best_i = 0;
    float best = 0.0f;
    for (uint32_t axis = 0; axis < 3; axis++) {
      if (!((valid_axes >> axis) & 1u)) continue;
      int abest = -1;
      float acost = 0.0f;
      uint32_t first = 64, last = 0;
      for (uint32_t j = 0; j < 64; j++)
        if (cnt[axis][j]) {
          if (first == 64) first = j + 1;
          last = j + 1;
        }
      for (uint32_t j = first; j < last; j++) {
        float c = cost[axis][j];
        if (abest < 0 || c < acost) {
          acost = c;
          abest = (int)j;
        }
      }
      if (best_axis < 0 || acost < best) {
        best = acost;
        best_axis = (int)axis;
        best_i = (uint32_t)abest;
      }
    }
    uint32_t f = flags & ~kPending;
    s_axis = -1;
    if (best_axis < 0) {
      atomicOr(&a.ctx->error, 2u);  // assert!(!is_first), bvh.rs:303
      f |= kLeaf;
    } else if ((float)n * 120.0f < best) {
      f |= kLeaf;  // bvh.rs:311-314
    } else {
      uint32_t lc = 0;
      for (uint32_t j = 0; j < best_i; j++) lc += cnt[best_axis][j];
      nd->axis = (uint32_t)best_axis;
      nd->split_bin = best_i;
      nd->left_count = lc;
      if (a.mode == 0) nd->uniform_mask = 0;
      f |= kSplit;
      s_axis = best_axis;
      s_bin = best_i;
    }
    nd->flags = f;
  }
  if (!nch) return;
  __syncthreads();
  // a wide node split on centroid bins: where each chunk's references of
  // each bin start in the stably sorted segment (bins in order, chunks in
  // order within a bin)
  const int ax = s_axis;
  if (ax < 0 || ((umask >> ax) & 1u) || tid >= 64) return;
  uint32_t run = 0;
  for (uint32_t j = 0; j < tid; j++) run += cnt[ax][j];
  const uint32_t* part = a.part + (size_t)cbase * kBinWords + (uint32_t)ax * 64 + tid;
  uint32_t* offs = a.offs + (size_t)cbase * 64 + tid;
  // 16 chunks' counts are loaded before they are summed, so that the scan
  // does not wait for one load per chunk
  for (uint32_t c0 = 0; c0 < nch; c0 += 16) {
    uint32_t v[16];
#pragma unroll
    for (uint32_t j = 0; j < 16; j++) v[j] = c0 + j < nch ? part[(size_t)(c0 + j) * kBinWords] : 0u;
#pragma unroll
    for (uint32_t j = 0; j < 16; j++) {
      if (c0 + j < nch) offs[(size_t)(c0 + j) * 64] = run;
      run += v[j];
    }
  }
}

// ---------------------------------------------------------------------------
// Stable counting-sort pass of a segment inside one workgroup: elements are
// taken in tiles of BS in order; within a tile a lane's rank among equal
// digits comes from ballots, across waves from LDS counts.  Digit d starts
// at start[d] when start is given (a chunk of a larger sort), else after
// the segment's smaller digits.
// ---------------------------------------------------------------------------

template <int BITS, uint32_t BS, typename DigitFn, typename LoadFn>
__device__ __forceinline__ void wg_stable_pass(uint32_t n, LoadFn load, DigitFn digit_of,
                                               uint32_t* dst, uint32_t* hist, uint32_t* wcount,
                                               const uint32_t* start = nullptr) {
  constexpr uint32_t R = 1u << BITS;
  constexpr uint32_t W = BS / 64;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (start) {
    for (uint32_t k = tid; k < R; k += BS) hist[k] = start[k];
  } else {
    for (uint32_t k = tid; k < R; k += BS) hist[k] = 0;
    __syncthreads();
    for (uint32_t e = tid; e < n; e += BS) atomicAdd(&hist[digit_of(load(e))], 1u);
    __syncthreads();
    if (tid == 0) {
      uint32_t s = 0;
      for (uint32_t k = 0; k < R; k++) {
        uint32_t c = hist[k];
        hist[k] = s;
        s += c;
      }
    }
  }
  for (uint32_t k = tid; k < W * R; k += BS) wcount[k] = 0;
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t t0 = 0; t0 < n; t0 += BS) {
    uint32_t e = t0 + tid;
    bool valid = e < n;
    uint32_t v = valid ? load(e) : 0u;
    uint32_t d = valid ? digit_of(v) : 0u;
    uint64_t match = __ballot(valid);
    for (int b = 0; b < BITS; b++) {
      uint64_t m = __ballot((d >> b) & 1u);
      match &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t rank = __popcll(match & lt);
    if (valid && rank == 0) wcount[wave * R + d] = __popcll(match);
    __syncthreads();
    if (valid) {
      uint32_t prior = 0;
      for (uint32_t w = 0; w < wave; w++) prior += wcount[w * R + d];
      dst[hist[d] + prior + rank] = v;
    }
    __syncthreads();
    for (uint32_t k = tid; k < R; k += BS) {
      uint32_t s = 0;
      for (uint32_t w = 0; w < W; w++) {
        s += wcount[w * R + k];
        wcount[w * R + k] = 0;
      }
      hist[k] += s;
    }
    __syncthreads();
  }
}

// Sorted binning input: a stable sort of the segment by barycenter
// coordinate (bvh.rs:176-190), as 4 stable 8-bit LSD passes.
__global__ __launch_bounds__(kBlock) void k_sort(Args a) {
  const uint32_t id = a.list[blockIdx.x / 3];
  const uint32_t axis = blockIdx.x % 3;
  const INode* nd = &a.nodes[id];
  if (!(nd->flags & kPending) || !((nd->uniform_mask >> axis) & 1u)) return;
  const uint32_t n = nd->count, begin = nd->begin;
  const uint32_t* refs = a.buf[(nd->flags & kParity) ? 1 : 0] + begin;
  uint32_t* X = a.tmp[axis] + begin;
  uint32_t* Y = a.sorted[axis] + begin;
  __shared__ uint32_t hist[256];
  __shared__ uint32_t wcount[(kBlock / 64) * 256];
  for (int pass = 0; pass < 4; pass++) {
    const uint32_t* src = pass == 0 ? refs : (pass & 1) ? X : Y;
    uint32_t* dst = (pass & 1) ? Y : X;
    const uint32_t shift = 8 * pass;
    auto load = [&](uint32_t e) { return pass == 0 ? src[e] : load_coherent(src + e); };
    auto digit = [&](uint32_t r) { return (sort_key(centroid(a, r, axis)) >> shift) & 255u; };
    wg_stable_pass<8, kBlock>(n, load, digit, dst, hist, wcount);
    __threadfence();
    __syncthreads();
  }
}

// Outer and barycenter bounds of the two children over order[lo, hi) (side
// 0: position below L when uniform, else bin below split), reduced into
// cb (LDS, initialised by the caller) with one atomic per wave and value.
template <uint32_t BS, typename BinFn>
__device__ __forceinline__ void side_bounds(const Args& a, const uint32_t* order, uint32_t lo,
                                            uint32_t hi, bool uniform, uint32_t L, uint32_t split,
                                            BinFn bin_of, uint32_t (*cb)[12]) {
  uint32_t lk[2][12];
  for (int s = 0; s < 2; s++)
    for (int k = 0; k < 12; k++) lk[s][k] = (k % 6) < 3 ? 0xffffffffu : 0u;
  for (uint32_t e = lo + threadIdx.x; e < hi; e += BS) {
    uint32_t r = order[e];
    int side = uniform ? (e < L ? 0 : 1) : (bin_of(r) < split ? 0 : 1);
    float4 A = a.A[r], B = a.B[r];
    float cz = a.C[r];
    uint32_t* l = lk[side];
    l[0] = min(l[0], okey(A.x));
    l[1] = min(l[1], okey(A.y));
    l[2] = min(l[2], okey(A.z));
    l[3] = max(l[3], okey(B.x));
    l[4] = max(l[4], okey(B.y));
    l[5] = max(l[5], okey(B.z));
    l[6] = min(l[6], okey(A.w));
    l[7] = min(l[7], okey(B.w));
    l[8] = min(l[8], okey(cz));
    l[9] = max(l[9], okey(A.w));
    l[10] = max(l[10], okey(B.w));
    l[11] = max(l[11], okey(cz));
  }
  for (int s = 0; s < 2; s++)
    for (int k = 0; k < 12; k++) {
      uint32_t v = lk[s][k];
      // wave reduction first to cut LDS atomics
      for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(v, off);
        v = (k % 6) < 3 ? min(v, o) : max(v, o);
      }
      if ((threadIdx.x & 63) == 0) {
        if ((k % 6) < 3)
          atomicMin(&cb[s][k], v);
        else
          atomicMax(&cb[s][k], v);
      }
    }
  __syncthreads();
}

// One chunk of a wide node that splits: its references go to the positions
// k_decide computed (a stable sort by bin, or the sorted order's copy for a
// uniform axis), and its part of the children's bounds to its side partial.
__global__ __launch_bounds__(kBlock) void k_part_chunks(Args a) {
  const uint32_t c = blockIdx.x;
  if (c >= a.ctx->total_chunks) return;
  const INode* nd = &a.nodes[a.chunk_node[c]];
  if (!(nd->flags & kSplit)) return;
  const uint32_t n = nd->count, begin = nd->begin;
  const uint32_t lo = (c - nd->chunk_base) * kChunk, hi = min(n, lo + kChunk);
  const uint32_t par = (nd->flags & kParity) ? 1u : 0u;
  const uint32_t* src = a.buf[par] + begin;
  uint32_t* dst = a.buf[par ^ 1u] + begin;
  const uint32_t axis = nd->axis, split = nd->split_bin, L = nd->left_count;
  const bool uniform = (nd->uniform_mask >> axis) & 1u;
  __shared__ uint32_t hist[64];
  __shared__ uint32_t wcount[(kBlock / 64) * 64];
  __shared__ uint32_t cb[2][12];
  const uint32_t tid = threadIdx.x;
  if (tid < 24) (&cb[0][0])[tid] = ((tid % 12) % 6) < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  const float mn = nd->cbox[axis], size = nd->cbox[3 + axis] - mn;
  auto bin_of = [&](uint32_t r) {
    return centroid_bin(centroid(a, r, axis), mn, size, a.nbins);
  };
  if (uniform) {
    const uint32_t* srt = a.sorted[axis] + begin;
    for (uint32_t s = lo + tid; s < hi; s += kBlock) dst[s] = srt[s];
  } else {
    auto load = [&](uint32_t e) { return src[lo + e]; };
    wg_stable_pass<6, kBlock>(hi - lo, load, bin_of, dst, hist, wcount, a.offs + (size_t)c * 64);
  }
  side_bounds<kBlock>(a, uniform ? a.sorted[axis] + begin : src, lo, hi, uniform, L, split, bin_of,
                     cb);
  if (tid < 24) a.side[(size_t)c * 24 + tid] = (&cb[0][0])[tid];
}

// Reorders the segment as split() collects the children (bins in order,
// references in push order) and creates the children.
template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_partition(Args a) {
  const uint32_t id = a.list[blockIdx.x];
  INode* nd = &a.nodes[id];
  const uint32_t flags = nd->flags;
  if (!(flags & kSplit)) return;
  const uint32_t n = nd->count, begin = nd->begin;
  const uint32_t par = (flags & kParity) ? 1u : 0u;
  const uint32_t* src = a.buf[par] + begin;
  uint32_t* dst = a.buf[par ^ 1u] + begin;
  const uint32_t axis = nd->axis, split = nd->split_bin, L = nd->left_count;
  const bool uniform = (nd->uniform_mask >> axis) & 1u;
  __shared__ uint32_t hist[64];
  __shared__ uint32_t wcount[(BS / 64) * 64];
  __shared__ uint32_t cb[2][12];  // per side: outer min.xyz max.xyz, barycenter min.xyz max.xyz
  const uint32_t tid = threadIdx.x;
  if (tid < 24) (&cb[0][0])[tid] = ((tid % 12) % 6) < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  const float mn = nd->cbox[axis], size = nd->cbox[3 + axis] - mn;
  auto bin_of = [&](uint32_t r) {
    return centroid_bin(centroid(a, r, axis), mn, size, a.nbins);
  };
  if (nd->nchunks) {
    // a wide node: k_part_chunks reordered it; its chunks' bounds are added
    // up here
    // (ordered min / max: any order gives the same bounds; BS / 32 groups of
    // 24 lanes take every (BS / 32)-th chunk, then 24 lanes combine them)
    const uint32_t* sp = a.side + (size_t)nd->chunk_base * 24;
    __shared__ uint32_t red[BS / 32][24];
    const uint32_t k = tid % 24, g = tid / 24;  // 21 groups of 24 lanes
    const bool lo = ((k % 12) % 6) < 3;
    if (g < BS / 32) {
      uint32_t v = lo ? 0xffffffffu : 0u;
      for (uint32_t c = g; c < nd->nchunks; c += BS / 32) {
        const uint32_t w = sp[(size_t)c * 24 + k];
        v = lo ? min(v, w) : max(v, w);
      }
      red[g][k] = v;
    }
    __syncthreads();
    if (tid < 24) {
      uint32_t v = lo ? 0xffffffffu : 0u;
      for (uint32_t j = 0; j < BS / 32; j++) v = lo ? min(v, red[j][tid]) : max(v, red[j][tid]);
      (&cb[0][0])[tid] = v;
    }
    __syncthreads();
  } else {
    if (uniform) {
      const uint32_t* srt = a.sorted[axis] + begin;
      for (uint32_t s = tid; s < n; s += BS) dst[s] = srt[s];
    } else {
      auto load = [&](uint32_t e) { return src[e]; };
      wg_stable_pass<6, BS>(n, load, bin_of, dst, hist, wcount);
    }
    side_bounds<BS>(a, uniform ? a.sorted[axis] + begin : src, 0, n, uniform, L, split, bin_of, cb);
  }
  if (tid == 0) {
    uint32_t c = atomicAdd(&a.ctx->node_count, 2u);
    INode ch[2];
    for (int s = 0; s < 2; s++) {
      for (int k = 0; k < 6; k++) {
        ch[s].box[k] = okey_dec(cb[s][k]);
        ch[s].cbox[k] = okey_dec(cb[s][6 + k]);
      }
      ch[s].begin = begin + (s == 0 ? 0 : L);
      ch[s].count = s == 0 ? L : n - L;
      ch[s].parent = (int32_t)id;
      ch[s].child[0] = ch[s].child[1] = -1;
      ch[s].flags = (par ^ 1u) ? kParity : 0u;
      ch[s].leaves = 0;
      ch[s].sar = 0.0f;
      ch[s].depth = 0;
      ch[s].axis = ch[s].split_bin = ch[s].left_count = ch[s].uniform_mask = 0;
      ch[s].chunk_base = ch[s].nchunks = 0;
    }
    // smaller area first (bvh.rs:337-345)
    bool swap = area6(ch[0].box) > area6(ch[1].box);
    uint32_t first = swap ? c + 1 : c, second = swap ? c : c + 1;
    ch[swap ? 0 : 1].flags |= kIsC1;
    for (int s = 0; s < 2; s++) {
      uint32_t cid = c + s;
      if (ch[s].count == 1) ch[s].flags |= kLeaf;
      a.nodes[cid] = ch[s];
      if (ch[s].count > a.small_max) {
        uint32_t slot = atomicAdd(&a.ctx->next_large, 1u);
        a.next_list[slot] = cid;
        if (ch[s].count > kWide) atomicAdd(&a.ctx->next_wide, 1u);
        atomicMax(&a.ctx->next_max, ch[s].count);
      } else if (ch[s].count >= 2) {
        uint32_t slot = atomicAdd(&a.ctx->small_count, 1u);
        a.small_list[slot] = cid;
      }
    }
    nd->child[0] = (int32_t)first;
    nd->child[1] = (int32_t)second;
  }
}

// ---------------------------------------------------------------------------
// k_small: one wave finishes the subtree of a node of 2..32 triangles.  With
// at most 32 triangles every split uses sorted single-triangle bins
// (bvh.rs:145-147), so a node is three sorts of the lanes and two scans.
// A sort ranks each lane of the node's segment [off, off + cnt) by its key
// (ties by lane: the stable order) with cnt readlanes, and one ds_permute
// sends every lane to its rank; lanes outside the segment keep their place.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t stable_rank_src(uint32_t key, bool inside, uint32_t off,
                                                    uint32_t cnt, uint32_t lane) {
  uint32_t rank = lane;
  if (inside) {
    rank = off;
    for (uint32_t j = off; j < off + cnt; j++) {
      const uint32_t kj = __builtin_amdgcn_readlane(key, j);
      rank += (kj < key || (kj == key && j < lane)) ? 1u : 0u;
    }
  }
  // lane `rank` receives this lane's index: the source of each sorted position
  return (uint32_t)__builtin_amdgcn_ds_permute((int)(rank * 4), (int)lane);
}

// tree_sah with the power terms of counts up to 32 from LDS
__device__ __forceinline__ float tree_sah_small(const float (*tp)[33], float parent_area,
                                                const float* b, uint32_t n) {
  const float p = 0.8f;
  float ac_ap = area6(b) / parent_area;
  float nf = (float)n;
  float aabb_term = 1.0f + (ac_ap * (2.0f * p - nf * tp[0][n])) / (p - (2.0f * p) * p);
  float tri_term = (nf * tp[1][n]) * ac_ap;
  return aabb_term * 40.0f + tri_term * 120.0f;
}

__global__ __launch_bounds__(256) void k_small(Args a) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ float tp[2][33];
  if (threadIdx.x < 66) {
    const uint32_t k = threadIdx.x % 33;
    tp[threadIdx.x / 33][k] = k <= a.ntris ? (threadIdx.x < 33 ? a.tpm[k] : a.tpm1[k]) : 0.0f;
  }
  __syncthreads();
  const uint32_t si = blockIdx.x * 4 + wave;
  if (si >= a.ctx->small_count) return;
  const uint32_t root = a.small_list[si];
  const INode R = a.nodes[root];
  const uint32_t n = R.count;
  const uint32_t sbegin = R.begin;
  const uint32_t par = (R.flags & kParity) ? 1u : 0u;
  uint32_t* seg = a.buf[par] + sbegin;
  uint32_t ref = 0;
  float bmn[3] = {0, 0, 0}, bmx[3] = {0, 0, 0}, c[3] = {0, 0, 0};
  if (lane < n) {
    ref = seg[lane];
    float4 A = a.A[ref], B = a.B[ref];
    bmn[0] = A.x; bmn[1] = A.y; bmn[2] = A.z;
    bmx[0] = B.x; bmx[1] = B.y; bmx[2] = B.z;
    c[0] = A.w; c[1] = B.w; c[2] = a.C[ref];
  }
  // The subtree has at most n - 1 inner nodes, so its nodes take ids from
  // 2 (n - 1) reserved with one atomic (one counter for every wave would
  // serialise them); ids left over are written as holes at the end.
  uint32_t ids = 0;
  if (lane == 0) ids = atomicAdd(&a.ctx->node_count, 2u * (n - 1u));
  ids = __shfl(ids, 0);
  const uint32_t ids_end = ids + 2u * (n - 1u);
  // Pending nodes: off, cnt, id, parent, base flags, outer box.  Every node
  // record of the subtree is written once, when its node is popped.
  __shared__ uint32_t stk[4][40][11];
  if (lane == 0) {
    uint32_t* e = stk[wave][0];
    e[0] = 0; e[1] = n; e[2] = root; e[3] = (uint32_t)R.parent; e[4] = R.flags;
    for (int k = 0; k < 6; k++) e[5 + k] = f2u(R.box[k]);
  }
  __builtin_amdgcn_wave_barrier();
  int sp = 1;
  while (sp > 0) {
    sp--;
    // the entry is the same for the whole wave: scalar values, so that the
    // loops over the segment below are scalar loops with readlanes
    const uint32_t* e = stk[wave][sp];
    const uint32_t off = __builtin_amdgcn_readfirstlane(e[0]);
    const uint32_t cnt = __builtin_amdgcn_readfirstlane(e[1]);
    const uint32_t id = __builtin_amdgcn_readfirstlane(e[2]);
    const int32_t parent = (int32_t)__builtin_amdgcn_readfirstlane(e[3]);
    const uint32_t bflags = __builtin_amdgcn_readfirstlane(e[4]);
    float nbox[6];
    for (int k = 0; k < 6; k++) nbox[k] = u2f(e[5 + k]);
    __builtin_amdgcn_wave_barrier();
    auto write_node = [&](uint32_t flags, int32_t c0, int32_t c1) {
      if (lane != 0) return;
      INode w;
      for (int k = 0; k < 6; k++) {
        w.box[k] = nbox[k];
        w.cbox[k] = id == root ? R.cbox[k] : 0.0f;
      }
      w.begin = sbegin + off;
      w.count = cnt;
      w.parent = parent;
      w.child[0] = c0;
      w.child[1] = c1;
      w.flags = flags;
      w.leaves = 0;
      w.sar = 0.0f;
      w.depth = 0;
      w.axis = w.split_bin = w.left_count = w.uniform_mask = 0;
      w.chunk_base = w.nchunks = 0;
      a.nodes[id] = w;
    };
    if (cnt <= 1) {
      write_node(bflags | kLeaf, -1, -1);
      continue;
    }
    const float parent_area = area6(nbox);
    const bool inside = lane >= off && lane < off + cnt;
    uint32_t perm[3];
    int best_axis = -1;
    uint32_t best_i = 0;
    float best = 0.0f;
    for (uint32_t axis = 0; axis < 3; axis++) {
      const uint32_t src = stable_rank_src(sort_key(c[axis]), inside, off, cnt, lane);
      perm[axis] = src;
      float s[6];
      for (int k = 0; k < 3; k++) {
        s[k] = __shfl(bmn[k], (int)src);
        s[3 + k] = __shfl(bmx[k], (int)src);
      }
      float pre[6], suf[6];
      for (int k = 0; k < 6; k++) {
        float idv = k < 3 ? INFINITY : -INFINITY;
        pre[k] = inside ? s[k] : idv;
        suf[k] = pre[k];
      }
      // the segment lies in lanes 0..31: scans over 32 lanes
      for (uint32_t d = 1; d < 32; d <<= 1)
        for (int k = 0; k < 6; k++) {
          float up = __shfl_up(pre[k], d);
          float dn = __shfl_down(suf[k], d);
          if (lane >= d) pre[k] = k < 3 ? tmin_f(pre[k], up) : tmax_f(pre[k], up);
          if (lane + d < 32) suf[k] = k < 3 ? tmin_f(suf[k], dn) : tmax_f(suf[k], dn);
        }
      float rsuf[6];
      for (int k = 0; k < 6; k++) rsuf[k] = __shfl_down(suf[k], 1);
      float cst = 0.0f;
      if (lane >= off && lane + 1 < off + cnt) {
        uint32_t lc = lane - off + 1;
        cst = tree_sah_small(tp, parent_area, pre, lc) +
              tree_sah_small(tp, parent_area, rsuf, cnt - lc);
      }
      // first strictly smallest position; then strictly smaller than earlier axes
      int abest = -1;
      float acost = 0.0f;
      for (uint32_t k = off; k + 1 < off + cnt; k++) {
        const float ck = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cst), k));
        if (abest < 0 || ck < acost) {
          acost = ck;
          abest = (int)(k - off + 1);
        }
      }
      if (best_axis < 0 || acost < best) {
        best = acost;
        best_axis = (int)axis;
        best_i = (uint32_t)abest;
      }
    }
    if ((float)cnt * 120.0f < best) {
      write_node(bflags | kLeaf, -1, -1);
      continue;
    }
    uint32_t p = best_axis == 0 ? perm[0] : best_axis == 1 ? perm[1] : perm[2];
    ref = __shfl(ref, (int)p);
    for (int k = 0; k < 3; k++) {
      bmn[k] = __shfl(bmn[k], (int)p);
      bmx[k] = __shfl(bmx[k], (int)p);
      c[k] = __shfl(c[k], (int)p);
    }
    float cbx[2][6];
    for (int side = 0; side < 2; side++) {
      uint32_t lo = side == 0 ? off : off + best_i;
      uint32_t hi = side == 0 ? off + best_i : off + cnt;
      bool in = lane >= lo && lane < hi;
      for (int k = 0; k < 6; k++) {
        float v = k < 3 ? (in ? bmn[k] : INFINITY) : (in ? bmx[k - 3] : -INFINITY);
        for (int o = 32; o > 0; o >>= 1) {
          float w = __shfl_xor(v, o);
          v = k < 3 ? tmin_f(v, w) : tmax_f(v, w);
        }
        cbx[side][k] = v;
      }
    }
    const uint32_t cid = ids;
    ids += 2;
    // smaller area first (bvh.rs:337-345)
    const bool swap = area6(cbx[0]) > area6(cbx[1]);
    write_node((bflags & ~kLeaf) | kSplit, (int32_t)(swap ? cid + 1 : cid),
               (int32_t)(swap ? cid : cid + 1));
    if (lane == 0) {
      for (int s = 0; s < 2; s++) {
        uint32_t* q = stk[wave][sp + (s == 0 ? 1 : 0)];
        q[0] = s == 0 ? off : off + best_i;
        q[1] = s == 0 ? best_i : cnt - best_i;
        q[2] = cid + s;
        q[3] = id;
        q[4] = (par ? kParity : 0u) | (((s == 1) != swap) ? kIsC1 : 0u);
        for (int k = 0; k < 6; k++) q[5 + k] = f2u(cbx[s][k]);
      }
    }
    __builtin_amdgcn_wave_barrier();
    sp += 2;
  }
  if (lane < n) seg[lane] = ref;
  for (uint32_t h = ids + lane; h < ids_end; h += 64) {
    INode w{};
    w.flags = kHole;
    w.parent = -1;
    w.child[0] = w.child[1] = -1;
    a.nodes[h] = w;
  }
}

// ---------------------------------------------------------------------------
// Statistics and layout
// ---------------------------------------------------------------------------

// Depth of every node (a walk to the root), the tree's depth, and the
// leaves' statistics (one leaf, summed_area_ratio 0 + 0).
// Depth of every node; leaves start the bottom-up statistics; inner nodes
// are counted per depth (dcount holds dcap >= max depth + 1 words), in LDS
// per workgroup first (one global atomic per workgroup and depth: a global
// counter per depth shared by all nodes serialises).
constexpr uint32_t kDepthLds = 512;
__global__ __launch_bounds__(256) void k_depth(Args a, uint32_t ninodes, uint32_t* dcount,
                                               uint32_t dcap) {
  __shared__ uint32_t cnt[kDepthLds];
  const bool local = dcap <= kDepthLds;
  if (local)
    for (uint32_t k = threadIdx.x; k < dcap; k += 256) cnt[k] = 0;
  __syncthreads();
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  INode* nodes = a.nodes;
  if (x < ninodes && !(nodes[x].flags & kHole)) {
    uint32_t d = 0;
    for (int32_t u = nodes[x].parent; u >= 0; u = nodes[u].parent) d++;
    nodes[x].depth = d;
    if (nodes[x].flags & kLeaf) {
      nodes[x].leaves = 1;
      nodes[x].sar = 0.0f + 0.0f;
      atomicMax(&a.ctx->max_depth, d);
    } else if (d < dcap) {
      atomicAdd(local ? &cnt[d] : &dcount[d], 1u);
    }
  }
  __syncthreads();
  if (local)
    for (uint32_t k = threadIdx.x; k < dcap; k += 256)
      if (cnt[k]) atomicAdd(&dcount[k], cnt[k]);
}

// The inner nodes listed by depth: cursor[d] starts at depth d's offset;
// a workgroup reserves one range per depth and ranks its nodes in LDS.
__global__ __launch_bounds__(256) void k_depth_lists(Args a, uint32_t ninodes, uint32_t* cursor,
                                                     uint32_t dcap, uint32_t* list) {
  __shared__ uint32_t cnt[kDepthLds];
  const bool local = dcap <= kDepthLds;
  if (local)
    for (uint32_t k = threadIdx.x; k < dcap; k += 256) cnt[k] = 0;
  __syncthreads();
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  bool inner = false;
  uint32_t d = 0, rank = 0;
  if (x < ninodes) {
    const INode& n = a.nodes[x];
    inner = !(n.flags & (kLeaf | kHole)) && n.depth < dcap;
    d = n.depth;
  }
  if (inner && local) rank = atomicAdd(&cnt[d], 1u);
  __syncthreads();
  if (local)
    for (uint32_t k = threadIdx.x; k < dcap; k += 256)
      if (cnt[k]) cnt[k] = atomicAdd(&cursor[k], cnt[k]);
  __syncthreads();
  if (inner) list[local ? cnt[d] + rank : atomicAdd(&cursor[d], 1u)] = x;
}

// Leaf counts and summed_area_ratio (bvh.rs:380-386) of the inner nodes of
// one depth (listed by k_depth_lists), from their children: one launch per
// depth, deepest first, each over that depth's nodes only.
__global__ __launch_bounds__(256) void k_bottom_up(Args a, const uint32_t* list, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  INode* nodes = a.nodes;
  const uint32_t x = list[k];
  const int32_t c0 = nodes[x].child[0], c1 = nodes[x].child[1];
  const float s0 = nodes[c0].sar, s1 = nodes[c1].sar;
  const float A = area6(nodes[x].box);
  const float cc = (0.0f + s0) + s1;
  const float cs = (0.0f + area6(nodes[c0].box) / A) + area6(nodes[c1].box) / A;
  nodes[x].leaves = nodes[c0].leaves + nodes[c1].leaves;
  nodes[x].sar = cs + cc;
}

// Counters of the next level's node lists back to zero.
__global__ void k_level_reset(Ctx* ctx) {
  if (threadIdx.x == 0) {
    ctx->next_large = 0;
    ctx->next_wide = 0;
    ctx->next_max = 0;
  }
}

// crystallize (bvh.rs:389-428): node x is written at L(parent) + [x is c1],
// where L(u) = nodes.len() when u is crystallised; a leaf's triangles start
// at T(x).  Both are sums over the path to the root.
__global__ __launch_bounds__(256) void k_layout(Args a, uint32_t ninodes, DevNode* out,
                                                uint32_t* order) {
  uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= ninodes || x == 0) return;
  const INode* nodes = a.nodes;
  const INode& nx = nodes[x];
  if (nx.flags & kHole) return;
  uint32_t Lx = 0, Tx = 0, depth = 0;
  uint32_t c = x;
  for (;;) {
    int32_t u = nodes[c].parent;
    if (u < 0) break;
    depth++;
    if (nodes[c].flags & kIsC1) {
      const INode& sib = nodes[nodes[u].child[0]];
      Lx += 2 + 2 * (sib.leaves - 1);
      Tx += sib.count;
    } else {
      Lx += 2;
    }
    c = (uint32_t)u;
  }
  const bool c1 = nx.flags & kIsC1;
  const INode& par = nodes[nx.parent];
  uint32_t Lp = Lx - 2 - (c1 ? 2 * (nodes[par.child[0]].leaves - 1) : 0);
  uint32_t slot = Lp + (c1 ? 1u : 0u);
  const bool leaf = nx.flags & kLeaf;
  uint32_t index = leaf ? Tx : Lx;
  uint32_t len = leaf ? nx.count : 0u;
  DevNode d;
  d.a = make_float4(nx.box[0], nx.box[1], nx.box[2], nx.box[3]);
  d.b = make_float4(nx.box[4], nx.box[5], u2f(index), u2f(len));
  out[slot] = d;
  if (leaf) {
    const uint32_t* src = a.buf[(nx.flags & kParity) ? 1 : 0] + nx.begin;
    for (uint32_t k = 0; k < nx.count; k++) order[Tx + k] = src[k];
    atomicMax(&a.ctx->max_leaf, nx.count);
    atomicMax(&a.ctx->max_depth, depth);
  }
}

}  // namespace bvh

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------

#define BVH_CHECK(x)                \
  do {                              \
    hipError_t e_ = (x);            \
    if (e_ != hipSuccess) {         \
      cleanup();                    \
      return e_;                    \
    }                               \
  } while (0)

hipError_t build_reference_bvh(const float* d_v9, uint32_t ntris, DevNode* nodes_out,
                               uint32_t* order_out, BuildResult* res, hipStream_t s,
                               uint32_t nbins) {
  using namespace bvh;
  res->error = 0;
  if (nbins != 16 && nbins != 32 && nbins != 64) return hipErrorInvalidValue;
  std::vector<void*> allocs;
  auto cleanup = [&]() {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
  };
  auto alloc = [&](void** p, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e == hipSuccess) allocs.push_back(*p);
    return e;
  };
  const size_t N = ntris;
  Args a{};
  a.ntris = ntris;
  a.nbins = nbins;
  a.small_max = nbins / 2;
  float4 *A, *B;
  float* C;
  BVH_CHECK(alloc((void**)&A, N * sizeof(float4)));
  BVH_CHECK(alloc((void**)&B, N * sizeof(float4)));
  BVH_CHECK(alloc((void**)&C, N * sizeof(float)));
  a.A = A;
  a.B = B;
  a.C = C;
  for (int k = 0; k < 2; k++) BVH_CHECK(alloc((void**)&a.buf[k], N * 4));
  for (int k = 0; k < 3; k++) {
    BVH_CHECK(alloc((void**)&a.sorted[k], N * 4));
    BVH_CHECK(alloc((void**)&a.tmp[k], N * 4));
  }
  // Node ids: the levels take 1 + 2 S for S splits of nodes of more than 32
  // triangles, and each small subtree of n triangles reserves 2 (n - 1).
  // Those S splits form a full binary tree whose leaves are the small
  // subtrees, the single triangles and the unsplit large nodes, so S is
  // their number - 1 and all ids together are 2 (singles + unsplit large
  // nodes + triangles of small subtrees) - 1 <= 2 N - 1.
  const size_t ninodes_cap = 2 * N + 2;
  BVH_CHECK(alloc((void**)&a.nodes, ninodes_cap * sizeof(INode)));
  Ctx* ctx;
  BVH_CHECK(alloc((void**)&ctx, sizeof(Ctx)));
  a.ctx = ctx;
  const size_t list_cap = N / (a.small_max + 1) + 2;
  uint32_t *listA, *listB, *small;
  BVH_CHECK(alloc((void**)&listA, list_cap * 4));
  BVH_CHECK(alloc((void**)&listB, list_cap * 4));
  BVH_CHECK(alloc((void**)&small, (N / 2 + 2) * 4));
  a.small_list = small;
  // chunks of the wide nodes of one level: their triangles are disjoint, so
  // at most N / kChunk + (number of wide nodes) chunks
  const size_t max_chunks = N / kChunk + N / kWide + 2;
  BVH_CHECK(alloc((void**)&a.chunk_node, max_chunks * 4));
  BVH_CHECK(alloc((void**)&a.part, max_chunks * kBinWords * 4));
  BVH_CHECK(alloc((void**)&a.offs, max_chunks * 64 * 4));
  BVH_CHECK(alloc((void**)&a.side, max_chunks * 24 * 4));
  // TreeSAH power tables with the C library (same values as the CPU build)
  std::vector<float> tpm(N + 1), tpm1(N + 1);
  for (size_t n = 1; n <= N; n++) {
    float m = log2f((float)n);
    tpm[n] = powf(0.8f, m);
    tpm1[n] = powf(0.8f, m - 1.0f);
  }
  tpm[0] = tpm1[0] = 0.0f;
  float *d_tpm, *d_tpm1;
  BVH_CHECK(alloc((void**)&d_tpm, (N + 1) * 4));
  BVH_CHECK(alloc((void**)&d_tpm1, (N + 1) * 4));
  BVH_CHECK(hipMemcpyAsync(d_tpm, tpm.data(), (N + 1) * 4, hipMemcpyHostToDevice, s));
  BVH_CHECK(hipMemcpyAsync(d_tpm1, tpm1.data(), (N + 1) * 4, hipMemcpyHostToDevice, s));
  a.tpm = d_tpm;
  a.tpm1 = d_tpm1;

  hipEvent_t ev0, ev1;
  BVH_CHECK(hipEventCreate(&ev0));
  BVH_CHECK(hipEventCreate(&ev1));
  BVH_CHECK(hipEventRecord(ev0, s));

  Ctx init{};
  for (int k = 0; k < 6; k++) {
    init.root_box[k] = k < 3 ? 0xffffffffu : 0u;
    init.root_cbox[k] = k < 3 ? 0xffffffffu : 0u;
  }
  BVH_CHECK(hipMemcpyAsync(ctx, &init, sizeof(Ctx), hipMemcpyHostToDevice, s));
  {
    uint32_t blocks = (ntris + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_trefs, dim3(blocks), dim3(256), 0, s, d_v9, ntris, A, B, C, a.buf[0],
                       ctx);
    BVH_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_root, dim3(1), dim3(64), 0, s, a.nodes, ctx, ntris);
    BVH_CHECK(hipGetLastError());
  }
  uint32_t zero = 0, one = 1;
  uint32_t nlist = 0;
  if (ntris > a.small_max) {
    BVH_CHECK(hipMemcpyAsync(listA, &zero, 4, hipMemcpyHostToDevice, s));
    nlist = 1;
  } else if (ntris >= 2) {
    BVH_CHECK(hipMemcpyAsync(small, &zero, 4, hipMemcpyHostToDevice, s));
    BVH_CHECK(hipMemcpyAsync(&ctx->small_count, &one, 4, hipMemcpyHostToDevice, s));
  }
  uint32_t* cur = listA;
  uint32_t* nxt = listB;
  bool wide = ntris > kWide;
  uint32_t level_max = ntris;
  uint32_t levels = 0;  // iterations of the level phase
  while (nlist > 0) {
    levels++;
    const bool narrow = level_max <= kNarrow;
    a.list = cur;
    a.nlist = nlist;
    a.next_list = nxt;
    a.mode = 0;
    if (wide) {
      hipLaunchKernelGGL(k_wide_plan, dim3(1), dim3(1024), 0, s, a);
      BVH_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_bin_chunks, dim3((uint32_t)max_chunks), dim3(kBlock), 0, s, a);
      BVH_CHECK(hipGetLastError());
    }
    auto decide = [&]() {
      if (narrow)
        hipLaunchKernelGGL(k_decide<64>, dim3(nlist), dim3(64), 0, s, a);
      else
        hipLaunchKernelGGL(k_decide<kBlock>, dim3(nlist), dim3(kBlock), 0, s, a);
    };
    decide();
    BVH_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sort, dim3(nlist * 3), dim3(kBlock), 0, s, a);
    BVH_CHECK(hipGetLastError());
    a.mode = 1;
    decide();
    BVH_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_level_reset, dim3(1), dim3(64), 0, s, ctx);
    BVH_CHECK(hipGetLastError());
    if (wide) {
      hipLaunchKernelGGL(k_part_chunks, dim3((uint32_t)max_chunks), dim3(kBlock), 0, s, a);
      BVH_CHECK(hipGetLastError());
    }
    // (64-thread blocks measured slower here: the stable pass's tiles)
    hipLaunchKernelGGL(k_partition<kBlock>, dim3(nlist), dim3(kBlock), 0, s, a);
    BVH_CHECK(hipGetLastError());
    uint32_t hctl[7];
    BVH_CHECK(hipMemcpyAsync(hctl, ctx, 28, hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipStreamSynchronize(s));
    if (hctl[3]) break;
    nlist = hctl[1];
    wide = hctl[4] > 0;
    level_max = hctl[6];
    uint32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
  Ctx hc;
  BVH_CHECK(hipMemcpyAsync(&hc, ctx, sizeof(Ctx), hipMemcpyDeviceToHost, s));
  BVH_CHECK(hipStreamSynchronize(s));
  if (hc.error == 0 && hc.small_count > 0) {
    hipLaunchKernelGGL(k_small, dim3((hc.small_count + 3) / 4), dim3(256), 0, s, a);
    BVH_CHECK(hipGetLastError());
  }
  BVH_CHECK(hipMemcpyAsync(&hc, ctx, sizeof(Ctx), hipMemcpyDeviceToHost, s));
  BVH_CHECK(hipStreamSynchronize(s));
  uint32_t ninodes = hc.node_count;
  INode root_h;
  if (hc.error == 0) {
    BVH_CHECK(hipMemcpyAsync(&root_h, a.nodes, sizeof(INode), hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipStreamSynchronize(s));
  }
  if (hc.error & 1u) {
    res->error = 9;  // CV_ERR_NAN
  } else if (hc.error & 2u) {
    res->error = 8;  // CV_ERR_BVH_NO_SPLIT
  } else if (root_h.child[0] < 0) {
    res->error = 8;  // the root must split (bvh.rs:512)
  }
  if (res->error) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    cleanup();
    return hipSuccess;
  }
  uint32_t blocks = (ninodes + 255) / 256;
  // depths: the level phase's nodes lie at most `levels` deep and a small
  // subtree of at most 32 triangles adds at most 31 levels
  const uint32_t dcap = levels + 34;
  uint32_t *dcount, *dlist;
  BVH_CHECK(alloc((void**)&dcount, (size_t)dcap * 4));
  BVH_CHECK(alloc((void**)&dlist, (size_t)ninodes * 4));
  BVH_CHECK(hipMemsetAsync(dcount, 0, (size_t)dcap * 4, s));
  hipLaunchKernelGGL(k_depth, dim3(blocks), dim3(256), 0, s, a, ninodes, dcount, dcap);
  BVH_CHECK(hipGetLastError());
  std::vector<uint32_t> hcount(dcap), hoff(dcap);
  BVH_CHECK(hipMemcpyAsync(hcount.data(), dcount, (size_t)dcap * 4, hipMemcpyDeviceToHost, s));
  BVH_CHECK(hipMemcpyAsync(&hc, ctx, sizeof(Ctx), hipMemcpyDeviceToHost, s));
  BVH_CHECK(hipStreamSynchronize(s));
  if (hc.max_depth >= dcap) {
    cleanup();
    return hipErrorInvalidValue;
  }
  uint32_t run = 0;
  for (uint32_t d = 0; d < dcap; d++) {
    hoff[d] = run;
    run += hcount[d];
  }
  BVH_CHECK(hipMemcpyAsync(dcount, hoff.data(), (size_t)dcap * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_depth_lists, dim3(blocks), dim3(256), 0, s, a, ninodes, dcount, dcap,
                     dlist);
  BVH_CHECK(hipGetLastError());
  for (uint32_t d = hc.max_depth; d-- > 0;) {
    if (!hcount[d]) continue;
    hipLaunchKernelGGL(k_bottom_up, dim3((hcount[d] + 255) / 256), dim3(256), 0, s, a,
                       dlist + hoff[d], hcount[d]);
    BVH_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_layout, dim3(blocks), dim3(256), 0, s, a, ninodes, nodes_out, order_out);
  BVH_CHECK(hipGetLastError());
  BVH_CHECK(hipEventRecord(ev1, s));
  BVH_CHECK(hipMemcpyAsync(&hc, ctx, sizeof(Ctx), hipMemcpyDeviceToHost, s));
  BVH_CHECK(hipMemcpyAsync(&root_h, a.nodes, sizeof(INode), hipMemcpyDeviceToHost, s));
  BVH_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  res->leaves = root_h.leaves;
  res->num_nodes = 2 * (root_h.leaves - 1);
  res->depth = hc.max_depth;
  res->max_leaf = hc.max_leaf;
  res->avg_area_ratio = root_h.sar / (float)(2 * root_h.leaves - 1);
  res->avg_tris_per_leaf = (float)ntris / (float)root_h.leaves;
  res->build_ms = ms;
  cleanup();
  return hipSuccess;
}

}  // namespace cvk
