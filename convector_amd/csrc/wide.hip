// wide.hip -- collapses a scene's reference BVH (node pairs, bvh.rs:20-30)
// into a tree of 4 or 8 children per node for the wide traversal of
// kernels.hip (no reference counterpart: a layout the autotuner may pick
// for speed, with results equal to the reference's except at exact ties).
//
// Each wide node takes the two children of a binary inner node and keeps
// replacing the inner child of largest surface area by its two children
// until it holds W children or only leaves.  Leaves are the reference's
// leaves, unchanged: a child reference names the same triangle range of the
// BVH order.  Nodes are laid out depth first, a parent before its children.
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "host.hpp"

namespace cvh {

namespace {

constexpr uint32_t kWideLeaf = 0x80000000u;

float area(const cv_bvh_node& n) {
  const float dx = n.max[0] - n.min[0], dy = n.max[1] - n.min[1], dz = n.max[2] - n.min[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct Collapser {
  const std::vector<cv_bvh_node>& bin;
  const uint32_t W;
  const uint32_t f4;  // float4 per node
  std::vector<float4> out;
  uint32_t depth = 0;  // wide levels below the root (the root is level 1)

  // The wide node over binary nodes c0, c1 (siblings) at `level`; returns
  // its index.
  uint32_t collapse(uint32_t c0, uint32_t c1, uint32_t level) {
    depth = level > depth ? level : depth;
    uint32_t kids[8] = {c0, c1};
    uint32_t nk = 2;
    while (nk < W) {
      int best = -1;
      float best_area = -1.0f;
      for (uint32_t k = 0; k < nk; k++) {
        const cv_bvh_node& n = bin[kids[k]];
        if (n.len == 0 && area(n) > best_area) {
          best_area = area(n);
          best = (int)k;
        }
      }
      if (best < 0) break;
      const uint32_t b = kids[best];
      kids[best] = bin[b].index;
      kids[nk++] = bin[b].index + 1;
    }
    const uint32_t me = (uint32_t)(out.size() / f4);
    out.resize(out.size() + f4);
    uint32_t refs[8];
    for (uint32_t k = 0; k < nk; k++) {
      const cv_bvh_node& n = bin[kids[k]];
      refs[k] = n.len ? kWideLeaf | (n.len << 26) | n.index
                      : collapse(n.index, n.index + 1, level + 1);
    }
    // SoA quads: min.x, max.x, min.y, max.y, min.z, max.z, references
    const uint32_t Q = W / 4;
    float* f = reinterpret_cast<float*>(out.data() + (size_t)me * f4);
    uint32_t* u = reinterpret_cast<uint32_t*>(f);
    memset(f, 0, f4 * sizeof(float4));
    for (uint32_t k = 0; k < W; k++) {
      const uint32_t q = k / 4, j = k % 4;
      float lo[3], hi[3];
      uint32_t r;
      if (k < nk) {
        const cv_bvh_node& n = bin[kids[k]];
        for (int a = 0; a < 3; a++) {
          lo[a] = n.min[a];
          hi[a] = n.max[a];
        }
        r = refs[k];
      } else {
        // an empty slot: a point box far outside any scene, never entered
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = 1.0e30f;
        r = kWideLeaf;
      }
      for (int a = 0; a < 3; a++) {
        f[((2 * a) * Q + q) * 4 + j] = lo[a];
        f[((2 * a + 1) * Q + q) * 4 + j] = hi[a];
      }
      u[(6 * Q + q) * 4 + j] = r;
    }
    return me;
  }
};

}  // namespace

cv_status ensure_wide_locked(cv_scene* s, uint32_t width) {
  if (width == 2) return CV_OK;
  if (width != 4 && width != 8) return fail(CV_ERR_INVALID_ARGUMENT, "bvh width must be 2, 4 or 8");
  const int slot = width == 4 ? 0 : 1;
  if (s->wide[slot].p) return CV_OK;
  if (s->info.max_leaf > 31 || s->ntris >= (1u << 26))
    return fail(CV_ERR_INVALID_ARGUMENT,
                "wide BVHs need leaves of at most 31 triangles and fewer than 2^26 triangles");
  std::vector<cv_bvh_node> bin(s->info.num_nodes);
  HIPTRY(hipStreamSynchronize(s->stream), "sync");
  HIPTRY(hipMemcpy(bin.data(), s->nodes.p, bin.size() * sizeof(cv_bvh_node), hipMemcpyDeviceToHost),
         "download bvh");
  Collapser c{bin, width, width == 4 ? 8u : 16u, {}, 0};
  c.out.reserve((size_t)bin.size() / (width == 4 ? 3 : 6) * c.f4 + c.f4);
  c.collapse(0, 1, 1);
  HIPTRY(s->wide[slot].alloc(c.out.size()), "alloc wide bvh");
  HIPTRY(hipMemcpy(s->wide[slot].p, c.out.data(), c.out.size() * sizeof(float4),
                   hipMemcpyHostToDevice),
         "upload wide bvh");
  s->wide_nodes[slot] = (uint32_t)(c.out.size() / c.f4);
  s->wide_depth[slot] = c.depth;
  return CV_OK;
}

cv_status ensure_wide(cv_scene* s, uint32_t width) {
  if (width == 2) return CV_OK;
  std::lock_guard<std::mutex> g(s->mu);
  return ensure_wide_locked(s, width);
}

// Stack entries a traversal of the given width can hold, less 4 (the spill
// areas hold depth + 4 entries per lane): the binary traversal keeps at most
// one entry per level below the root, the nearest-first wide traversal at
// most W - 1 per wide level on the current path.
uint32_t stack_depth(const cv_scene* s, uint32_t width) {
  if (width == 4) return 3 * s->wide_depth[0];
  if (width == 8) return 7 * s->wide_depth[1];
  return s->info.depth;
}

const cvk::DevNode* traversal_nodes(const cv_scene* s, uint32_t width) {
  if (width == 4) return reinterpret_cast<const cvk::DevNode*>(s->wide[0].p);
  if (width == 8) return reinterpret_cast<const cvk::DevNode*>(s->wide[1].p);
  return s->nodes.p;
}

}  // namespace cvh
