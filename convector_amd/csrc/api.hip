// api.cpp -- the C ABI of include/convector.h: meshes, scenes (GPU BVH
// build + device upload), batched nearest-hit queries and the wavefront
// renderer.  Every compute step runs in the kernels of kernels.hip and
// build.hip; this file only parses input, owns device memory and orders
// launches on one HIP stream per object.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/convector.h"
#include "internal.hpp"

using cvk::DevLight;
using cvk::DevNode;
using cvk::DevTri;

namespace {

thread_local std::string g_last_error;

cv_status fail(cv_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

cv_status hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(e == hipErrorOutOfMemory ? CV_ERR_OUT_OF_MEMORY : CV_ERR_DEVICE, m);
}

#define HIPTRY(x, what)                   \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

int g_num_cus = 0;

cv_status ensure_device() {
  if (g_num_cus > 0) return CV_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(CV_ERR_NO_DEVICE, "no HIP device available");
  int dev = 0;
  HIPTRY(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  HIPTRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  g_num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  return CV_OK;
}

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
};

// ----------------------------------------------------------- kernels local

__global__ void k_gather_u32(const uint32_t* src, const uint32_t* order, uint32_t n,
                             uint32_t stride, uint32_t* dst) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    for (uint32_t j = 0; j < stride; j++) dst[(size_t)k * stride + j] = src[(size_t)order[k] * stride + j];
}

// Scene::from_meshes's direct_sample list (scene.rs:114-129): indices in BVH
// order whose material has bit 30 set, ascending.  One workgroup, in order.
__global__ __launch_bounds__(1024) void k_direct_list(const uint32_t* mat, uint32_t n,
                                                      uint32_t* out, uint32_t* count) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t t0 = 0; t0 < n; t0 += 1024) {
    uint32_t k = t0 + threadIdx.x;
    bool f = k < n && (mat[k] & 0x40000000u) == 0x40000000u;
    uint64_t b = __ballot(f);
    if (lane == 0) wsum[wave] = __popcll(b);
    __syncthreads();
    uint32_t prior = 0;
    for (uint32_t w = 0; w < wave; w++) prior += wsum[w];
    if (f) out[base + prior + __popcll(b & ((1ull << lane) - 1ull))] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int w = 0; w < 16; w++) s += wsum[w];
      base += s;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = base;
}

uint32_t grid_for(size_t n, uint32_t block, uint32_t cap = 1u << 20) {
  size_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (uint32_t)b;
}

}  // namespace

// ------------------------------------------------------------------ objects

struct cv_mesh {
  std::vector<float> v9;
  std::vector<float> uv6;
  std::vector<uint32_t> mat;
  size_t ntris() const { return mat.size(); }
};

struct cv_scene {
  hipStream_t stream = nullptr;
  uint32_t ntris = 0;
  DevBuf<DevNode> nodes;
  DevBuf<DevTri> tris;
  DevBuf<uint32_t> tri_mat;
  DevBuf<float> tri_uv;
  DevBuf<uint32_t> tri_src;
  DevBuf<uint32_t> direct;
  uint32_t ndirect = 0;
  cv_bvh_info info{};
  bool packed = true;
  int stack = 16;
  cvk::CameraParams cam{};
  bool has_override = false;
  float emission[3] = {0, 0, 0};
  DevBuf<uint2> spill;
  uint32_t persistent_blocks = 0;
  uint32_t refill = 16;
  uint32_t fetch_batches = 1;
  DevBuf<uint32_t> cursor;
  std::mutex mu;
};

struct cv_renderer {
  cv_scene* scene = nullptr;
  uint32_t width = 0, height = 0;
  cv_render_opts opts{};
  hipStream_t stream = nullptr;
  DevBuf<float4> r0[2], r1[2], col[2];
  DevBuf<uint64_t> rng[2];
  DevBuf<uint4> hits;
  DevBuf<cvk::Ctrl> ctrl;
  DevBuf<DevLight> lights;
  DevBuf<float> hdr;
  DevBuf<uint32_t> bitmap, gbuffer;
  cv_render_stats stats{};
  std::vector<hipEvent_t> events;
  std::mutex mu;
};

// -------------------------------------------------------------- OBJ loader

namespace {

// SVector3::cross (cross_fma) and norm_squared (dot_fma) of vector3.rs:62-108
bool degenerate(const float* v0, const float* v1, const float* v2) {
  float e1[3] = {v0[0] - v2[0], v0[1] - v2[1], v0[2] - v2[2]};
  float e2[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  float c[3] = {fmaf(e1[1], e2[2], -(e1[2] * e2[1])), fmaf(e1[2], e2[0], -(e1[0] * e2[2])),
                fmaf(e1[0], e2[1], -(e1[1] * e2[0]))};
  float n2 = fmaf(c[0], c[0], fmaf(c[1], c[1], c[2] * c[2]));
  return n2 == 0.0f;
}

bool parse_u32(const std::string& s, uint32_t* out) {
  if (s.empty()) return false;
  uint64_t v = 0;
  size_t i = 0;
  if (s[0] == '+') i = 1;
  if (i >= s.size()) return false;
  for (; i < s.size(); i++) {
    if (s[i] < '0' || s[i] > '9') return false;
    v = v * 10 + (uint64_t)(s[i] - '0');
    if (v > 0xffffffffull) return false;
  }
  *out = (uint32_t)v;
  return true;
}

bool parse_f32(const std::string& s, float* out) {
  if (s.empty()) return false;
  char* end = nullptr;
  *out = strtof(s.c_str(), &end);
  return end && *end == '\0';
}

void split_ws(const std::string& line, std::vector<std::string>& out) {
  out.clear();
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && isspace((unsigned char)line[i])) i++;
    if (i >= n) break;
    size_t j = i;
    while (j < n && !isspace((unsigned char)line[j])) j++;
    out.emplace_back(line.substr(i, j - i));
    i = j;
  }
}

}  // namespace

extern "C" {

const char* cv_last_error(void) { return g_last_error.c_str(); }

const char* cv_version(void) { return "convector-mi355x 0.1 (gfx950)"; }

const char* cv_status_string(cv_status s) {
  switch (s) {
    case CV_OK: return "ok";
    case CV_ERR_INVALID_ARGUMENT: return "invalid argument";
    case CV_ERR_IO: return "i/o error";
    case CV_ERR_PARSE: return "parse error";
    case CV_ERR_DEGENERATE_TRIANGLE: return "degenerate triangle";
    case CV_ERR_UNKNOWN_MATERIAL: return "unknown material";
    case CV_ERR_NOT_MULTIPLE_OF_16: return "size not a multiple of 16";
    case CV_ERR_LIGHT_COUNT: return "fewer than 8 direct-sample triangles";
    case CV_ERR_BVH_NO_SPLIT: return "bvh node cannot be split";
    case CV_ERR_NAN: return "NaN in geometry";
    case CV_ERR_DEVICE: return "device error";
    case CV_ERR_OUT_OF_MEMORY: return "out of device memory";
    case CV_ERR_NO_DEVICE: return "no device";
    case CV_ERR_OUT_OF_RANGE: return "out of range";
  }
  return "unknown status";
}

cv_status cv_mesh_load_obj(const char* path, const cv_material_entry* table, size_t n,
                           cv_mesh** out) {
  if (!path || !out || (n && !table)) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(CV_ERR_IO, std::string("failed to open file ") + path);
  std::string text;
  {
    char buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    fclose(f);
  }
  std::unordered_map<std::string, uint32_t> mats;
  for (size_t i = 0; i < n; i++) {
    if (!table[i].name) return fail(CV_ERR_INVALID_ARGUMENT, "null material name");
    mats[table[i].name] = table[i].material;
  }
  std::vector<float> verts;  // 3 per vertex
  std::vector<float> tex;    // 2 per tex coord
  cv_mesh* m = new cv_mesh();
  uint32_t material = 0x00ffffffu;  // SMaterial::white()
  std::vector<std::string> pieces;
  size_t pos = 0;
  uint32_t line_nr = 0;
  struct Idx {
    uint32_t v;
    bool has_t;
    uint32_t t;
  };
  auto parse_index = [&](const std::string& s, Idx* r) -> bool {
    // parse_vertex_index (wavefront.rs:48-55): every '/'-separated part is a
    // u32; the first is the vertex, the second (if any) the tex coord.
    std::vector<std::string> parts;
    size_t a = 0;
    for (;;) {
      size_t b = s.find('/', a);
      parts.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
      if (b == std::string::npos) break;
      a = b + 1;
    }
    uint32_t vals[8];
    for (size_t k = 0; k < parts.size() && k < 8; k++)
      if (!parse_u32(parts[k], &vals[k])) return false;
    if (vals[0] == 0) return false;
    r->v = vals[0] - 1;
    r->has_t = parts.size() >= 2;
    if (r->has_t) {
      if (vals[1] == 0) return false;
      r->t = vals[1] - 1;
    }
    return true;
  };
  auto push_tri = [&](const Idx& i0, const Idx& i1, const Idx& i2) -> cv_status {
    size_t nv = verts.size() / 3;
    if (i0.v >= nv || i1.v >= nv || i2.v >= nv)
      return fail(CV_ERR_PARSE, "vertex index out of range at line " + std::to_string(line_nr));
    const float* v0 = &verts[3 * (size_t)i0.v];
    const float* v1 = &verts[3 * (size_t)i1.v];
    const float* v2 = &verts[3 * (size_t)i2.v];
    if (degenerate(v0, v1, v2))
      return fail(CV_ERR_DEGENERATE_TRIANGLE,
                  "degenerate triangle at line " + std::to_string(line_nr));
    for (int k = 0; k < 3; k++) m->v9.push_back(v0[k]);
    for (int k = 0; k < 3; k++) m->v9.push_back(v1[k]);
    for (int k = 0; k < 3; k++) m->v9.push_back(v2[k]);
    float uv[6] = {0, 0, 0, 0, 0, 0};
    if (i0.has_t && i1.has_t && i2.has_t) {
      size_t nt = tex.size() / 2;
      if (i0.t >= nt || i1.t >= nt || i2.t >= nt)
        return fail(CV_ERR_PARSE, "tex coord index out of range at line " + std::to_string(line_nr));
      uv[0] = tex[2 * i0.t];
      uv[1] = tex[2 * i0.t + 1];
      uv[2] = tex[2 * i1.t];
      uv[3] = tex[2 * i1.t + 1];
      uv[4] = tex[2 * i2.t];
      uv[5] = tex[2 * i2.t + 1];
    }
    for (int k = 0; k < 6; k++) m->uv6.push_back(uv[k]);
    m->mat.push_back(material);
    return CV_OK;
  };
  while (pos <= text.size()) {
    size_t e = text.find('\n', pos);
    std::string line = text.substr(pos, e == std::string::npos ? std::string::npos : e - pos);
    pos = e == std::string::npos ? text.size() + 1 : e + 1;
    line_nr++;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    split_ws(line, pieces);
    if (pieces.empty()) continue;
    const std::string& kw = pieces[0];
    cv_status st = CV_OK;
    if (kw == "v") {
      float c[3];
      if (pieces.size() < 4) st = fail(CV_ERR_PARSE, "missing coordinate at line " + std::to_string(line_nr));
      for (int k = 0; k < 3 && st == CV_OK; k++)
        if (!parse_f32(pieces[1 + k], &c[k])) st = fail(CV_ERR_PARSE, "bad float at line " + std::to_string(line_nr));
      for (size_t k = 4; k < pieces.size() && st == CV_OK; k++) {
        float x;
        if (!parse_f32(pieces[k], &x)) st = fail(CV_ERR_PARSE, "bad float at line " + std::to_string(line_nr));
      }
      if (st == CV_OK) verts.insert(verts.end(), c, c + 3);
    } else if (kw == "vt") {
      float c[2];
      if (pieces.size() < 3) st = fail(CV_ERR_PARSE, "missing coordinate at line " + std::to_string(line_nr));
      for (int k = 0; k < 2 && st == CV_OK; k++)
        if (!parse_f32(pieces[1 + k], &c[k])) st = fail(CV_ERR_PARSE, "bad float at line " + std::to_string(line_nr));
      if (st == CV_OK) tex.insert(tex.end(), c, c + 2);
    } else if (kw == "usemtl") {
      if (pieces.size() < 2) {
        st = fail(CV_ERR_PARSE, "missing material name at line " + std::to_string(line_nr));
      } else {
        auto it = mats.find(pieces[1]);
        if (it == mats.end())
          st = fail(CV_ERR_UNKNOWN_MATERIAL,
                    "material '" + pieces[1] + "' not present in material dictionary");
        else
          material = it->second;
      }
    } else if (kw == "f") {
      if (pieces.size() < 4) {
        st = fail(CV_ERR_PARSE, "missing triangle index at line " + std::to_string(line_nr));
      } else {
        Idx i0, i1, i2;
        if (!parse_index(pieces[1], &i0) || !parse_index(pieces[2], &i1) ||
            !parse_index(pieces[3], &i2)) {
          st = fail(CV_ERR_PARSE, "bad index at line " + std::to_string(line_nr));
        } else {
          st = push_tri(i0, i1, i2);
          // fan triangulation of n-gons (wavefront.rs:133-140)
          for (size_t k = 4; k < pieces.size() && st == CV_OK; k++) {
            Idx i3;
            if (!parse_index(pieces[k], &i3)) {
              st = fail(CV_ERR_PARSE, "bad index at line " + std::to_string(line_nr));
              break;
            }
            st = push_tri(i0, i2, i3);
            i2 = i3;
          }
        }
      }
    }
    if (st != CV_OK) {
      delete m;
      return st;
    }
  }
  *out = m;
  return CV_OK;
}

cv_status cv_mesh_from_triangles(const float* v9, const float* uv6, const uint32_t* materials,
                                 size_t ntris, cv_mesh** out) {
  if (!out || (ntris && (!v9 || !materials))) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  for (size_t i = 0; i < ntris; i++)
    if (degenerate(v9 + 9 * i, v9 + 9 * i + 3, v9 + 9 * i + 6))
      return fail(CV_ERR_DEGENERATE_TRIANGLE, "degenerate triangle " + std::to_string(i));
  cv_mesh* m = new cv_mesh();
  m->v9.assign(v9, v9 + 9 * ntris);
  if (uv6)
    m->uv6.assign(uv6, uv6 + 6 * ntris);
  else
    m->uv6.assign(6 * ntris, 0.0f);
  m->mat.assign(materials, materials + ntris);
  *out = m;
  return CV_OK;
}

size_t cv_mesh_num_triangles(const cv_mesh* mesh) { return mesh ? mesh->ntris() : 0; }

cv_status cv_mesh_copy_triangles(const cv_mesh* mesh, float* v9, float* uv6, uint32_t* materials) {
  if (!mesh) return fail(CV_ERR_INVALID_ARGUMENT, "null mesh");
  if (v9) memcpy(v9, mesh->v9.data(), mesh->v9.size() * 4);
  if (uv6) memcpy(uv6, mesh->uv6.data(), mesh->uv6.size() * 4);
  if (materials) memcpy(materials, mesh->mat.data(), mesh->mat.size() * 4);
  return CV_OK;
}

void cv_mesh_destroy(cv_mesh* mesh) { delete mesh; }

// ------------------------------------------------------------------ scenes

// Stack depth and persistent grid sizes; the spill area holds the entries
// beyond the LDS part of every lane's stack (depth + 2 entries suffice: a pop
// pushes at most two children).
static cv_status scene_apply_tuning(cv_scene* s, const cv_tuning& t) {
  if (t.stack_entries && t.stack_entries != 8 && t.stack_entries != 16 && t.stack_entries != 24 &&
      t.stack_entries != 32)
    return fail(CV_ERR_INVALID_ARGUMENT, "stack_entries must be 8, 16, 24 or 32");
  if (t.extend_waves_per_cu > 64 || t.refill_threshold > 65 || t.fetch_batches > 64)
    return fail(CV_ERR_INVALID_ARGUMENT, "tuning value out of range");
  if (t.stack_entries) s->stack = (int)t.stack_entries;
  if (t.extend_waves_per_cu) s->persistent_blocks = (uint32_t)g_num_cus * t.extend_waves_per_cu;
  if (t.refill_threshold) s->refill = t.refill_threshold;
  if (t.fetch_batches) s->fetch_batches = t.fetch_batches;
  uint32_t need = s->info.depth + 4;
  size_t spill_entries = need > (uint32_t)s->stack ? (size_t)(need - s->stack) : 0;
  HIPTRY(hipStreamSynchronize(s->stream), "sync");
  HIPTRY(s->spill.alloc(spill_entries * s->persistent_blocks * 64 + 1), "alloc spill");
  return CV_OK;
}

static cv_status scene_finish(cv_scene* s, const std::vector<float>& v9,
                              const std::vector<float>& uv6, const std::vector<uint32_t>& mat,
                              const cv_bvh_node* nodes_in, size_t nnodes_in,
                              const uint32_t* order_in) {
  const uint32_t N = s->ntris;
  hipStream_t st = s->stream;
  DevBuf<float> d_v9;
  DevBuf<float> d_uv;
  DevBuf<uint32_t> d_mat;
  HIPTRY(d_v9.alloc(v9.size()), "alloc");
  HIPTRY(hipMemcpyAsync(d_v9.p, v9.data(), v9.size() * 4, hipMemcpyHostToDevice, st), "upload");
  HIPTRY(s->tri_src.alloc(N), "alloc");
  if (nodes_in) {
    HIPTRY(s->nodes.alloc(nnodes_in), "alloc");
    HIPTRY(hipMemcpyAsync(s->nodes.p, nodes_in, nnodes_in * sizeof(DevNode), hipMemcpyHostToDevice, st),
           "upload");
    HIPTRY(hipMemcpyAsync(s->tri_src.p, order_in, N * 4ull, hipMemcpyHostToDevice, st), "upload");
    s->info.num_nodes = (uint32_t)nnodes_in;
    uint32_t depth = 0, max_leaf = 0;
    // depth of an imported tree: walk the pairs on the host copy
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    stack.push_back({0, 1});
    stack.push_back({1, 1});
    uint32_t leaves = 0;
    while (!stack.empty()) {
      auto [i, d] = stack.back();
      stack.pop_back();
      if (i >= nnodes_in) return fail(CV_ERR_INVALID_ARGUMENT, "bvh node index out of range");
      depth = d > depth ? d : depth;
      if (nodes_in[i].len == 0) {
        if (nodes_in[i].index + 1 >= nnodes_in)
          return fail(CV_ERR_INVALID_ARGUMENT, "bvh child index out of range");
        stack.push_back({nodes_in[i].index, d + 1});
        stack.push_back({nodes_in[i].index + 1, d + 1});
      } else {
        if ((uint64_t)nodes_in[i].index + nodes_in[i].len > N)
          return fail(CV_ERR_INVALID_ARGUMENT, "bvh leaf range out of range");
        max_leaf = nodes_in[i].len > max_leaf ? nodes_in[i].len : max_leaf;
        leaves++;
      }
      if (stack.size() > (1u << 24)) return fail(CV_ERR_INVALID_ARGUMENT, "bvh is not a tree");
    }
    s->info.depth = depth;
    s->info.max_leaf = max_leaf;
    s->info.avg_tris_per_leaf = leaves ? (float)N / (float)leaves : 0.0f;
    s->info.avg_area_ratio = 0.0f;
    s->info.build_ms = 0.0f;
  } else {
    HIPTRY(s->nodes.alloc(2ull * N + 2), "alloc");
    cvk::BuildResult br{};
    HIPTRY(cvk::build_reference_bvh(d_v9.p, N, s->nodes.p, s->tri_src.p, &br, st), "bvh build");
    if (br.error) {
      return fail((cv_status)br.error, br.error == CV_ERR_NAN
                                           ? "NaN in triangle coordinates"
                                           : "no valid BVH split (coincident barycenters or a "
                                             "single-triangle scene)");
    }
    s->info.num_nodes = br.num_nodes;
    s->info.depth = br.depth;
    s->info.max_leaf = br.max_leaf;
    s->info.avg_area_ratio = br.avg_area_ratio;
    s->info.avg_tris_per_leaf = br.avg_tris_per_leaf;
    s->info.build_ms = br.build_ms;
  }
  s->info.num_triangles = N;
  HIPTRY(s->tris.alloc(N), "alloc");
  HIPTRY(cvk::launch_prepare_tris(d_v9.p, s->tri_src.p, N, s->tris.p, st), "prepare");
  HIPTRY(d_mat.alloc(N), "alloc");
  HIPTRY(hipMemcpyAsync(d_mat.p, mat.data(), N * 4ull, hipMemcpyHostToDevice, st), "upload");
  HIPTRY(d_uv.alloc(6ull * N), "alloc");
  HIPTRY(hipMemcpyAsync(d_uv.p, uv6.data(), N * 24ull, hipMemcpyHostToDevice, st), "upload");
  HIPTRY(s->tri_mat.alloc(N), "alloc");
  HIPTRY(s->tri_uv.alloc(6ull * N), "alloc");
  hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(N, 256, 65535)), dim3(256), 0, st, d_mat.p,
                     s->tri_src.p, N, 1u, s->tri_mat.p);
  HIPTRY(hipGetLastError(), "gather");
  hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(N, 256, 65535)), dim3(256), 0, st,
                     (const uint32_t*)d_uv.p, s->tri_src.p, N, 6u, (uint32_t*)s->tri_uv.p);
  HIPTRY(hipGetLastError(), "gather");
  HIPTRY(s->direct.alloc(N + 1), "alloc");
  hipLaunchKernelGGL(k_direct_list, dim3(1), dim3(1024), 0, st, s->tri_mat.p, N, s->direct.p,
                     s->direct.p + N);
  HIPTRY(hipGetLastError(), "direct list");
  HIPTRY(hipMemcpyAsync(&s->ndirect, s->direct.p + N, 4, hipMemcpyDeviceToHost, st), "download");
  HIPTRY(hipStreamSynchronize(st), "sync");
  s->info.num_direct = s->ndirect;
  // traversal configuration: packed stack entries need index < 2^27 and
  // leaves of at most 31 triangles
  s->packed = s->info.max_leaf <= 31 && N < (1u << 27) && s->info.num_nodes < (1u << 27);
  HIPTRY(s->cursor.alloc(4), "alloc");
  cv_tuning t{8, 32, 32, 1};
  return scene_apply_tuning(s, t);
}

static cv_status scene_create_impl(const cv_mesh* const* meshes, size_t n,
                                   const cv_scene_opts* opts, const cv_bvh_node* nodes,
                                   size_t nnodes, const uint32_t* order, cv_scene** out) {
  if (!out || (n && !meshes)) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  cv_status ds = ensure_device();
  if (ds != CV_OK) return ds;
  if (opts && opts->builder != CV_BUILDER_REFERENCE)
    return fail(CV_ERR_INVALID_ARGUMENT, "unknown builder");
  std::vector<float> v9, uv6;
  std::vector<uint32_t> mat;
  for (size_t i = 0; i < n; i++) {
    if (!meshes[i]) return fail(CV_ERR_INVALID_ARGUMENT, "null mesh");
    v9.insert(v9.end(), meshes[i]->v9.begin(), meshes[i]->v9.end());
    uv6.insert(uv6.end(), meshes[i]->uv6.begin(), meshes[i]->uv6.end());
    mat.insert(mat.end(), meshes[i]->mat.begin(), meshes[i]->mat.end());
  }
  if (mat.empty()) return fail(CV_ERR_BVH_NO_SPLIT, "scene has no triangles");
  if (mat.size() >= (1ull << 31)) return fail(CV_ERR_INVALID_ARGUMENT, "too many triangles");
  cv_scene* s = new cv_scene();
  s->ntris = (uint32_t)mat.size();
  // Camera::new (scene.rs:33-42)
  s->cam.q[0] = 1.0f;
  s->cam.screen_distance = 1.0f / sinf(3.14159274f / 5.0f);
  if (opts && opts->has_emission_override) {
    s->has_override = true;
    memcpy(s->emission, opts->emission_override, 12);
  }
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "hipStreamCreate");
  }
  cv_status st = scene_finish(s, v9, uv6, mat, nodes, nnodes, order);
  if (st != CV_OK) {
    cv_scene_destroy(s);
    return st;
  }
  *out = s;
  return CV_OK;
}

cv_status cv_scene_create(const cv_mesh* const* meshes, size_t n, const cv_scene_opts* opts,
                          cv_scene** out) {
  return scene_create_impl(meshes, n, opts, nullptr, 0, nullptr, out);
}

cv_status cv_scene_create_with_bvh(const cv_mesh* const* meshes, size_t n,
                                   const cv_scene_opts* opts, const cv_bvh_node* nodes,
                                   size_t num_nodes, const uint32_t* order, cv_scene** out) {
  if (!nodes || !order || num_nodes < 2) return fail(CV_ERR_INVALID_ARGUMENT, "null bvh");
  return scene_create_impl(meshes, n, opts, nodes, num_nodes, order, out);
}

cv_status cv_scene_set_camera(cv_scene* s, const float pos[3], const float dpos[3],
                              const float q[4], const float dq[4], float screen_distance) {
  if (!s || !pos || !dpos || !q || !dq) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  memcpy(s->cam.pos, pos, 12);
  memcpy(s->cam.dpos, dpos, 12);
  memcpy(s->cam.q, q, 16);
  memcpy(s->cam.dq, dq, 16);
  s->cam.screen_distance = screen_distance;
  return CV_OK;
}

cv_status cv_scene_set_tuning(cv_scene* s, const cv_tuning* t) {
  if (!s || !t) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  return scene_apply_tuning(s, *t);
}

cv_status cv_scene_get_tuning(const cv_scene* s, cv_tuning* t) {
  if (!s || !t) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  t->stack_entries = (uint32_t)s->stack;
  t->extend_waves_per_cu = s->persistent_blocks / (uint32_t)g_num_cus;
  t->refill_threshold = s->refill;
  t->fetch_batches = s->fetch_batches;
  return CV_OK;
}

cv_status cv_scene_get_bvh_info(const cv_scene* s, cv_bvh_info* info) {
  if (!s || !info) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *info = s->info;
  return CV_OK;
}

cv_status cv_scene_copy_bvh(const cv_scene* s, cv_bvh_node* nodes, uint32_t* order) {
  if (!s) return fail(CV_ERR_INVALID_ARGUMENT, "null scene");
  if (nodes)
    HIPTRY(hipMemcpy(nodes, s->nodes.p, (size_t)s->info.num_nodes * 32, hipMemcpyDeviceToHost),
           "download");
  if (order)
    HIPTRY(hipMemcpy(order, s->tri_src.p, (size_t)s->ntris * 4, hipMemcpyDeviceToHost), "download");
  return CV_OK;
}

void cv_scene_destroy(cv_scene* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  s->nodes.reset();
  s->tris.reset();
  s->tri_mat.reset();
  s->tri_uv.reset();
  s->tri_src.reset();
  s->direct.reset();
  s->spill.reset();
  s->cursor.reset();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// --------------------------------------------------------------- traversal

cv_status cv_bvh_intersect_nearest(const cv_scene* sc, const cv_ray* rays, size_t n, cv_hit* hits,
                                   cv_isect* isects, uint32_t flags) {
  if (!sc || (n && !rays)) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return CV_OK;
  if (n >= (1ull << 31)) return fail(CV_ERR_INVALID_ARGUMENT, "too many rays");
  cv_scene* s = const_cast<cv_scene*>(sc);
  std::lock_guard<std::mutex> g(s->mu);
  hipStream_t st = s->stream;
  const bool dev = flags & CV_PTR_DEVICE;
  DevBuf<float> d_rays;
  DevBuf<uint32_t> d_hits, d_isects;
  const float* rp = reinterpret_cast<const float*>(rays);
  uint32_t* hp = reinterpret_cast<uint32_t*>(hits);
  uint32_t* ip = reinterpret_cast<uint32_t*>(isects);
  if (!dev) {
    HIPTRY(d_rays.alloc(6 * n), "alloc");
    HIPTRY(hipMemcpyAsync(d_rays.p, rays, 24 * n, hipMemcpyHostToDevice, st), "upload");
    rp = d_rays.p;
    if (hits) {
      HIPTRY(d_hits.alloc(4 * n), "alloc");
      hp = d_hits.p;
    }
    if (isects) {
      HIPTRY(d_isects.alloc(16 * n), "alloc");
      ip = d_isects.p;
    }
  }
  cvk::IntersectArgs a{};
  a.rays = rp;
  a.n = (uint32_t)n;
  a.cursor = s->cursor.p;
  a.nodes = s->nodes.p;
  a.tris = s->tris.p;
  a.tri_src = s->tri_src.p;
  a.tri_mat = s->tri_mat.p;
  a.tri_uv = s->tri_uv.p;
  a.spill = s->spill.p;
  a.hits = hp;
  a.isects = ip;
  HIPTRY(cvk::launch_zero_u32(s->cursor.p, 1, st), "zero");
  HIPTRY(cvk::launch_intersect(a, isects != nullptr, s->packed, s->stack, s->persistent_blocks, st),
         "intersect");
  if (!dev) {
    if (hits) HIPTRY(hipMemcpyAsync(hits, hp, 16 * n, hipMemcpyDeviceToHost, st), "download");
    if (isects) HIPTRY(hipMemcpyAsync(isects, ip, 64 * n, hipMemcpyDeviceToHost, st), "download");
  }
  HIPTRY(hipStreamSynchronize(st), "sync");
  return CV_OK;
}

// ---------------------------------------------------------------- renderer

cv_status cv_renderer_create(cv_scene* s, uint32_t width, uint32_t height,
                             const cv_render_opts* opts, cv_renderer** out) {
  if (!s || !out) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  // The 16x4-block HDR layout needs width % 16 == 0 and height % 4 == 0
  // (renderer.rs:366-374, 384-385); the reference's RenderBuffer also
  // rounds heights to 16, which the per-pixel GPU path does not need.
  if ((width & 15) || (height & 3) || width == 0 || height == 0)
    return fail(CV_ERR_NOT_MULTIPLE_OF_16, "width must be a multiple of 16 and height of 4");
  if (s->ndirect < 8)
    return fail(CV_ERR_LIGHT_COUNT, "the scene has " + std::to_string(s->ndirect) +
                                        " direct-sample triangles; direct sampling needs 8");
  cv_render_opts o{CV_RNG_PIXEL, CV_APPROX_PORTABLE, 5};
  if (opts) o = *opts;
  if (o.max_bounces == 0) o.max_bounces = 5;
  if (o.max_bounces >= (uint32_t)cvk::kMaxBounceSlots)
    return fail(CV_ERR_INVALID_ARGUMENT, "max_bounces too large");
  if (o.rng_mode != CV_RNG_PIXEL)
    return fail(CV_ERR_INVALID_ARGUMENT, "the GPU renderer implements CV_RNG_PIXEL");
  if (o.approx_mode != CV_APPROX_PORTABLE)
    return fail(CV_ERR_INVALID_ARGUMENT, "the GPU renderer implements CV_APPROX_PORTABLE");
  cv_renderer* r = new cv_renderer();
  r->scene = s;
  r->width = width;
  r->height = height;
  r->opts = o;
  auto bail = [&](cv_status st) {
    cv_renderer_destroy(r);
    return st;
  };
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
  const size_t npix = (size_t)width * height;
  for (int k = 0; k < 2; k++) {
    if ((e = r->r0[k].alloc(npix)) || (e = r->r1[k].alloc(npix)) || (e = r->col[k].alloc(npix)) ||
        (e = r->rng[k].alloc(npix)))
      return bail(hip_fail(e, "alloc queues"));
  }
  if ((e = r->hits.alloc(npix)) || (e = r->ctrl.alloc(1)) || (e = r->lights.alloc(s->ndirect)) ||
      (e = r->hdr.alloc(npix * 3)) || (e = r->bitmap.alloc(npix)) || (e = r->gbuffer.alloc(npix)))
    return bail(hip_fail(e, "alloc buffers"));
  if ((e = cvk::launch_prepare_lights(s->tris.p, s->direct.p, s->ndirect, r->lights.p, r->stream)))
    return bail(hip_fail(e, "prepare lights"));
  if ((e = hipStreamSynchronize(r->stream))) return bail(hip_fail(e, "sync"));
  *out = r;
  return CV_OK;
}

void cv_renderer_destroy(cv_renderer* r) {
  if (!r) return;
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  for (hipEvent_t ev : r->events) (void)hipEventDestroy(ev);
  for (int k = 0; k < 2; k++) {
    r->r0[k].reset();
    r->r1[k].reset();
    r->col[k].reset();
    r->rng[k].reset();
  }
  r->hits.reset();
  r->ctrl.reset();
  r->lights.reset();
  r->hdr.reset();
  r->bitmap.reset();
  r->gbuffer.reset();
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

namespace {

struct Rect {
  uint32_t x, y, w, h;
};

// One frame over a rectangle: generate, then extend + shade per bounce.
cv_status render_rect(cv_renderer* r, const Rect& rc, uint32_t frame, float* hdr_dev,
                      uint32_t* bitmap_dev, uint32_t* gbuffer_dev, uint32_t flags,
                      std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>>* evs) {
  cv_scene* s = r->scene;
  hipStream_t st = r->stream;
  auto ev_pair = [&](int kind) -> std::pair<hipEvent_t, hipEvent_t>* {
    if (!evs) return nullptr;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return nullptr;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return nullptr;
    }
    evs->push_back({kind, {a, b}});
    return &evs->back().second;
  };
  cvk::GenerateArgs g{};
  g.ctrl = r->ctrl.p;
  g.q = cvk::Queue{r->r0[0].p, r->r1[0].p, r->col[0].p, r->rng[0].p};
  g.width = r->width;
  g.height = r->height;
  g.rect_x = rc.x;
  g.rect_y = rc.y;
  g.rect_w = rc.w;
  g.rect_h = rc.h;
  g.frame = frame;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    g.cam = s->cam;
  }
  g.gbuffer = gbuffer_dev;
  g.bitmap = bitmap_dev;
  const size_t npix = (size_t)rc.w * rc.h;
  auto* ep = ev_pair(0);
  if (ep) (void)hipEventRecord(ep->first, st);
  HIPTRY(cvk::launch_generate(g, grid_for(npix, 256, 8192), st), "generate");
  if (ep) (void)hipEventRecord(ep->second, st);
  const bool count = flags & CV_RENDER_COUNT;
  for (uint32_t b = 0; b < r->opts.max_bounces; b++) {
    const int cur = b & 1, nxt = cur ^ 1;
    cvk::ExtendArgs x{};
    x.ctrl = r->ctrl.p;
    x.bounce = b;
    x.r0 = r->r0[cur].p;
    x.r1 = r->r1[cur].p;
    x.hits = r->hits.p;
    x.nodes = s->nodes.p;
    x.tris = s->tris.p;
    x.spill = s->spill.p;
    x.batches = s->fetch_batches;
    x.refill = s->refill <= 64 ? s->refill : 0u;
    ep = ev_pair(1);
    if (ep) (void)hipEventRecord(ep->first, st);
    HIPTRY(cvk::launch_extend(x, count, s->packed, s->stack, s->persistent_blocks, st), "extend");
    if (ep) (void)hipEventRecord(ep->second, st);
    cvk::ShadeArgs sh{};
    sh.ctrl = r->ctrl.p;
    sh.bounce = b;
    sh.max_bounces = r->opts.max_bounces;
    sh.width = r->width;
    sh.in = cvk::Queue{r->r0[cur].p, r->r1[cur].p, r->col[cur].p, r->rng[cur].p};
    sh.out = cvk::Queue{r->r0[nxt].p, r->r1[nxt].p, r->col[nxt].p, r->rng[nxt].p};
    sh.hits = r->hits.p;
    sh.tris = s->tris.p;
    sh.tri_mat = s->tri_mat.p;
    sh.tri_uv = s->tri_uv.p;
    sh.lights = r->lights.p;
    sh.nlights = s->ndirect;
    sh.has_override = s->has_override ? 1u : 0u;
    memcpy(sh.emission, s->emission, 12);
    sh.hdr = hdr_dev;
    sh.bitmap = bitmap_dev;
    sh.gbuffer = gbuffer_dev;
    ep = ev_pair(2);
    if (ep) (void)hipEventRecord(ep->first, st);
    HIPTRY(cvk::launch_shade(sh, (uint32_t)npix, st), "shade");
    if (ep) (void)hipEventRecord(ep->second, st);
  }
  return CV_OK;
}

cv_status run_frames(cv_renderer* r, const Rect& rc, uint32_t first, uint32_t nframes, float* hdr,
                     uint32_t* bitmap, uint32_t* gbuffer, uint32_t flags, bool hdr_mode) {
  hipStream_t st = r->stream;
  const bool dev = flags & CV_PTR_DEVICE;
  const uint32_t W = r->width;
  // HDR block region of the rectangle (rc.x, rc.w multiples of 16; rc.y, rc.h of 4)
  const size_t blk_row_bytes = (size_t)(W / 16) * 768;
  const size_t rect_blk_off = (size_t)(rc.y / 4) * blk_row_bytes + (size_t)(rc.x / 16) * 768;
  const size_t rect_blk_w = (size_t)(rc.w / 16) * 768;
  float* hdr_dev = nullptr;
  if (hdr_mode) {
    hdr_dev = dev ? hdr : r->hdr.p;
    if (!dev)
      HIPTRY(hipMemcpy2DAsync((char*)r->hdr.p + rect_blk_off, blk_row_bytes,
                              (const char*)hdr + rect_blk_off, blk_row_bytes, rect_blk_w, rc.h / 4,
                              hipMemcpyHostToDevice, st),
             "upload hdr");
  }
  uint32_t* bm_dev = nullptr;
  if (bitmap) bm_dev = dev ? bitmap : r->bitmap.p;
  uint32_t* gb_dev = nullptr;
  if (gbuffer) gb_dev = dev ? gbuffer : r->gbuffer.p;
  HIPTRY(cvk::launch_zero_u32(reinterpret_cast<uint32_t*>(&r->ctrl.p->sum_aabb), 6, st), "zero");
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;
  const bool timing = flags & CV_RENDER_TIMING;
  hipEvent_t t0, t1;
  HIPTRY(hipEventCreate(&t0), "event");
  HIPTRY(hipEventCreate(&t1), "event");
  HIPTRY(hipEventRecord(t0, st), "event");
  cv_status status = CV_OK;
  for (uint32_t f = 0; f < nframes && status == CV_OK; f++)
    status = render_rect(r, rc, first + f, hdr_dev, bm_dev, gb_dev, flags, timing ? &evs : nullptr);
  (void)hipEventRecord(t1, st);
  if (status == CV_OK && !dev) {
    const size_t row = (size_t)W * 4;
    const size_t off = ((size_t)rc.y * W + rc.x) * 4;
    if (hdr_mode)
      status = hipMemcpy2DAsync((char*)hdr + rect_blk_off, blk_row_bytes,
                                (const char*)r->hdr.p + rect_blk_off, blk_row_bytes, rect_blk_w,
                                rc.h / 4, hipMemcpyDeviceToHost, st) == hipSuccess
                   ? CV_OK
                   : fail(CV_ERR_DEVICE, "download hdr");
    if (status == CV_OK && bitmap)
      status = hipMemcpy2DAsync((char*)bitmap + off, row, (const char*)r->bitmap.p + off, row,
                                (size_t)rc.w * 4, rc.h, hipMemcpyDeviceToHost, st) == hipSuccess
                   ? CV_OK
                   : fail(CV_ERR_DEVICE, "download bitmap");
    if (status == CV_OK && gbuffer)
      status = hipMemcpy2DAsync((char*)gbuffer + off, row, (const char*)r->gbuffer.p + off, row,
                                (size_t)rc.w * 4, rc.h, hipMemcpyDeviceToHost, st) == hipSuccess
                   ? CV_OK
                   : fail(CV_ERR_DEVICE, "download gbuffer");
  }
  hipError_t se = hipStreamSynchronize(st);
  if (status == CV_OK && se != hipSuccess) status = hip_fail(se, "render");
  cv_render_stats stt{};
  stt.paths = (uint64_t)rc.w * rc.h * nframes;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  stt.total_ms = ms;
  for (auto& e : evs) {
    float m = 0.0f;
    (void)hipEventElapsedTime(&m, e.second.first, e.second.second);
    if (e.first == 0) stt.generate_ms += m;
    if (e.first == 1) {
      stt.extend_ms += m;
      stt.extend_launches++;
    }
    if (e.first == 2) stt.shade_ms += m;
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (!timing) stt.extend_launches = (uint64_t)nframes * r->opts.max_bounces;
  if (status == CV_OK) {
    cvk::Ctrl hc;
    if (hipMemcpy(&hc, r->ctrl.p, sizeof hc, hipMemcpyDeviceToHost) == hipSuccess) {
      stt.sum_numi_aabb = hc.sum_aabb;
      stt.sum_numi_tri = hc.sum_tri;
      stt.extend_rays = hc.total_rays;
    }
  }
  r->stats = stt;
  return status;
}

}  // namespace

static cv_status patch_args(cv_renderer* r, uint32_t pw, uint32_t x, uint32_t y) {
  if (!r) return fail(CV_ERR_INVALID_ARGUMENT, "null renderer");
  if ((pw & 15) || pw == 0) return fail(CV_ERR_NOT_MULTIPLE_OF_16, "patch width must be a multiple of 16");
  if ((x & 15) || (y & 15)) return fail(CV_ERR_NOT_MULTIPLE_OF_16, "patch origin must be a multiple of 16");
  if (x + pw > r->width || y + pw > r->height) return fail(CV_ERR_OUT_OF_RANGE, "patch outside the frame");
  return CV_OK;
}

cv_status cv_renderer_render_patch_u8(cv_renderer* r, uint32_t* bitmap, uint32_t* gbuffer,
                                      uint32_t patch_width, uint32_t x, uint32_t y,
                                      uint32_t frame, uint32_t flags) {
  cv_status st = patch_args(r, patch_width, x, y);
  if (st != CV_OK) return st;
  if (!bitmap) return fail(CV_ERR_INVALID_ARGUMENT, "null bitmap");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{x, y, patch_width, patch_width}, frame, 1, nullptr, bitmap, gbuffer,
                    flags, false);
}

cv_status cv_renderer_accumulate_patch_f32(cv_renderer* r, float* hdr, uint32_t* gbuffer,
                                           uint32_t patch_width, uint32_t x, uint32_t y,
                                           uint32_t frame, uint32_t flags) {
  cv_status st = patch_args(r, patch_width, x, y);
  if (st != CV_OK) return st;
  if (!hdr) return fail(CV_ERR_INVALID_ARGUMENT, "null hdr");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{x, y, patch_width, patch_width}, frame, 1, hdr, nullptr, gbuffer,
                    flags, true);
}

cv_status cv_renderer_render_frame_u8(cv_renderer* r, uint32_t* bitmap, uint32_t* gbuffer,
                                      uint32_t frame, uint32_t flags) {
  if (!r || !bitmap) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{0, 0, r->width, r->height}, frame, 1, nullptr, bitmap, gbuffer, flags,
                    false);
}

cv_status cv_renderer_accumulate_frames_f32(cv_renderer* r, float* hdr, uint32_t* gbuffer,
                                            uint32_t first_frame, uint32_t n_frames,
                                            uint32_t flags) {
  if (!r || !hdr) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{0, 0, r->width, r->height}, first_frame, n_frames, hdr, nullptr,
                    gbuffer, flags, true);
}

cv_status cv_renderer_buffer_f32_into_bitmap(cv_renderer* r, const float* hdr, uint32_t* bitmap,
                                             uint32_t num_samples, uint32_t flags) {
  if (!r || !hdr || !bitmap) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(r->mu);
  hipStream_t st = r->stream;
  const bool dev = flags & CV_PTR_DEVICE;
  const size_t npix = (size_t)r->width * r->height;
  const float* hp = hdr;
  uint32_t* bp = bitmap;
  if (!dev) {
    HIPTRY(hipMemcpyAsync(r->hdr.p, hdr, npix * 12, hipMemcpyHostToDevice, st), "upload");
    hp = r->hdr.p;
    bp = r->bitmap.p;
  }
  HIPTRY(cvk::launch_hdr_to_bitmap(hp, bp, r->width, r->height, num_samples, st), "convert");
  if (!dev) HIPTRY(hipMemcpyAsync(bitmap, bp, npix * 4, hipMemcpyDeviceToHost, st), "download");
  HIPTRY(hipStreamSynchronize(st), "sync");
  return CV_OK;
}

cv_status cv_renderer_get_stats(const cv_renderer* r, cv_render_stats* stats) {
  if (!r || !stats) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *stats = r->stats;
  return CV_OK;
}

cv_status cv_device_synchronize(void) {
  cv_status ds = ensure_device();
  if (ds != CV_OK) return ds;
  HIPTRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return CV_OK;
}

}  // extern "C"
