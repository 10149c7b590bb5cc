// api.hip -- the C ABI of include/convector.h for meshes, scenes (GPU BVH
// build + device upload) and batched nearest-hit queries (the renderer and
// batches are in render.hip).  Every compute step runs in the kernels of
// kernels.hip and build.hip; this file only parses input, owns device memory
// and orders launches on one HIP stream per object.
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
#include "host.hpp"
#include "internal.hpp"

using cvk::DevLight;
using cvk::DevNode;
using cvk::DevTri;
using cvh::DevBuf;
using cvh::fail;
using cvh::grid_for;
using cvh::hip_fail;

namespace cvh {

namespace {
thread_local std::string g_last_error;
uint32_t g_num_cus = 0;
}  // namespace

cv_status fail(cv_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

cv_status hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(e == hipErrorOutOfMemory ? CV_ERR_OUT_OF_MEMORY : CV_ERR_DEVICE, m);
}

cv_status ensure_device() {
  if (g_num_cus > 0) return CV_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(CV_ERR_NO_DEVICE, "no HIP device available");
  int dev = 0;
  HIPTRY(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  HIPTRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  g_num_cus = prop.multiProcessorCount > 0 ? (uint32_t)prop.multiProcessorCount : 256u;
  return CV_OK;
}

uint32_t num_cus() { return g_num_cus ? g_num_cus : 256u; }

const char* last_error() { return g_last_error.c_str(); }

namespace {
// One pinned bounce buffer for the whole process, allocated on first use and
// kept until the process ends (freeing pinned memory at exit can run after
// the HIP runtime is torn down).
constexpr size_t kBounceBytes = 16u << 20;
std::mutex g_bounce_mu;
char* g_bounce = nullptr;
}  // namespace

cv_status copy_caller_memory(void* dst, const void* src, size_t bytes, bool to_device,
                             hipStream_t st) {
  if (bytes == 0) return CV_OK;
  std::lock_guard<std::mutex> lk(g_bounce_mu);
  if (!g_bounce) HIPTRY(hipHostMalloc((void**)&g_bounce, kBounceBytes, hipHostMallocDefault),
                        "pinned alloc");
  for (size_t off = 0; off < bytes; off += kBounceBytes) {
    const size_t c = bytes - off < kBounceBytes ? bytes - off : kBounceBytes;
    if (to_device) {
      memcpy(g_bounce, (const char*)src + off, c);
      HIPTRY(hipMemcpyAsync((char*)dst + off, g_bounce, c, hipMemcpyHostToDevice, st), "upload");
      HIPTRY(hipStreamSynchronize(st), "upload");
    } else {
      HIPTRY(hipMemcpyAsync(g_bounce, (const char*)src + off, c, hipMemcpyDeviceToHost, st),
             "download");
      HIPTRY(hipStreamSynchronize(st), "download");
      memcpy((char*)dst + off, g_bounce, c);
    }
  }
  return CV_OK;
}

cv_status validate_tuning(const cv_tuning& t) {
  if (t.stack_entries && t.stack_entries != 8 && t.stack_entries != 16 && t.stack_entries != 24 &&
      t.stack_entries != 32)
    return fail(CV_ERR_INVALID_ARGUMENT, "stack_entries must be 8, 16, 24 or 32");
  if (t.extend_waves_per_cu > 64 || t.refill_threshold > 65 || t.fetch_batches > 64 ||
      (t.block_waves && t.block_waves != 1 && t.block_waves != 2 && t.block_waves != 4) ||
      (t.frames_per_step > 64 && t.frames_per_step != CV_FRAMES_AUTO) || t.refill_interval > 64 ||
      t.leaf_batch > 65)
    return fail(CV_ERR_INVALID_ARGUMENT, "tuning value out of range");
  if (t.bvh_width && t.bvh_width != 2 && t.bvh_width != 4 && t.bvh_width != 8)
    return fail(CV_ERR_INVALID_ARGUMENT, "bvh_width must be 2, 4 or 8");
  if (t.ray_sort_bits && t.ray_sort_bits != CV_RAY_SORT_NONE &&
      (t.ray_sort_bits < 3 || t.ray_sort_bits > 24 || t.ray_sort_bits % 3))
    return fail(CV_ERR_INVALID_ARGUMENT,
                "ray_sort_bits must be a multiple of 3 from 3 to 24, or CV_RAY_SORT_NONE");
  if (t.queue_parts && t.queue_parts != 1 && t.queue_parts != cvk::kParts)
    return fail(CV_ERR_INVALID_ARGUMENT, "queue_parts must be 1 or 8");
  return CV_OK;
}

void apply_tuning(LaunchCfg& c, const cv_tuning& t) {
  if (t.stack_entries) c.stack = (int)t.stack_entries;
  if (t.extend_waves_per_cu) c.blocks = num_cus() * t.extend_waves_per_cu;
  if (t.refill_threshold) c.refill = t.refill_threshold;
  if (t.fetch_batches) c.fetch_batches = t.fetch_batches;
  if (t.block_waves) c.block_waves = t.block_waves;
  if (t.frames_per_step) c.frames_per_step = t.frames_per_step;
  if (t.refill_interval) c.refill_interval = t.refill_interval;
  if (t.leaf_batch) c.leaf_batch = t.leaf_batch;
  if (t.bvh_width) c.width = t.bvh_width;
  if (t.ray_sort_bits) c.sort_bits = t.ray_sort_bits == CV_RAY_SORT_NONE ? 0u : t.ray_sort_bits;
  if (t.queue_parts) c.parts = t.queue_parts;
}

cv_tuning tuning_of(const LaunchCfg& c) {
  cv_tuning t{};
  t.stack_entries = (uint32_t)c.stack;
  t.extend_waves_per_cu = c.blocks / num_cus();
  t.refill_threshold = c.refill;
  t.fetch_batches = c.fetch_batches;
  t.block_waves = c.block_waves;
  t.frames_per_step = c.frames_per_step;
  t.refill_interval = c.refill_interval;
  t.leaf_batch = c.leaf_batch;
  t.bvh_width = c.width;
  t.ray_sort_bits = c.sort_bits ? c.sort_bits : CV_RAY_SORT_NONE;
  t.queue_parts = c.parts;
  return t;
}

cvk::JobDesc job_for_scene(cv_scene* s, uint32_t width) {
  cvk::JobDesc j{};
  j.nodes = traversal_nodes(s, width);
  j.tris = s->tris.p;
  j.tri_mat = s->tri_mat.p;
  j.tri_uv = s->tri_uv.p;
  j.lights = s->lights.p;
  j.nlights = s->ndirect;
  j.has_override = s->has_override ? 1u : 0u;
  memcpy(j.emission, s->emission, 12);
  for (int k = 0; k < 3; k++) {
    j.lo[k] = s->lo[k];
    j.inv_ext[k] = 1.0f / std::max(s->hi[k] - s->lo[k], 1e-30f);
  }
  std::lock_guard<std::mutex> g(s->mu);
  j.cam = s->cam;
  return j;
}

}  // namespace cvh

namespace {

// ----------------------------------------------------------- kernels local

__global__ void k_gather_u32(const uint32_t* src, const uint32_t* order, uint32_t n,
                             uint32_t stride, uint32_t* dst) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    for (uint32_t j = 0; j < stride; j++) dst[(size_t)k * stride + j] = src[(size_t)order[k] * stride + j];
}

// Scene::from_meshes's direct_sample list (scene.rs:114-129): indices in BVH
// order whose material has bit 30 set, ascending.  One workgroup, in order,
// 16 consecutive triangles per thread and tile (a prefix sum of the
// threads' counts places each thread's indices).
__global__ __launch_bounds__(1024) void k_direct_list(const uint32_t* mat, uint32_t n,
                                                      uint32_t* out, uint32_t* count) {
  constexpr uint32_t E = 16;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t t0 = 0; t0 < n; t0 += 1024 * E) {
    const uint32_t k0 = t0 + threadIdx.x * E;
    uint32_t bits = 0;
    if (k0 + E <= n) {
#pragma unroll
      for (uint32_t q = 0; q < E / 4; q++) {
        const uint4 m = reinterpret_cast<const uint4*>(mat + k0)[q];
        const uint32_t mm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
        for (uint32_t e = 0; e < 4; e++)
          bits |= ((mm[e] & 0x40000000u) ? 1u : 0u) << (4 * q + e);
      }
    } else {
      for (uint32_t e = 0; e < E; e++)
        if (k0 + e < n && (mat[k0 + e] & 0x40000000u)) bits |= 1u << e;
    }
    const uint32_t c = __popc(bits);
    uint32_t inc = c;
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
    uint32_t pos = base + prior + inc - c;
    for (uint32_t e = 0; e < E; e++)
      if ((bits >> e) & 1u) out[pos++] = k0 + e;
    __syncthreads();
    if (threadIdx.x == 0) base += tile;
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = base;
}

}  // namespace

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

const char* cv_last_error(void) { return cvh::last_error(); }

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
  cv_status v = cvh::validate_tuning(t);
  if (v != CV_OK) return v;
  if (t.bvh_width > 2 && (v = cvh::ensure_wide_locked(s, t.bvh_width)) != CV_OK) return v;
  cvh::apply_tuning(s->cfg, t);
  uint32_t need = cvh::stack_depth(s, s->cfg.width) + 4;
  size_t spill_entries = need > (uint32_t)s->cfg.stack ? (size_t)(need - s->cfg.stack) : 0;
  HIPTRY(hipStreamSynchronize(s->stream), "sync");
  // lanes of a launch: the waves rounded up to whole 4-wave workgroups
  HIPTRY(s->spill.alloc(spill_entries * (((size_t)s->cfg.blocks + 3) / 4 * 4) * 64 + 1),
         "alloc spill");
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
    HIPTRY(cvk::build_reference_bvh(d_v9.p, N, s->nodes.p, s->tri_src.p, &br, st, s->info.bins),
           "bvh build");
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
  HIPTRY(s->lights.alloc(s->ndirect), "alloc");
  HIPTRY(cvk::launch_prepare_lights(s->tris.p, s->direct.p, s->ndirect, s->lights.p, st),
         "prepare lights");
  HIPTRY(hipStreamSynchronize(st), "sync");
  // the scene's bounds: the union of the root's two children
  {
    cv_bvh_node root[2];
    HIPTRY(hipMemcpy(root, s->nodes.p, sizeof(root), hipMemcpyDeviceToHost), "download");
    for (int k = 0; k < 3; k++) {
      s->lo[k] = std::min(root[0].min[k], root[1].min[k]);
      s->hi[k] = std::max(root[0].max[k], root[1].max[k]);
    }
  }
  // traversal configuration: packed stack entries need index < 2^27 and
  // leaves of at most 31 triangles
  s->packed = s->info.max_leaf <= 31 && N < (1u << 27) && s->info.num_nodes < (1u << 27);
  HIPTRY(s->cursor.alloc(cvk::kIsectGroups * cvk::kIsectCursorStride), "alloc");
  cv_tuning t{8, 32, 16, 1, 1};
  return scene_apply_tuning(s, t);
}

static cv_status scene_create_impl(const cv_mesh* const* meshes, size_t n,
                                   const cv_scene_opts* opts, const cv_bvh_node* nodes,
                                   size_t nnodes, const uint32_t* order, cv_scene** out) {
  if (!out || (n && !meshes)) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  cv_status ds = cvh::ensure_device();
  if (ds != CV_OK) return ds;
  if (opts && opts->builder != CV_BUILDER_REFERENCE && opts->builder != CV_BUILDER_BINS_32 &&
      opts->builder != CV_BUILDER_BINS_16)
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
  s->info.bins = nodes ? 0u : (opts && opts->builder ? opts->builder : 64u);
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
  *t = cvh::tuning_of(s->cfg);
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
  s->lights.reset();
  s->spill.reset();
  s->cursor.reset();
  s->q_ctrl.reset();
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
  if (hits && !isects) {
    // Hits only: the renderer's lane-refill traversal (idle lanes take new
    // rays, lanes at leaves wait for a common leaf round) as a one-bounce
    // step over the caller's rays, writing source indices; each lane's
    // operations are those of the single-ray traversal, so the hits are
    // the same bits.
    if (!s->q_ctrl.p) HIPTRY(s->q_ctrl.alloc(1), "alloc");
    HIPTRY(cvk::launch_query_ctrl(s->q_ctrl.p, (uint32_t)n, st), "query control");
    cvk::ExtendArgs x{};
    x.ctrl = s->q_ctrl.p;
    x.bounce = 0;
    x.rays6 = rp;
    x.tri_src = s->tri_src.p;
    x.hits = reinterpret_cast<uint4*>(hp);
    cv_status ws = cvh::ensure_wide_locked(s, s->cfg.width);
    if (ws != CV_OK) return ws;
    x.nodes = cvh::traversal_nodes(s, s->cfg.width);
    x.tris = s->tris.p;
    x.jobs = nullptr;
    x.spill = s->spill.p;
    // At least 4 batches of 64 rays per cursor fetch: queries end after a
    // few dozen box tests, so one fetch per 64 rays serialises on the
    // cursor (3.4 ms for 2^24 rays whatever the scene; 2.0 ms with 4 on
    // 968 triangles).  Rendering keeps the tuned value (1 is fastest there).
    x.batches = s->cfg.fetch_batches > 4 ? s->cfg.fetch_batches : 4u;
    x.refill = s->cfg.refill <= 64 ? s->cfg.refill : 0u;
    // Refill checks 4 times as often as the renderer's: a query ends after
    // a few dozen box tests, so a finished lane would otherwise idle for a
    // large share of a ray's steps (2^24 rays on suzanne: 1.79 -> 1.65 ms
    // with 4 steps between checks instead of 16; the bunny's within 1 %).
    x.interval = s->cfg.refill_interval > 4 ? s->cfg.refill_interval / 4 : 1u;
    x.leaf_batch = s->cfg.leaf_batch <= 64 ? s->cfg.leaf_batch : 0u;
    x.parts = s->cfg.parts;
    HIPTRY(cvk::launch_extend(x, false, s->packed, false, s->cfg.stack, s->cfg.blocks,
                              s->cfg.block_waves, st, s->cfg.width),
           "intersect");
    if (!dev) HIPTRY(hipMemcpyAsync(hits, hp, 16 * n, hipMemcpyDeviceToHost, st), "download");
    HIPTRY(hipStreamSynchronize(st), "sync");
    return CV_OK;
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
  HIPTRY(cvk::launch_zero_u32(s->cursor.p, cvk::kIsectGroups * cvk::kIsectCursorStride, st),
         "zero");
  HIPTRY(cvk::launch_intersect(a, isects != nullptr, s->packed, s->cfg.stack, s->cfg.blocks, st),
         "intersect");
  if (!dev) {
    if (hits) HIPTRY(hipMemcpyAsync(hits, hp, 16 * n, hipMemcpyDeviceToHost, st), "download");
    if (isects) HIPTRY(hipMemcpyAsync(isects, ip, 64 * n, hipMemcpyDeviceToHost, st), "download");
  }
  HIPTRY(hipStreamSynchronize(st), "sync");
  return CV_OK;
}

}  // extern "C"

// ------------------------------------------------ allocation-free queries

namespace {
// k_intersect with a fixed launch shape: 2048 persistent 64-lane workgroups
// (8 per CU), 32 LDS stack entries per lane, and the deeper entries of
// trees up to CV_NOALLOC_MAX_DEPTH levels in the spill area.
constexpr uint32_t kNoAllocWaves = 2048;
constexpr int kNoAllocStack = 32;
constexpr size_t kNoAllocSpillPerLane = CV_NOALLOC_MAX_DEPTH + 4 - kNoAllocStack;
constexpr size_t kNoAllocCursorBytes =
    (size_t)cvk::kIsectGroups * cvk::kIsectCursorStride * sizeof(uint32_t);
constexpr size_t kNoAllocBytes =
    kNoAllocCursorBytes + (size_t)kNoAllocWaves * 64 * kNoAllocSpillPerLane * sizeof(uint2);
static_assert(kNoAllocBytes == CV_NOALLOC_SCRATCH_BYTES, "scratch size and header disagree");
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
// per device: the device address of `scratch` (each GPU holds its own copy)
void* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

// The queries' scratch: work cursors, then the stack spill area.  In host
// code `scratch` names a host-side stand-in; kernels get the device address
// that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

extern "C" {

cv_status cv_bvh_intersect_nearest_noalloc(const cv_scene* sc, const cv_ray* rays, size_t n,
                                           cv_hit* hits, cv_isect* isects) {
  if (!sc || (n && !rays)) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return CV_OK;
  if (n >= (1ull << 31)) return fail(CV_ERR_INVALID_ARGUMENT, "too many rays");
  if (sc->info.depth > CV_NOALLOC_MAX_DEPTH)
    return fail(CV_ERR_OUT_OF_RANGE,
                "BVH depth " + std::to_string(sc->info.depth) + " exceeds the " +
                    std::to_string(CV_NOALLOC_MAX_DEPTH) +
                    " levels the fixed scratch of the allocation-free query holds");
  int dev = 0;
  HIPTRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(CV_ERR_INVALID_ARGUMENT, "device index");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    // once per device, with that device current
    void* p = nullptr;
    HIPTRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
    g_noalloc_ptr[dev] = p;
  }
  char* base = static_cast<char*>(g_noalloc_ptr[dev]);
  cv_scene* s = const_cast<cv_scene*>(sc);
  std::lock_guard<std::mutex> g(s->mu);
  hipStream_t st = s->stream;
  cvk::IntersectArgs a{};
  a.rays = reinterpret_cast<const float*>(rays);
  a.n = (uint32_t)n;
  a.cursor = reinterpret_cast<uint32_t*>(base);
  a.nodes = s->nodes.p;
  a.tris = s->tris.p;
  a.tri_src = s->tri_src.p;
  a.tri_mat = s->tri_mat.p;
  a.tri_uv = s->tri_uv.p;
  a.spill = reinterpret_cast<uint2*>(base + kNoAllocCursorBytes);
  a.hits = reinterpret_cast<uint32_t*>(hits);
  a.isects = reinterpret_cast<uint32_t*>(isects);
  // the cursors start at zero for every call (a fill kernel, no memset node)
  HIPTRY(cvk::launch_zero_u32(a.cursor, cvk::kIsectGroups * cvk::kIsectCursorStride, st),
         "zero");
  HIPTRY(cvk::launch_intersect(a, isects != nullptr, s->packed, kNoAllocStack, kNoAllocWaves,
                               st),
         "intersect");
  HIPTRY(hipStreamSynchronize(st), "sync");
  return CV_OK;
}

cv_status cv_device_synchronize(void) {
  cv_status ds = cvh::ensure_device();
  if (ds != CV_OK) return ds;
  HIPTRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return CV_OK;
}

}  // extern "C"
