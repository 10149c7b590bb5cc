// render.hip -- the wavefront pipeline and the two C-ABI objects built on it:
// cv_renderer (the reference's Renderer: one job per step, a frame or a
// patch) and cv_batch (many jobs of different scenes, sizes and sample
// counts sharing every launch).
//
// A step renders one frame of each of its jobs: generate, then per bounce
// extend + shade over queues that hold the paths of all jobs.  The step's
// job table and path count are copied to the device before its launches, so
// the launch sequence of a step is the same for every step and can be
// captured once as a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/convector.h"
#include "host.hpp"
#include "internal.hpp"

using cvh::DevBuf;
using cvh::fail;
using cvh::HostBuf;
using cvh::hip_fail;

namespace {

constexpr uint32_t kRingSlots = 8;

// After a step: hdr += slots[0] + ... + slots[count-1] over the rectangle
// (further frames of one image rendered in the same launches, each into its
// own buffer, added in frame order).
struct FoldDesc {
  float* hdr = nullptr;
  float* slots = nullptr;
  size_t stride = 0;
  uint32_t count = 0, width = 0, rect[4] = {0, 0, 0, 0};
  // non-null: the features of the same frames, folded likewise
  float* feat = nullptr;
  float* fslots = nullptr;
  size_t fstride = 0;
  uint32_t fnpix = 0;
};

struct StepPlan {
  std::vector<cvk::JobDesc> jobs;  // path_base filled, sorted
  uint32_t paths = 0;
  std::vector<FoldDesc> folds;
};

struct RunOpts {
  cvh::LaunchCfg cfg;
  uint32_t max_bounces = 5;
  uint32_t flags = 0;
  bool packed = true;
  const cvk::DevNode* nodes = nullptr;  // shared scene, or null for per-job scenes
  const cvk::DevTri* tris = nullptr;
  bool multi_job = false;
  uint32_t max_depth = 0;
  bool graph = false;
  // held around graph capture: a HIP call from another thread of the
  // caller (a copy, an allocation) during the capture would invalidate it
  std::mutex* capture_mu = nullptr;
};

struct GraphKey {
  int stack = 0;
  uint32_t blocks = 0, refill = 0, fetch_batches = 0, max_bounces = 0, block_waves = 0;
  uint32_t interval = 0, leaf_batch = 0, width = 2, sort_bits = 0, parts = 1;
  bool packed = false, count = false, multi_job = false;
  const void* nodes = nullptr;
  const void* tris = nullptr;
  size_t capacity = 0;
  bool operator==(const GraphKey& o) const {
    return stack == o.stack && blocks == o.blocks && refill == o.refill &&
           block_waves == o.block_waves && interval == o.interval && leaf_batch == o.leaf_batch &&
           fetch_batches == o.fetch_batches && max_bounces == o.max_bounces &&
           packed == o.packed && count == o.count && multi_job == o.multi_job &&
           nodes == o.nodes && tris == o.tris && capacity == o.capacity && width == o.width &&
           sort_bits == o.sort_bits && parts == o.parts;
  }
};

// Header of a step upload, copied over the first words of Ctrl.
struct StepHeader {
  uint32_t paths;
  uint32_t jobs;
};

}  // namespace

// Queues, control block, job table and the staging of one pipeline.  Owned
// by a renderer or a batch; used by one call at a time (the owner's mutex).
struct Pipeline {
  hipStream_t stream = nullptr;
  size_t capacity = 0;  // paths per step the queues hold
  DevBuf<float4> r0[2], r1[2], col[2];
  DevBuf<uint64_t> rng[2];
  DevBuf<uint4> hits;
  DevBuf<cvk::Ctrl> ctrl;
  DevBuf<cvk::JobDesc> jobs[kRingSlots];
  size_t max_jobs = 0;
  HostBuf<StepHeader> head[kRingSlots];
  HostBuf<cvk::JobDesc> stage[kRingSlots];
  hipEvent_t slot_done[kRingSlots] = {};
  uint32_t next_slot = 0;
  DevBuf<uint2> spill;
  // the ray sort: keys and traversal order per path, key buckets and the
  // scan's block totals
  DevBuf<uint32_t> sort_keys, sort_order, sort_buckets, sort_sums;
  // one captured step per job-table slot (the slot's table address is part
  // of the recorded launches)
  hipGraphExec_t graphs[kRingSlots] = {};
  hipGraph_t graph_defs[kRingSlots] = {};
  GraphKey keys[kRingSlots];

  ~Pipeline() { release(); }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (uint32_t k = 0; k < kRingSlots; k++) {
      if (slot_done[k]) (void)hipEventDestroy(slot_done[k]);
      slot_done[k] = nullptr;
      jobs[k].reset();
      head[k].reset();
      stage[k].reset();
    }
    for (int k = 0; k < 2; k++) {
      r0[k].reset();
      r1[k].reset();
      col[k].reset();
      rng[k].reset();
    }
    hits.reset();
    ctrl.reset();
    spill.reset();
    sort_keys.reset();
    sort_order.reset();
    sort_buckets.reset();
    sort_sums.reset();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  void drop_graphs() {
    for (uint32_t k = 0; k < kRingSlots; k++) {
      if (graphs[k]) (void)hipGraphExecDestroy(graphs[k]);
      if (graph_defs[k]) (void)hipGraphDestroy(graph_defs[k]);
      graphs[k] = nullptr;
      graph_defs[k] = nullptr;
    }
  }

  cv_status init() {
    HIPTRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    HIPTRY(ctrl.alloc(1), "alloc ctrl");
    for (uint32_t k = 0; k < kRingSlots; k++) {
      HIPTRY(hipEventCreateWithFlags(&slot_done[k], hipEventDisableTiming), "event");
      HIPTRY(head[k].reserve(1), "pinned alloc");
    }
    return CV_OK;
  }

  cv_status reserve(size_t paths, size_t njobs) {
    if (paths > capacity) {
      HIPTRY(hipStreamSynchronize(stream), "sync");
      drop_graphs();
      for (int k = 0; k < 2; k++) {
        HIPTRY(r0[k].alloc(paths), "alloc queues");
        HIPTRY(r1[k].alloc(paths), "alloc queues");
        HIPTRY(col[k].alloc(paths), "alloc queues");
        HIPTRY(rng[k].alloc(paths), "alloc queues");
      }
      HIPTRY(hits.alloc(paths), "alloc hits");
      capacity = paths;
    }
    if (njobs > max_jobs) {
      HIPTRY(hipStreamSynchronize(stream), "sync");
      drop_graphs();
      for (uint32_t k = 0; k < kRingSlots; k++) {
        HIPTRY(jobs[k].alloc(njobs), "alloc job table");
        HIPTRY(stage[k].reserve(njobs), "pinned alloc");
      }
      max_jobs = njobs;
    }
    return CV_OK;
  }

  // Key bits of the ray sort for sort_bits requested bits: the job index
  // (multi-job steps, up to max_jobs jobs) above the octant and Morton code,
  // at most 24 bits in all (fewer Morton bits per axis when jobs need them).
  static void sort_key_bits(uint32_t sort_bits, bool multi_job, size_t max_jobs,
                            uint32_t* axis_bits, uint32_t* key_bits) {
    uint32_t jb = 0;
    while (multi_job && (1ull << jb) < max_jobs) jb++;
    uint32_t ab = (sort_bits - 3) / 3;
    while (ab > 0 && jb + 3 + 3 * ab > 24) ab--;
    *axis_bits = ab;
    *key_bits = jb + 3 + 3 * ab;
  }

  cv_status reserve_sort(uint32_t key_bits) {
    const size_t nb = (size_t)1 << key_bits;
    if (sort_keys.n < capacity || sort_order.n < capacity || sort_buckets.n < nb ||
        sort_sums.n < cvk::kSortScanBlock + 1) {
      HIPTRY(hipStreamSynchronize(stream), "sync");
      drop_graphs();
      if (sort_keys.n < capacity) HIPTRY(sort_keys.alloc(capacity), "alloc sort");
      if (sort_order.n < capacity) HIPTRY(sort_order.alloc(capacity), "alloc sort");
      if (sort_buckets.n < nb) HIPTRY(sort_buckets.alloc(nb), "alloc sort");
      if (sort_sums.n < cvk::kSortScanBlock + 1)
        HIPTRY(sort_sums.alloc(cvk::kSortScanBlock + 1), "alloc sort");
    }
    return CV_OK;
  }

  cv_status reserve_spill(const cvh::LaunchCfg& cfg, uint32_t max_depth) {
    uint32_t need = max_depth + 4;
    size_t per_lane = need > (uint32_t)cfg.stack ? (size_t)(need - cfg.stack) : 0;
    // lanes of the launch: the waves rounded up to whole 4-wave workgroups
    size_t entries = per_lane * (((size_t)cfg.blocks + 3) / 4 * 4) * 64 + 1;
    if (spill.n < entries) {
      HIPTRY(hipStreamSynchronize(stream), "sync");
      drop_graphs();
      HIPTRY(spill.alloc(entries), "alloc spill");
    }
    return CV_OK;
  }
};

namespace {

// A recorded launch: kind (0 generate, 1 extend, 2 shade, 3 ray sort), bounce, the
// frame of the step's first job, and its start / end events.
struct LaunchEvent {
  int kind;
  uint32_t bounce, frame;
  hipEvent_t start, end;
};
typedef std::vector<LaunchEvent> EventList;

// The trace log (trace.rs:22-152 restated for device spans): a bounded
// ring of kernel spans, exported in the Chrome trace format.
struct TraceSpan {
  int kind;
  double ts_us, dur_us;
  uint32_t frame, id;
};
std::mutex g_trace_mu;
bool g_trace_on = false;
size_t g_trace_limit = 0;
std::deque<TraceSpan> g_trace;
hipEvent_t g_trace_epoch = nullptr;

bool trace_enabled() {
  std::lock_guard<std::mutex> lk(g_trace_mu);
  return g_trace_on;
}

void trace_record(const EventList& evs) {
  std::lock_guard<std::mutex> lk(g_trace_mu);
  if (!g_trace_on) return;
  for (const LaunchEvent& e : evs) {
    float t0 = 0.0f, d = 0.0f;
    if (hipEventElapsedTime(&t0, g_trace_epoch, e.start) != hipSuccess) continue;
    if (hipEventElapsedTime(&d, e.start, e.end) != hipSuccess) continue;
    if (g_trace.size() == g_trace_limit) g_trace.pop_front();
    g_trace.push_back(TraceSpan{e.kind, t0 * 1e3, d * 1e3, e.frame, e.bounce});
  }
}

// The launches of one step (generate, then extend + shade per bounce) for
// job-table slot `slot`.  With graph == true every grid is sized for the
// pipeline's capacity, so that the recorded launches serve any step.
cv_status launch_step(Pipeline& p, const RunOpts& o, uint32_t slot, uint32_t paths, bool graph,
                      EventList* evs, uint32_t frame = 0) {
  hipStream_t st = p.stream;
  uint32_t bounce = 0;
  struct Pair {
    hipEvent_t first, second;
  };
  Pair pair{};
  auto ev_pair = [&](int kind) -> Pair* {
    if (!evs) return nullptr;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return nullptr;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return nullptr;
    }
    evs->push_back(LaunchEvent{kind, bounce, frame, a, b});
    pair = Pair{a, b};
    return &pair;
  };
  const size_t grid_paths = graph ? p.capacity : paths;
  const cvk::JobDesc* jobs = p.jobs[slot].p;
  cvk::GenerateArgs g{};
  g.ctrl = p.ctrl.p;
  g.q = cvk::Queue{p.r0[0].p, p.r1[0].p, p.col[0].p, p.rng[0].p};
  g.jobs = jobs;
  auto* ep = ev_pair(0);
  if (ep) (void)hipEventRecord(ep->first, st);
  HIPTRY(cvk::launch_generate(g, cvh::grid_for(grid_paths, 256, 8192), st), "generate");
  if (ep) (void)hipEventRecord(ep->second, st);
  const bool count = o.flags & CV_RENDER_COUNT;
  const bool multi_scene = o.nodes == nullptr;
  for (uint32_t b = 0; b < o.max_bounces; b++) {
    const int cur = b & 1, nxt = cur ^ 1;
    bounce = b;
    cvk::ExtendArgs x{};
    x.ctrl = p.ctrl.p;
    x.bounce = b;
    x.r0 = p.r0[cur].p;
    x.r1 = p.r1[cur].p;
    x.hits = p.hits.p;
    x.nodes = o.nodes;
    x.tris = o.tris;
    x.jobs = jobs;
    x.spill = p.spill.p;
    x.batches = o.cfg.fetch_batches;
    x.refill = o.cfg.refill <= 64 ? o.cfg.refill : (multi_scene ? 64u : 0u);
    x.interval = o.cfg.refill_interval;
    x.leaf_batch = o.cfg.leaf_batch <= 64 ? o.cfg.leaf_batch : 0u;
    x.parts = o.cfg.parts;
    if (o.cfg.sort_bits && b > 0) {
      // after the camera rays: traverse the bounce's paths in key order
      cvk::SortArgs so{};
      so.ctrl = p.ctrl.p;
      so.bounce = b;
      so.r0 = x.r0;
      so.r1 = x.r1;
      so.jobs = jobs;
      so.multi_job = o.multi_job ? 1u : 0u;
      Pipeline::sort_key_bits(o.cfg.sort_bits, o.multi_job, p.max_jobs, &so.axis_bits,
                              &so.key_bits);
      so.keys = p.sort_keys.p;
      so.buckets = p.sort_buckets.p;
      so.block_sums = p.sort_sums.p;
      so.order = p.sort_order.p;
      ep = ev_pair(3);
      if (ep) (void)hipEventRecord(ep->first, st);
      HIPTRY(cvk::launch_sort(so, (uint32_t)grid_paths, st), "ray sort");
      if (ep) (void)hipEventRecord(ep->second, st);
      x.order = p.sort_order.p;
    }
    ep = ev_pair(1);
    if (ep) (void)hipEventRecord(ep->first, st);
    HIPTRY(cvk::launch_extend(x, count, o.packed, multi_scene, o.cfg.stack, o.cfg.blocks,
                              o.cfg.block_waves, st, o.cfg.width),
           "extend");
    if (ep) (void)hipEventRecord(ep->second, st);
    cvk::ShadeArgs sh{};
    sh.ctrl = p.ctrl.p;
    sh.bounce = b;
    sh.max_bounces = o.max_bounces;
    sh.in = cvk::Queue{p.r0[cur].p, p.r1[cur].p, p.col[cur].p, p.rng[cur].p};
    sh.out = cvk::Queue{p.r0[nxt].p, p.r1[nxt].p, p.col[nxt].p, p.rng[nxt].p};
    sh.hits = p.hits.p;
    sh.jobs = jobs;
    ep = ev_pair(2);
    if (ep) (void)hipEventRecord(ep->first, st);
    HIPTRY(cvk::launch_shade(sh, o.multi_job, !multi_scene, (uint32_t)grid_paths, st), "shade");
    if (ep) (void)hipEventRecord(ep->second, st);
  }
  return CV_OK;
}

// Runs the steps in order on the pipeline's stream and waits for them.
cv_status pipeline_run(Pipeline& p, const std::vector<StepPlan>& steps, const RunOpts& o,
                       cv_render_stats* out) {
  size_t max_paths = 0, max_jobs = 0;
  for (const StepPlan& s : steps) {
    max_paths = std::max<size_t>(max_paths, s.paths);
    max_jobs = std::max(max_jobs, s.jobs.size());
  }
  cv_status cs = p.reserve(max_paths, max_jobs);
  if (cs != CV_OK) return cs;
  cs = p.reserve_spill(o.cfg, o.max_depth);
  if (cs != CV_OK) return cs;
  if (o.cfg.sort_bits && o.max_bounces > 1) {
    uint32_t ab = 0, kb = 0;
    Pipeline::sort_key_bits(o.cfg.sort_bits, o.multi_job, p.max_jobs, &ab, &kb);
    if ((cs = p.reserve_sort(kb)) != CV_OK) return cs;
  }
  hipStream_t st = p.stream;
  const bool timing = (o.flags & CV_RENDER_TIMING) || trace_enabled();
  const bool use_graph = o.graph && !timing;
  HIPTRY(cvk::launch_zero_u32(reinterpret_cast<uint32_t*>(&p.ctrl.p->sum_aabb), 6, st), "zero");
  GraphKey key;
  key.stack = o.cfg.stack;
  key.blocks = o.cfg.blocks;
  key.refill = o.cfg.refill;
  key.block_waves = o.cfg.block_waves;
  key.interval = o.cfg.refill_interval;
  key.leaf_batch = o.cfg.leaf_batch;
  key.width = o.cfg.width;
  key.sort_bits = o.cfg.sort_bits;
  key.parts = o.cfg.parts;
  key.fetch_batches = o.cfg.fetch_batches;
  key.max_bounces = o.max_bounces;
  key.packed = o.packed;
  key.count = o.flags & CV_RENDER_COUNT;
  key.multi_job = o.multi_job;
  key.nodes = o.nodes;
  key.tris = o.tris;
  key.capacity = p.capacity;
  EventList evs;
  // t0 / t1 bracket the steps; every exit below the loop destroys them and
  // waits for the stream, also when a step fails part-way
  hipEvent_t t0 = nullptr, t1 = nullptr;
  HIPTRY(hipEventCreate(&t0), "event");
  if (hipEventCreate(&t1) != hipSuccess) {
    (void)hipEventDestroy(t0);
    return fail(CV_ERR_DEVICE, "event");
  }
  cv_status status = CV_OK;
  if (hipEventRecord(t0, st) != hipSuccess) status = fail(CV_ERR_DEVICE, "event");
  for (const StepPlan& s : steps) {
    if (status != CV_OK) break;
    if (s.paths == 0) continue;
    const uint32_t slot = p.next_slot;
    p.next_slot = (p.next_slot + 1) % kRingSlots;
    // the slot's previous upload must have been read before it is rewritten
    if (hipEventSynchronize(p.slot_done[slot]) != hipSuccess) {
      status = fail(CV_ERR_DEVICE, "event");
      break;
    }
    p.head[slot].p[0] = StepHeader{s.paths, (uint32_t)s.jobs.size()};
    memcpy(p.stage[slot].p, s.jobs.data(), s.jobs.size() * sizeof(cvk::JobDesc));
    if (hipMemcpyAsync(p.ctrl.p, p.head[slot].p, sizeof(StepHeader), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(p.jobs[slot].p, p.stage[slot].p, s.jobs.size() * sizeof(cvk::JobDesc),
                       hipMemcpyHostToDevice, st) ||
        hipEventRecord(p.slot_done[slot], st)) {
      status = fail(CV_ERR_DEVICE, "step upload");
      break;
    }
    if (use_graph) {
      if (!p.graphs[slot] || !(p.keys[slot] == key)) {
        if (p.graphs[slot]) (void)hipGraphExecDestroy(p.graphs[slot]);
        if (p.graph_defs[slot]) (void)hipGraphDestroy(p.graph_defs[slot]);
        p.graphs[slot] = nullptr;
        p.graph_defs[slot] = nullptr;
        std::unique_lock<std::mutex> cl;
        if (o.capture_mu) cl = std::unique_lock<std::mutex>(*o.capture_mu);
        hipError_t be = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
        if (be != hipSuccess) {
          status = hip_fail(be, "graph capture");
          break;
        }
        cv_status ls = launch_step(p, o, slot, s.paths, true, nullptr);
        hipGraph_t gdef = nullptr;
        hipError_t ce = hipStreamEndCapture(st, &gdef);
        cl = std::unique_lock<std::mutex>();
        if (ls != CV_OK || ce != hipSuccess) {
          // nothing of a failed capture ran: clear the error and launch the
          // step directly (a real launch error fails there again)
          if (gdef) (void)hipGraphDestroy(gdef);
          (void)hipGetLastError();
          status = launch_step(p, o, slot, s.paths, false, nullptr);
          if (status != CV_OK) break;
        } else {
          p.graph_defs[slot] = gdef;
          hipError_t ie = hipGraphInstantiate(&p.graphs[slot], gdef, nullptr, nullptr, 0);
          if (ie != hipSuccess) {
            p.graphs[slot] = nullptr;
            status = hip_fail(ie, "graph instantiate");
            break;
          }
          p.keys[slot] = key;
        }
      }
      if (p.graphs[slot]) {
        hipError_t le = hipGraphLaunch(p.graphs[slot], st);
        if (le != hipSuccess) {
          status = hip_fail(le, "graph launch");
          break;
        }
      }
    } else {
      status = launch_step(p, o, slot, s.paths, false, timing ? &evs : nullptr,
                           s.jobs.empty() ? 0u : s.jobs[0].frame);
      if (status != CV_OK) break;
    }
    for (const FoldDesc& fd : s.folds) {
      hipError_t fe = cvk::launch_fold_hdr(fd.hdr, fd.slots, fd.stride, fd.count, fd.width,
                                           fd.rect[0], fd.rect[1], fd.rect[2], fd.rect[3], st);
      if (fe == hipSuccess && fd.feat)
        fe = cvk::launch_fold_features(fd.feat, fd.fslots, fd.fstride, fd.count, fd.fnpix, st);
      if (fe != hipSuccess) {
        status = hip_fail(fe, "fold");
        break;
      }
    }
  }
  (void)hipEventRecord(t1, st);
  hipError_t se = hipStreamSynchronize(st);
  if (status == CV_OK && se != hipSuccess) status = hip_fail(se, "render");
  cv_render_stats stt{};
  uint32_t nsteps = 0;
  for (const StepPlan& s : steps) {
    stt.paths += s.paths;
    nsteps += s.paths ? 1 : 0;
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  stt.total_ms = ms;
  trace_record(evs);
  for (auto& e : evs) {
    float m = 0.0f;
    (void)hipEventElapsedTime(&m, e.start, e.end);
    if (e.kind == 0) stt.generate_ms += m;
    if (e.kind == 1) {
      stt.extend_ms += m;
      stt.extend_launches++;
    }
    if (e.kind == 2) stt.shade_ms += m;
    if (e.kind == 3) stt.sort_ms += m;
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.end);
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (!timing) stt.extend_launches = (uint64_t)nsteps * o.max_bounces;
  if (status == CV_OK) {
    cvk::Ctrl hc;
    if (hipMemcpy(&hc, p.ctrl.p, sizeof hc, hipMemcpyDeviceToHost) == hipSuccess) {
      stt.sum_numi_aabb = hc.sum_aabb;
      stt.sum_numi_tri = hc.sum_tri;
      stt.extend_rays = hc.total_rays;
    }
  }
  if (out) *out = stt;
  return status;
}

}  // namespace

// ---------------------------------------------------------------- renderer

struct cv_renderer {
  cv_scene* scene = nullptr;
  uint32_t width = 0, height = 0;
  cv_render_opts opts{};
  Pipeline pipe;
  DevBuf<float> hdr, features;
  // accumulation buffers of the 2nd..Kth frame of a step (zero between calls)
  DevBuf<float> slots;
  bool slots_clean = false;
  DevBuf<uint32_t> bitmap, gbuffer;
  DevBuf<uint4> ref_patches;  // CV_RNG_REF_PATCH: the call's patches (x, y, width, 0)
  cv_render_stats stats{};
  std::mutex mu;
};

namespace {

// allow_ref_patch: renderers render CV_RNG_REF_PATCH too (patch-serial, a
// correctness mode); batches and servers share launches between jobs and
// take CV_RNG_PIXEL only.
cv_status check_opts(const cv_render_opts* opts, cv_render_opts* o, bool allow_ref_patch) {
  *o = cv_render_opts{CV_RNG_PIXEL, CV_APPROX_PORTABLE, 5};
  if (opts) *o = *opts;
  if (o->max_bounces == 0) o->max_bounces = 5;
  if (o->max_bounces >= (uint32_t)cvk::kMaxBounceSlots)
    return fail(CV_ERR_INVALID_ARGUMENT, "max_bounces too large");
  if (o->rng_mode != CV_RNG_PIXEL && !(allow_ref_patch && o->rng_mode == CV_RNG_REF_PATCH))
    return fail(CV_ERR_INVALID_ARGUMENT,
                allow_ref_patch ? "rng_mode must be CV_RNG_PIXEL or CV_RNG_REF_PATCH"
                                : "batches and servers render CV_RNG_PIXEL");
  if (o->approx_mode != CV_APPROX_PORTABLE)
    return fail(CV_ERR_INVALID_ARGUMENT, "the GPU renderer implements CV_APPROX_PORTABLE");
  return CV_OK;
}

cv_status check_frame(const cv_scene* s, uint32_t width, uint32_t height) {
  // The 16x4-block HDR layout needs width % 16 == 0 and height % 4 == 0
  // (renderer.rs:366-374, 384-385); the reference's RenderBuffer also
  // rounds heights to 16, which the per-pixel GPU path does not need.
  if ((width & 15) || (height & 3) || width == 0 || height == 0)
    return fail(CV_ERR_NOT_MULTIPLE_OF_16, "width must be a multiple of 16 and height of 4");
  if ((uint64_t)width * height >= (1ull << 31)) return fail(CV_ERR_INVALID_ARGUMENT, "frame too large");
  if (s->ndirect < 8)
    return fail(CV_ERR_LIGHT_COUNT, "the scene has " + std::to_string(s->ndirect) +
                                        " direct-sample triangles; direct sampling needs 8");
  return CV_OK;
}

// The rectangle of a render call; patch > 0 for the patch calls (one
// reference patch of that width at x, y), 0 for whole frames.
struct Rect {
  uint32_t x, y, w, h;
  uint32_t patch;
};

// CV_RNG_REF_PATCH: the reference's patch streams.  A patch call renders
// its one patch; a frame call renders the patch loop of main.rs:174-214:
// 32 x 32 patches, columns outer, rows inner, and where fewer than 32
// columns or rows remain, that 32 x 32 cell tiled with 16 x 16 patches
// (the reference needs frames of whole patches).  Each patch seeds its
// stream with (x, y, frame), so the partition is part of the result.
cv_status ref_patch_render(cv_renderer* r, const Rect& rc, uint32_t first, uint32_t nframes,
                           const cvk::JobDesc& job, bool features) {
  if (features)
    return fail(CV_ERR_INVALID_ARGUMENT, "first-hit features are rendered with CV_RNG_PIXEL");
  cv_scene* s = r->scene;
  const uint32_t W = r->width, H = r->height;
  std::vector<uint4> patches;
  if (rc.patch) {
    patches.push_back(make_uint4(rc.x, rc.y, rc.patch, 0u));
  } else {
    if ((H & 15) != 0)
      return fail(CV_ERR_NOT_MULTIPLE_OF_16,
                  "CV_RNG_REF_PATCH frames are whole 16 x 16 patches: height must be a multiple "
                  "of 16");
    const uint32_t npx = (W + 31) / 32, npy = (H + 31) / 32;
    for (uint32_t p = 0; p < npx * npy; p++) {
      const uint32_t x = (p / npy) * 32, y = (p % npy) * 32;
      const uint32_t pw = (W - x >= 32 && H - y >= 32) ? 32u : 16u;
      const uint32_t xe = std::min(x + 32, W), ye = std::min(y + 32, H);
      for (uint32_t yy = y; yy < ye; yy += pw)
        for (uint32_t xx = x; xx < xe; xx += pw) patches.push_back(make_uint4(xx, yy, pw, 0u));
    }
  }
  hipStream_t st = r->pipe.stream;
  cv_status cs = r->pipe.reserve_spill(s->cfg, s->info.depth);
  if (cs != CV_OK) return cs;
  if (r->ref_patches.n < patches.size()) HIPTRY(r->ref_patches.alloc(patches.size()), "alloc");
  cs = cvh::copy_caller_memory(r->ref_patches.p, patches.data(), patches.size() * sizeof(uint4),
                               true, st);
  if (cs != CV_OK) return cs;
  cvk::RefPatchArgs a{};
  a.job = job;
  a.patches = r->ref_patches.p;
  a.npatches = (uint32_t)patches.size();
  a.first_frame = first;
  a.nframes = nframes;
  a.max_bounces = r->opts.max_bounces;
  a.spill = r->pipe.spill.p;
  // 8 patches per 64-lane workgroup, within the spill area's waves
  const uint32_t blocks = std::max<uint32_t>(
      1u, std::min<uint32_t>((a.npatches + 7) / 8, s->cfg.blocks));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  HIPTRY(hipEventCreate(&t0), "event");
  if (hipEventCreate(&t1) != hipSuccess) {
    (void)hipEventDestroy(t0);
    return fail(CV_ERR_DEVICE, "event");
  }
  cv_status status = CV_OK;
  (void)hipEventRecord(t0, st);
  hipError_t le = cvk::launch_ref_patch(a, s->packed, s->cfg.stack, blocks, st);
  if (le != hipSuccess) status = hip_fail(le, "ref patch render");
  (void)hipEventRecord(t1, st);
  hipError_t se = hipStreamSynchronize(st);
  if (status == CV_OK && se != hipSuccess) status = hip_fail(se, "ref patch render");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  cv_render_stats stt{};
  stt.paths = (uint64_t)rc.w * rc.h * nframes;
  stt.total_ms = ms;
  stt.extend_ms = ms;
  stt.extend_launches = 1;
  r->stats = stt;
  return status;
}

cv_status run_frames(cv_renderer* r, const Rect& rc, uint32_t first, uint32_t nframes, float* hdr,
                     uint32_t* bitmap, uint32_t* gbuffer, uint32_t flags, bool hdr_mode,
                     float* features = nullptr) {
  hipStream_t st = r->pipe.stream;
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
  float* ft_dev = nullptr;
  const size_t npix_all = (size_t)W * r->height;
  if (features) {
    if (dev) {
      ft_dev = features;
    } else {
      if (r->features.n < npix_all * 8) HIPTRY(r->features.alloc(npix_all * 8), "alloc features");
      HIPTRY(hipMemcpyAsync(r->features.p, features, npix_all * 32, hipMemcpyHostToDevice, st),
             "upload features");
      ft_dev = r->features.p;
    }
  }
  cv_scene* s = r->scene;
  cv_status ws = cvh::ensure_wide(s, s->cfg.width);
  if (ws != CV_OK) return ws;
  cvk::JobDesc base = cvh::job_for_scene(s, s->cfg.width);
  base.width = r->width;
  base.height = r->height;
  base.rect_x = rc.x;
  base.rect_y = rc.y;
  base.rect_w = rc.w;
  base.rect_h = rc.h;
  base.hdr = hdr_dev;
  base.bitmap = bm_dev;
  base.features = ft_dev;
  base.path_base = 0;
  cv_status status = CV_OK;
  if (r->opts.rng_mode == CV_RNG_REF_PATCH) {
    // every frame writes the gbuffer; the last frame's remains
    cvk::JobDesc job = base;
    job.gbuffer = gb_dev;
    status = ref_patch_render(r, rc, first, nframes, job, features != nullptr);
  } else {
  // Frames per step: an accumulation renders K frames in the same launches
  // (a launch's last rays keep most of the GPU idle, so larger launches are
  // faster per ray); frame k > 0 of a step accumulates into its own zeroed
  // buffer and a fold adds those to the image in frame order afterwards.
  // The features of the denoiser are summed in place, so they keep K = 1.
  const uint32_t npix = rc.w * rc.h;
  uint32_t K = s->cfg.frames_per_step;
  if (K == CV_FRAMES_AUTO) K = std::max<uint32_t>(1u, (1u << 25) / std::max(npix, 1u));
  K = std::min({K, nframes, 64u});
  if (!hdr_mode || features || bitmap) K = 1;
  const size_t slot_floats = (size_t)W * r->height * 3;
  if (K > 1) {
    const size_t need = (size_t)(K - 1) * slot_floats;
    if (r->slots.n < need) {
      HIPTRY(r->slots.alloc(need), "alloc accumulation buffers");
      r->slots_clean = false;
    }
    if (!r->slots_clean)
      HIPTRY(hipMemsetAsync(r->slots.p, 0, r->slots.n * sizeof(float), st), "clear");
    r->slots_clean = false;
  }
  std::vector<StepPlan> steps;
  for (uint32_t f0 = 0; f0 < nframes; f0 += K) {
    const uint32_t kk = std::min(K, nframes - f0);
    StepPlan sp;
    for (uint32_t k = 0; k < kk; k++) {
      cvk::JobDesc j = base;
      j.frame = first + f0 + k;
      j.hdr = k == 0 ? hdr_dev : r->slots.p + (size_t)(k - 1) * slot_floats;
      // every frame rewrites the gbuffer; the last one is what remains
      j.gbuffer = f0 + k + 1 == nframes ? gb_dev : nullptr;
      j.path_base = k * npix;
      sp.jobs.push_back(j);
    }
    sp.paths = kk * npix;
    if (kk > 1) {
      FoldDesc fd;
      fd.hdr = hdr_dev;
      fd.slots = r->slots.p;
      fd.stride = slot_floats;
      fd.count = kk - 1;
      fd.width = W;
      fd.rect[0] = rc.x;
      fd.rect[1] = rc.y;
      fd.rect[2] = rc.w;
      fd.rect[3] = rc.h;
      sp.folds.push_back(fd);
    }
    steps.push_back(std::move(sp));
  }
  RunOpts o;
  o.cfg = s->cfg;
  o.max_bounces = r->opts.max_bounces;
  o.flags = flags;
  o.packed = s->packed;
  o.nodes = cvh::traversal_nodes(s, s->cfg.width);
  o.tris = s->tris.p;
  o.multi_job = K > 1;
  o.max_depth = cvh::stack_depth(s, s->cfg.width);
  o.graph = flags & CV_RENDER_GRAPH;
  status = pipeline_run(r->pipe, steps, o, &r->stats);
  if (status == CV_OK && K > 1) r->slots_clean = true;
  }
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
    if (status == CV_OK && features)
      status = hipMemcpyAsync(features, r->features.p, npix_all * 32, hipMemcpyDeviceToHost, st) ==
                       hipSuccess
                   ? CV_OK
                   : fail(CV_ERR_DEVICE, "download features");
    hipError_t se = hipStreamSynchronize(st);
    if (status == CV_OK && se != hipSuccess) status = hip_fail(se, "download");
  }
  return status;
}

cv_status patch_args(cv_renderer* r, uint32_t pw, uint32_t x, uint32_t y) {
  if (!r) return fail(CV_ERR_INVALID_ARGUMENT, "null renderer");
  if ((pw & 15) || pw == 0) return fail(CV_ERR_NOT_MULTIPLE_OF_16, "patch width must be a multiple of 16");
  if ((x & 15) || (y & 15)) return fail(CV_ERR_NOT_MULTIPLE_OF_16, "patch origin must be a multiple of 16");
  if (x + pw > r->width || y + pw > r->height) return fail(CV_ERR_OUT_OF_RANGE, "patch outside the frame");
  return CV_OK;
}

}  // namespace

// ------------------------------------------------------------------ batches

struct cv_batch {
  cv_render_opts opts{};
  cvh::LaunchCfg cfg;
  uint64_t max_paths_per_step = 0;
  Pipeline pipe;
  cv_render_stats stats{};
  uint32_t last_steps = 0;
  std::mutex mu;
};

extern "C" {

cv_status cv_renderer_create(cv_scene* s, uint32_t width, uint32_t height,
                             const cv_render_opts* opts, cv_renderer** out) {
  if (!s || !out) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  cv_status cs = check_frame(s, width, height);
  if (cs != CV_OK) return cs;
  cv_render_opts o;
  cs = check_opts(opts, &o, true);
  if (cs != CV_OK) return cs;
  cv_renderer* r = new cv_renderer();
  r->scene = s;
  r->width = width;
  r->height = height;
  r->opts = o;
  auto bail = [&](cv_status st) {
    cv_renderer_destroy(r);
    return st;
  };
  if ((cs = r->pipe.init()) != CV_OK) return bail(cs);
  const size_t npix = (size_t)width * height;
  if ((cs = r->pipe.reserve(npix, 1)) != CV_OK) return bail(cs);
  hipError_t e;
  if ((e = r->hdr.alloc(npix * 3)) || (e = r->bitmap.alloc(npix)) || (e = r->gbuffer.alloc(npix)))
    return bail(hip_fail(e, "alloc buffers"));
  *out = r;
  return CV_OK;
}

void cv_renderer_destroy(cv_renderer* r) {
  if (!r) return;
  r->pipe.release();
  r->hdr.reset();
  r->features.reset();
  r->bitmap.reset();
  r->gbuffer.reset();
  r->ref_patches.reset();
  delete r;
}

cv_status cv_renderer_render_patch_u8(cv_renderer* r, uint32_t* bitmap, uint32_t* gbuffer,
                                      uint32_t patch_width, uint32_t x, uint32_t y,
                                      uint32_t frame, uint32_t flags) {
  cv_status st = patch_args(r, patch_width, x, y);
  if (st != CV_OK) return st;
  if (!bitmap) return fail(CV_ERR_INVALID_ARGUMENT, "null bitmap");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{x, y, patch_width, patch_width, patch_width}, frame, 1, nullptr, bitmap, gbuffer,
                    flags, false);
}

cv_status cv_renderer_accumulate_patch_f32(cv_renderer* r, float* hdr, uint32_t* gbuffer,
                                           uint32_t patch_width, uint32_t x, uint32_t y,
                                           uint32_t frame, uint32_t flags) {
  cv_status st = patch_args(r, patch_width, x, y);
  if (st != CV_OK) return st;
  if (!hdr) return fail(CV_ERR_INVALID_ARGUMENT, "null hdr");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{x, y, patch_width, patch_width, patch_width}, frame, 1, hdr, nullptr, gbuffer,
                    flags, true);
}

cv_status cv_renderer_render_frame_u8(cv_renderer* r, uint32_t* bitmap, uint32_t* gbuffer,
                                      uint32_t frame, uint32_t flags) {
  if (!r || !bitmap) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{0, 0, r->width, r->height, 0}, frame, 1, nullptr, bitmap, gbuffer, flags,
                    false);
}

cv_status cv_renderer_accumulate_frames_f32(cv_renderer* r, float* hdr, uint32_t* gbuffer,
                                            uint32_t first_frame, uint32_t n_frames,
                                            uint32_t flags) {
  if (!r || !hdr) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{0, 0, r->width, r->height, 0}, first_frame, n_frames, hdr, nullptr,
                    gbuffer, flags, true);
}

cv_status cv_renderer_accumulate_aov_f32(cv_renderer* r, float* hdr, float* features,
                                         uint32_t first_frame, uint32_t n_frames,
                                         uint32_t flags) {
  if (!r || !hdr || !features) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(r->mu);
  return run_frames(r, Rect{0, 0, r->width, r->height, 0}, first_frame, n_frames, hdr, nullptr,
                    nullptr, flags, true, features);
}

cv_status cv_renderer_buffer_f32_into_bitmap(cv_renderer* r, const float* hdr, uint32_t* bitmap,
                                             uint32_t num_samples, uint32_t flags) {
  if (!r || !hdr || !bitmap) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(r->mu);
  hipStream_t st = r->pipe.stream;
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

// ------------------------------------------------------------------ batches

cv_status cv_batch_create(const cv_render_opts* opts, uint64_t max_paths_per_step,
                          cv_batch** out) {
  if (!out) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  cv_status cs = cvh::ensure_device();
  if (cs != CV_OK) return cs;
  cv_render_opts o;
  if ((cs = check_opts(opts, &o, false)) != CV_OK) return cs;
  if (max_paths_per_step >= (1ull << 31))
    return fail(CV_ERR_INVALID_ARGUMENT, "max_paths_per_step must be below 2^31");
  cv_batch* b = new cv_batch();
  b->opts = o;
  b->max_paths_per_step = max_paths_per_step ? max_paths_per_step : (1ull << 27);
  cv_tuning t{8, 32, 16, 1, 1};
  cvh::apply_tuning(b->cfg, t);
  if ((cs = b->pipe.init()) != CV_OK) {
    cv_batch_destroy(b);
    return cs;
  }
  *out = b;
  return CV_OK;
}

void cv_batch_destroy(cv_batch* b) {
  if (!b) return;
  b->pipe.release();
  delete b;
}

cv_status cv_batch_set_tuning(cv_batch* b, const cv_tuning* t) {
  if (!b || !t) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  cv_status v = cvh::validate_tuning(*t);
  if (v != CV_OK) return v;
  std::lock_guard<std::mutex> g(b->mu);
  cvh::apply_tuning(b->cfg, *t);
  return CV_OK;
}

cv_status cv_batch_get_tuning(const cv_batch* b, cv_tuning* t) {
  if (!b || !t) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *t = cvh::tuning_of(b->cfg);
  return CV_OK;
}

cv_status cv_batch_render(cv_batch* b, const cv_job* jobs, size_t n, uint32_t flags) {
  if (!b || (n && !jobs)) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return CV_OK;
  if (n > 65536) return fail(CV_ERR_INVALID_ARGUMENT, "too many jobs");
  std::lock_guard<std::mutex> g(b->mu);
  const bool dev = flags & CV_PTR_DEVICE;
  uint32_t max_frames = 0, max_depth = 0;
  size_t largest = 0, total = 0;
  bool packed = true, one_scene = true;
  for (size_t i = 0; i < n; i++) {
    const cv_job& J = jobs[i];
    if (!J.scene || !J.hdr) return fail(CV_ERR_INVALID_ARGUMENT, "job without scene or hdr");
    cv_status cs = check_frame(J.scene, J.width, J.height);
    if (cs != CV_OK) return cs;
    if (J.num_frames == 0) return fail(CV_ERR_INVALID_ARGUMENT, "job with zero frames");
    if ((cs = cvh::ensure_wide(J.scene, b->cfg.width)) != CV_OK) return cs;
    max_frames = std::max(max_frames, J.num_frames);
    max_depth = std::max(max_depth, cvh::stack_depth(J.scene, b->cfg.width));
    largest = std::max(largest, (size_t)J.width * J.height);
    total += (size_t)J.width * J.height * J.num_frames;
    packed = packed && J.scene->packed;
    one_scene = one_scene && J.scene == jobs[0].scene;
  }
  // queue capacity: all the batch's paths, within the bound (a job larger
  // than the bound still fits)
  const size_t cap = std::max<size_t>(largest, std::min<size_t>(b->max_paths_per_step, total));
  if (cap >= (1ull << 31)) return fail(CV_ERR_INVALID_ARGUMENT, "step too large");
  // host outputs: device staging per job
  std::vector<DevBuf<float>> d_hdr(dev ? 0 : n);
  std::vector<DevBuf<uint32_t>> d_gb(dev ? 0 : n);
  std::vector<DevBuf<float>> d_ft(dev ? 0 : n);
  hipStream_t st = b->pipe.stream;
  std::vector<cvk::JobDesc> base(n);
  for (size_t i = 0; i < n; i++) {
    const cv_job& J = jobs[i];
    const size_t npix = (size_t)J.width * J.height;
    cvk::JobDesc d = cvh::job_for_scene(J.scene, b->cfg.width);
    d.width = J.width;
    d.height = J.height;
    d.rect_x = d.rect_y = 0;
    d.rect_w = J.width;
    d.rect_h = J.height;
    d.hdr = J.hdr;
    d.gbuffer = J.gbuffer;
    d.features = J.features;
    d.bitmap = nullptr;
    if (!dev) {
      // caller arrays: hdr and features npix * 12 and npix * 32 bytes (read
      // and written), gbuffer npix * 4 (written)
      HIPTRY(d_hdr[i].alloc(npix * 3), "alloc hdr");
      cv_status cs = cvh::copy_caller_memory(d_hdr[i].p, J.hdr, npix * 12, true, st);
      if (cs != CV_OK) return cs;
      d.hdr = d_hdr[i].p;
      if (J.gbuffer) {
        HIPTRY(d_gb[i].alloc(npix), "alloc gbuffer");
        d.gbuffer = d_gb[i].p;
      }
      if (J.features) {
        HIPTRY(d_ft[i].alloc(npix * 8), "alloc features");
        cs = cvh::copy_caller_memory(d_ft[i].p, J.features, npix * 32, true, st);
        if (cs != CV_OK) return cs;
        d.features = d_ft[i].p;
      }
    }
    base[i] = d;
  }
  // Steps take (job, frame) pairs in frame-major order (frame 0 of every
  // job, then frame 1 of every job with more than one frame, ...) up to the
  // queue capacity, so a step can hold several frames of one job once few
  // jobs are left: the first of them accumulates into the job's image, the
  // k-th further one into the job's k-th zeroed buffer, and the step's fold
  // adds those in frame order.  A job with features appears once per step
  // (its features are summed in place).
  std::vector<StepPlan> steps;
  std::vector<std::vector<uint32_t>> extra_of;   // per step entry: 0 (image) or k (k-th buffer)
  std::vector<std::vector<size_t>> job_of_entry; // per step entry: the job
  std::vector<std::vector<size_t>> fold_jobs;    // per step: the job of each fold
  std::vector<uint32_t> occ(n, 0), max_extra(n, 0);
  StepPlan cur;
  std::vector<uint32_t> cur_extra;
  std::vector<size_t> cur_job;
  auto flush = [&]() {
    if (cur.jobs.empty()) return;
    std::vector<size_t> fj;
    for (size_t i = 0; i < n; i++) {
      if (occ[i] > 1) {
        FoldDesc fd;
        fd.hdr = base[i].hdr;
        fd.stride = (size_t)jobs[i].width * jobs[i].height * 3;
        fd.count = occ[i] - 1;
        fd.width = jobs[i].width;
        fd.rect[2] = jobs[i].width;
        fd.rect[3] = jobs[i].height;
        cur.folds.push_back(fd);
        fj.push_back(i);
      }
      occ[i] = 0;
    }
    steps.push_back(std::move(cur));
    extra_of.push_back(std::move(cur_extra));
    job_of_entry.push_back(std::move(cur_job));
    fold_jobs.push_back(std::move(fj));
    cur = StepPlan();
    cur_extra.clear();
    cur_job.clear();
  };
  for (uint32_t f = 0; f < max_frames; f++) {
    for (size_t i = 0; i < n; i++) {
      if (jobs[i].num_frames <= f) continue;
      const uint32_t npix = jobs[i].width * jobs[i].height;
      if (cur.paths + (size_t)npix > cap || (occ[i] > 0 && jobs[i].features)) flush();
      cvk::JobDesc d = base[i];
      d.frame = jobs[i].first_frame + f;
      d.gbuffer = f + 1 == jobs[i].num_frames ? base[i].gbuffer : nullptr;
      d.path_base = cur.paths;
      cur_extra.push_back(occ[i]);
      cur_job.push_back(i);
      max_extra[i] = std::max(max_extra[i], occ[i]);
      occ[i]++;
      cur.jobs.push_back(d);
      cur.paths += npix;
    }
  }
  flush();
  // the further frames' buffers, zeroed (each fold clears what it read)
  std::vector<DevBuf<float>> slot_bufs(n);
  for (size_t i = 0; i < n; i++) {
    if (!max_extra[i]) continue;
    const size_t floats = (size_t)max_extra[i] * jobs[i].width * jobs[i].height * 3;
    HIPTRY(slot_bufs[i].alloc(floats), "alloc accumulation buffers");
    HIPTRY(hipMemsetAsync(slot_bufs[i].p, 0, floats * sizeof(float), st), "clear");
  }
  bool multi_job = false;
  for (size_t k = 0; k < steps.size(); k++) {
    StepPlan& sp = steps[k];
    multi_job = multi_job || sp.jobs.size() > 1;
    for (size_t e = 0; e < sp.jobs.size(); e++) {
      const uint32_t x = extra_of[k][e];
      if (!x) continue;
      const size_t i = job_of_entry[k][e];
      sp.jobs[e].hdr = slot_bufs[i].p + (size_t)(x - 1) * jobs[i].width * jobs[i].height * 3;
    }
    for (size_t q = 0; q < sp.folds.size(); q++) sp.folds[q].slots = slot_bufs[fold_jobs[k][q]].p;
  }
  RunOpts o;
  o.cfg = b->cfg;
  o.max_bounces = b->opts.max_bounces;
  o.flags = flags;
  o.packed = packed;
  o.nodes = one_scene ? cvh::traversal_nodes(jobs[0].scene, b->cfg.width) : nullptr;
  o.tris = one_scene ? jobs[0].scene->tris.p : nullptr;
  o.multi_job = multi_job;
  o.max_depth = max_depth;
  o.graph = flags & CV_RENDER_GRAPH;
  cv_status cs = b->pipe.reserve(cap, 1);
  if (cs != CV_OK) return cs;
  cs = pipeline_run(b->pipe, steps, o, &b->stats);
  b->last_steps = (uint32_t)steps.size();
  if (cs == CV_OK && !dev) {
    for (size_t i = 0; i < n && cs == CV_OK; i++) {
      const size_t npix = (size_t)jobs[i].width * jobs[i].height;
      cs = cvh::copy_caller_memory(jobs[i].hdr, d_hdr[i].p, npix * 12, false, st);
      if (cs == CV_OK && jobs[i].gbuffer)
        cs = cvh::copy_caller_memory(jobs[i].gbuffer, d_gb[i].p, npix * 4, false, st);
      if (cs == CV_OK && jobs[i].features)
        cs = cvh::copy_caller_memory(jobs[i].features, d_ft[i].p, npix * 32, false, st);
    }
  }
  return cs;
}

cv_status cv_batch_autotune(cv_batch* b, const cv_job* jobs, size_t n, uint64_t max_sample_paths,
                            uint32_t flags, cv_tuning* chosen, double* chosen_ms) {
  if (!b || !jobs || n == 0) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  if (n > 65536) return fail(CV_ERR_INVALID_ARGUMENT, "too many jobs");
  if (flags & ~CV_TUNE_TOLERANCE) return fail(CV_ERR_INVALID_ARGUMENT, "unknown autotune flag");
  const bool sweep_width = flags & CV_TUNE_TOLERANCE;
  std::lock_guard<std::mutex> g(b->mu);
  if (max_sample_paths == 0) max_sample_paths = 1ull << 22;
  // sample: the first frame of the jobs in order, up to max_sample_paths
  // paths (at least one job), written to scratch outputs
  StepPlan step;
  std::vector<cv_scene*> sample_scenes;  // the scene of each sample entry
  uint32_t max_depth = 0;
  bool packed = true, one_scene = true, wide_ok = true;
  for (size_t i = 0; i < n; i++) {
    const cv_job& J = jobs[i];
    if (!J.scene) return fail(CV_ERR_INVALID_ARGUMENT, "job without scene");
    cv_status cs = check_frame(J.scene, J.width, J.height);
    if (cs != CV_OK) return cs;
    const uint32_t npix = J.width * J.height;
    if (!step.jobs.empty() && step.paths + (uint64_t)npix > max_sample_paths) continue;
    cvk::JobDesc d = cvh::job_for_scene(J.scene, 2);
    d.width = J.width;
    d.height = J.height;
    d.rect_w = J.width;
    d.rect_h = J.height;
    d.frame = J.first_frame;
    d.path_base = step.paths;
    step.jobs.push_back(d);
    sample_scenes.push_back(J.scene);
    step.paths += npix;
    wide_ok = wide_ok && J.scene->info.max_leaf <= 31 && J.scene->ntris < (1u << 26) &&
              (sweep_width || b->cfg.width != 2);
    max_depth = std::max(max_depth, J.scene->info.depth);
    packed = packed && J.scene->packed;
    one_scene = one_scene && J.scene == jobs[0].scene;
  }
  DevBuf<float> scratch;
  HIPTRY(scratch.alloc((size_t)step.paths * 3), "alloc scratch");
  for (cvk::JobDesc& d : step.jobs) d.hdr = scratch.p + (size_t)d.path_base * 3;
  std::vector<StepPlan> steps(1, step);
  RunOpts o;
  o.max_bounces = b->opts.max_bounces;
  o.flags = 0;
  o.packed = packed;
  o.tris = one_scene ? jobs[0].scene->tris.p : nullptr;
  o.multi_job = step.jobs.size() > 1;
  o.max_depth = max_depth;
  if (wide_ok) {
    // every candidate width: the spill area is sized for the deepest stack
    for (cv_scene* sc : sample_scenes) {
      cv_status ws;
      if ((ws = cvh::ensure_wide(sc, 4)) != CV_OK || (ws = cvh::ensure_wide(sc, 8)) != CV_OK)
        return ws;
      o.max_depth = std::max({o.max_depth, cvh::stack_depth(sc, 4), cvh::stack_depth(sc, 8)});
    }
  }
  auto time_cfg = [&](const cvh::LaunchCfg& c, double* ms) -> cv_status {
    o.cfg = c;
    // the tree of the candidate's width for every entry of the sample step
    for (size_t k = 0; k < steps[0].jobs.size(); k++)
      steps[0].jobs[k].nodes = cvh::traversal_nodes(sample_scenes[k], c.width);
    o.nodes = one_scene ? cvh::traversal_nodes(jobs[0].scene, c.width) : nullptr;
    double best = 1e30;
    for (int rep = 0; rep < 2; rep++) {
      cv_render_stats st{};
      cv_status cs = pipeline_run(b->pipe, steps, o, &st);
      if (cs != CV_OK) return cs;
      best = std::min(best, st.total_ms);
    }
    *ms = best;
    return CV_OK;
  };
  // coordinate descent over the launch knobs, two rounds
  static const uint32_t stacks[] = {8, 16, 24, 32};
  static const uint32_t waves[] = {16, 24, 32, 40, 48};
  static const uint32_t refills[] = {16, 24, 32, 48, 64, 65};
  static const uint32_t block_waves[] = {1, 2, 4};
  static const uint32_t intervals[] = {1, 4, 16, 32};
  static const uint32_t leaf_batches[] = {8, 16, 24, 65};
  static const uint32_t widths[] = {2, 4, 8};
  static const uint32_t sort_bits[] = {0, 12, 24};
  static const uint32_t parts[] = {1, cvk::kParts};
  cvh::LaunchCfg best = b->cfg;
  if (!wide_ok) best.width = 2;
  double best_ms = 0;
  cv_status cs = time_cfg(best, &best_ms);
  if (cs != CV_OK) return cs;
  for (int round = 0; round < 2; round++) {
    for (uint32_t v : stacks) {
      cvh::LaunchCfg c = best;
      c.stack = (int)v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : waves) {
      cvh::LaunchCfg c = best;
      c.blocks = cvh::num_cus() * v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : refills) {
      cvh::LaunchCfg c = best;
      c.refill = v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : block_waves) {
      cvh::LaunchCfg c = best;
      c.block_waves = v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : intervals) {
      cvh::LaunchCfg c = best;
      c.refill_interval = v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : leaf_batches) {
      cvh::LaunchCfg c = best;
      c.leaf_batch = v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : sort_bits) {
      cvh::LaunchCfg c = best;
      c.sort_bits = v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : parts) {
      cvh::LaunchCfg c = best;
      c.parts = v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
    for (uint32_t v : widths) {
      if (!sweep_width || (v != 2 && !wide_ok)) continue;
      cvh::LaunchCfg c = best;
      c.width = v;
      double ms;
      if ((cs = time_cfg(c, &ms)) != CV_OK) return cs;
      if (ms < best_ms) best = c, best_ms = ms;
    }
  }
  b->cfg = best;
  if (chosen) *chosen = cvh::tuning_of(best);
  if (chosen_ms) *chosen_ms = best_ms;
  return CV_OK;
}

cv_status cv_batch_get_stats(const cv_batch* b, cv_render_stats* stats) {
  if (!b || !stats) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *stats = b->stats;
  return CV_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ server

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct ServeReq {
  uint64_t ticket = 0;
  cv_request req{};
  uint32_t frames_done = 0;
  DevBuf<float> hdr, feat;
  // buffers of the further frames a step renders for this request (zero
  // between steps); slot_bytes of them count against the budget
  DevBuf<float> slots;
  size_t bytes = 0, slot_bytes = 0;
  double t_submit = 0, t_done = 0;
  bool done = false;
  cv_status status = CV_OK;
  std::string error;
};

size_t scene_bytes(const cv_scene* s) {
  return (size_t)s->info.num_nodes * sizeof(cvk::DevNode) +
         (size_t)s->ntris * (sizeof(cvk::DevTri) + 4 + 24 + 4 + 4) +
         (size_t)s->ndirect * sizeof(cvk::DevLight) +
         (s->wide[0].n + s->wide[1].n) * sizeof(float4);
}

size_t request_bytes(const cv_request& r) {
  const size_t npix = (size_t)r.width * r.height;
  return npix * 12 + (r.denoise ? npix * (32 + 40 + 12 + 12) : 0);
}

}  // namespace

struct cv_server {
  uint64_t budget = 0, max_paths = 0;
  cv_render_opts opts{};
  cvh::LaunchCfg cfg;
  cv_denoiser* denoiser = nullptr;
  std::vector<cv_scene*> scenes;
  size_t fixed_bytes = 0;  // scenes + queues
  size_t used_bytes = 0;   // admitted requests
  size_t peak_bytes = 0;
  Pipeline pipe;
  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  std::deque<std::shared_ptr<ServeReq>> pending;
  std::vector<std::shared_ptr<ServeReq>> active;
  std::unordered_map<uint64_t, std::shared_ptr<ServeReq>> tickets;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  // requests whose frames are done, denoised and delivered by the finisher
  // thread while the worker runs the next steps
  std::deque<std::shared_ptr<ServeReq>> finishing;
  std::condition_variable fin_cv;
  bool fin_stop = false;
  std::thread finisher;
  hipStream_t fin_stream = nullptr;
  // held by the finisher around its HIP calls and by the worker around
  // graph capture, which another thread's HIP call would invalidate
  std::mutex hip_mu;
  // device buffers of released requests, freed by the worker between steps
  // (hipFree waits for the device, which the finisher should not)
  std::vector<void*> to_free;
  // statistics
  uint64_t submitted = 0, completed = 0, steps = 0;
  double busy_ms = 0;
  std::vector<double> latencies;
};

namespace {

// Denoise (optional) and deliver one finished request; called by the
// finisher thread without the server lock, holding hip_mu.
cv_status finish_request(cv_server* sv, cv_denoiser* dn, ServeReq& r) {
  const cv_request& q = r.req;
  const size_t npix = (size_t)q.width * q.height;
  const bool dev = q.flags & CV_PTR_DEVICE;
  if (q.denoise && dn)
    return cvh::denoise_image(dn, r.hdr.p, r.feat.p, q.width, q.height, q.spp, q.hdr, dev);
  HIPTRY(hipMemcpyAsync(q.hdr, r.hdr.p, npix * 12, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        sv->fin_stream),
         "deliver result");
  HIPTRY(hipStreamSynchronize(sv->fin_stream), "deliver result");
  return CV_OK;
}

void serve_loop(cv_server* sv) {
  for (;;) {
    std::vector<std::shared_ptr<ServeReq>> act;
    std::vector<void*> frees;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      // a request that could never fit fails; one that does not fit now
      // waits for active requests to finish and finished ones to be released
      auto never_fits = [&](const ServeReq& r) {
        return sv->fixed_bytes + r.bytes > sv->budget ||
               (size_t)r.req.width * r.req.height > sv->max_paths;
      };
      auto fits_now = [&](const ServeReq& r, size_t pixels) {
        return sv->fixed_bytes + sv->used_bytes + r.bytes <= sv->budget &&
               pixels + (size_t)r.req.width * r.req.height <= sv->max_paths;
      };
      sv->work_cv.wait(lk, [&] {
        if (!sv->active.empty() || !sv->to_free.empty()) return true;
        if (sv->pending.empty()) return sv->stop;
        const ServeReq& f = *sv->pending.front();
        return never_fits(f) || fits_now(f, 0);
      });
      frees.swap(sv->to_free);
      if (frees.empty() && sv->stop && sv->pending.empty() && sv->active.empty()) return;
    }
    for (void* p : frees) (void)hipFree(p);
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      auto never_fits = [&](const ServeReq& r) {
        return sv->fixed_bytes + r.bytes > sv->budget ||
               (size_t)r.req.width * r.req.height > sv->max_paths;
      };
      auto fits_now = [&](const ServeReq& r, size_t pixels) {
        return sv->fixed_bytes + sv->used_bytes + r.bytes <= sv->budget &&
               pixels + (size_t)r.req.width * r.req.height <= sv->max_paths;
      };
      // admission in arrival order: memory budget and queue capacity
      size_t pixels = 0;
      for (auto& r : sv->active) pixels += (size_t)r->req.width * r->req.height;
      while (!sv->pending.empty()) {
        auto r = sv->pending.front();
        const size_t npix = (size_t)r->req.width * r->req.height;
        const bool never = never_fits(*r);
        if (!never && !fits_now(*r, pixels)) break;
        sv->pending.pop_front();
        if (never || r->hdr.alloc(npix * 3) != hipSuccess ||
            hipMemset(r->hdr.p, 0, npix * 12) != hipSuccess ||
            (r->req.denoise && (r->feat.alloc(npix * 8) != hipSuccess ||
                                hipMemset(r->feat.p, 0, npix * 32) != hipSuccess))) {
          r->status = never ? CV_ERR_OUT_OF_MEMORY : CV_ERR_DEVICE;
          r->error = never ? "request exceeds the memory budget or the step capacity"
                           : "request buffers";
          r->hdr.reset();
          r->feat.reset();
          r->done = true;
          r->t_done = now_ms();
          sv->done_cv.notify_all();
          continue;
        }
        sv->used_bytes += r->bytes;
        sv->peak_bytes = std::max(sv->peak_bytes, sv->fixed_bytes + sv->used_bytes);
        pixels += npix;
        sv->active.push_back(r);
      }
      act = sv->active;
    }
    if (act.empty()) continue;
    // one step: the next frame of every active request, then further
    // frames (frame-major, in admission order) while the step's path
    // capacity lasts, so that few requests still fill the GPU.  A further
    // frame accumulates into the request's own zeroed buffer (and, for a
    // denoised request, its first-hit features into another) and is folded
    // into its image (features) in frame order after the step.
    std::vector<uint32_t> take(act.size(), 1);
    {
      size_t paths = 0;
      for (auto& r : act) paths += (size_t)r->req.width * r->req.height;
      for (bool more = true; more;) {
        more = false;
        for (size_t k = 0; k < act.size(); k++) {
          const ServeReq& r = *act[k];
          const size_t npix = (size_t)r.req.width * r.req.height;
          if (r.frames_done + take[k] >= r.req.spp || take[k] >= 64 ||
              paths + npix > sv->max_paths)
            continue;
          take[k]++;
          paths += npix;
          more = true;
        }
      }
      // the further frames' buffers, within the memory budget
      std::lock_guard<std::mutex> lk(sv->mu);
      for (size_t k = 0; k < act.size(); k++) {
        ServeReq& r = *act[k];
        // slots: the further frames' HDR buffers, then their features
        const size_t npix = (size_t)r.req.width * r.req.height;
        const size_t frame_floats = npix * 3 + (r.req.denoise ? npix * 8 : 0);
        const size_t need = (size_t)(take[k] - 1) * frame_floats;
        if (need <= r.slots.n) continue;
        const size_t extra = (need - r.slots.n) * sizeof(float);
        if (sv->fixed_bytes + sv->used_bytes + extra > sv->budget ||
            r.slots.alloc(need) != hipSuccess ||
            hipMemset(r.slots.p, 0, need * sizeof(float)) != hipSuccess) {
          sv->used_bytes -= r.slot_bytes;
          r.slots.reset();
          r.slot_bytes = 0;
          (void)hipGetLastError();
          take[k] = 1;
          continue;
        }
        sv->used_bytes += need * sizeof(float) - r.slot_bytes;
        r.slot_bytes = need * sizeof(float);
        sv->peak_bytes = std::max(sv->peak_bytes, sv->fixed_bytes + sv->used_bytes);
      }
    }
    StepPlan step;
    uint32_t max_depth = 0;
    bool packed = true;
    // the tree width of this step (every resident scene holds that tree:
    // built when the scene is added or the width set)
    uint32_t width = 2;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      width = sv->cfg.width;
    }
    for (size_t k = 0; k < act.size(); k++) {
      const ServeReq& r = *act[k];
      const cv_request& q = r.req;
      cv_scene* sc = sv->scenes[q.scene];
      cvk::JobDesc d = cvh::job_for_scene(sc, width);
      memcpy(d.cam.pos, q.camera_pos, 12);
      memset(d.cam.dpos, 0, 12);
      memcpy(d.cam.q, q.camera_q, 16);
      memset(d.cam.dq, 0, 16);
      if (q.screen_distance > 0.0f) d.cam.screen_distance = q.screen_distance;
      d.width = q.width;
      d.height = q.height;
      d.rect_w = q.width;
      d.rect_h = q.height;
      const size_t npix = (size_t)q.width * q.height;
      const size_t frame_floats = npix * 3, feat_floats = npix * 8;
      // the feature slots follow the HDR slots of this step's frames
      float* const fslots = r.slots.p + (size_t)(take[k] - 1) * frame_floats;
      for (uint32_t f = 0; f < take[k]; f++) {
        d.frame = 1 + r.frames_done + f;
        d.hdr = f == 0 ? r.hdr.p : r.slots.p + (size_t)(f - 1) * frame_floats;
        d.features = !q.denoise ? nullptr
                     : f == 0   ? r.feat.p
                                : fslots + (size_t)(f - 1) * feat_floats;
        d.path_base = step.paths;
        step.jobs.push_back(d);
        step.paths += q.width * q.height;
      }
      if (take[k] > 1) {
        FoldDesc fd;
        fd.hdr = r.hdr.p;
        fd.slots = r.slots.p;
        fd.stride = frame_floats;
        fd.count = take[k] - 1;
        fd.width = q.width;
        fd.rect[2] = q.width;
        fd.rect[3] = q.height;
        if (q.denoise) {
          fd.feat = r.feat.p;
          fd.fslots = fslots;
          fd.fstride = feat_floats;
          fd.fnpix = (uint32_t)npix;
        }
        step.folds.push_back(fd);
      }
      max_depth = std::max(max_depth, cvh::stack_depth(sc, width));
      packed = packed && sc->packed;
    }
    RunOpts o;
    o.cfg = sv->cfg;
    o.cfg.width = width;
    o.max_bounces = sv->opts.max_bounces;
    o.packed = packed;
    o.multi_job = true;
    {
      // one scene for the whole step: the shared-scene traversal and shading
      const uint32_t s0 = act[0]->req.scene;
      bool one = true;
      for (auto& r : act) one = one && r->req.scene == s0;
      if (one) {
        o.nodes = cvh::traversal_nodes(sv->scenes[s0], width);
        o.tris = sv->scenes[s0]->tris.p;
      }
    }
    o.max_depth = max_depth;
    o.graph = true;
    o.capture_mu = &sv->hip_mu;
    std::vector<StepPlan> steps(1, std::move(step));
    cv_render_stats stt{};
    cv_status st = pipeline_run(sv->pipe, steps, o, &stt);
    std::string err = st != CV_OK ? cv_last_error() : "";
    std::vector<std::shared_ptr<ServeReq>> finished;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->steps++;
      sv->busy_ms += stt.total_ms;
      for (size_t k = 0; k < act.size(); k++) {
        auto& r = act[k];
        r->frames_done += take[k];
        if (st != CV_OK || r->frames_done >= r->req.spp) finished.push_back(r);
      }
      for (auto& r : finished)
        sv->active.erase(std::find(sv->active.begin(), sv->active.end(), r));
    }
    if (finished.empty()) continue;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (auto& r : finished) {
        if (st != CV_OK) {
          r->status = st;
          r->error = err;
        }
        sv->finishing.push_back(r);
      }
    }
    sv->fin_cv.notify_all();
  }
}

// Finisher thread: denoises and delivers finished requests (on the
// denoiser's stream and synchronous copies) while the worker's next steps
// run, then releases their memory and wakes their waiters.
void finish_loop(cv_server* sv) {
  for (;;) {
    std::shared_ptr<ServeReq> r;
    cv_denoiser* dn = nullptr;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->fin_cv.wait(lk, [&] { return sv->fin_stop || !sv->finishing.empty(); });
      if (sv->finishing.empty()) return;
      r = sv->finishing.front();
      sv->finishing.pop_front();
      dn = sv->denoiser;
    }
    if (r->status == CV_OK) {
      std::lock_guard<std::mutex> hl(sv->hip_mu);
      const cv_status st = finish_request(sv, dn, *r);
      if (st != CV_OK) {
        r->status = st;
        r->error = cv_last_error();
      }
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (DevBuf<float>* b : {&r->hdr, &r->feat, &r->slots}) {
        if (b->p) sv->to_free.push_back(b->p);
        b->p = nullptr;
        b->n = 0;
      }
      r->done = true;
      r->t_done = now_ms();
      sv->used_bytes -= r->bytes + r->slot_bytes;
      r->slot_bytes = 0;
      if (r->status == CV_OK) {
        sv->completed++;
        sv->latencies.push_back(r->t_done - r->t_submit);
      }
    }
    sv->done_cv.notify_all();
    sv->work_cv.notify_all();  // freed memory may admit waiting requests
  }
}

}  // namespace

extern "C" {

cv_status cv_server_create(const cv_render_opts* opts, uint64_t hbm_budget_bytes,
                           uint64_t max_paths_per_step, cv_server** out) {
  if (!out) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  cv_status cs = cvh::ensure_device();
  if (cs != CV_OK) return cs;
  cv_render_opts o;
  if ((cs = check_opts(opts, &o, false)) != CV_OK) return cs;
  if (max_paths_per_step == 0) max_paths_per_step = 1ull << 23;
  if (max_paths_per_step >= (1ull << 31))
    return fail(CV_ERR_INVALID_ARGUMENT, "max_paths_per_step must be below 2^31");
  cv_server* sv = new cv_server();
  sv->opts = o;
  sv->budget = hbm_budget_bytes ? hbm_budget_bytes : (200ull << 30);
  sv->max_paths = max_paths_per_step;
  cv_tuning t{8, 32, 16, 1, 1};
  cvh::apply_tuning(sv->cfg, t);
  if ((cs = sv->pipe.init()) != CV_OK || (cs = sv->pipe.reserve(sv->max_paths, 256)) != CV_OK) {
    delete sv;
    return cs;
  }
  sv->fixed_bytes = sv->max_paths * (3 * 32 + 2 * 8 + 16);
  if (sv->fixed_bytes > sv->budget) {
    delete sv;
    return fail(CV_ERR_OUT_OF_MEMORY, "the path queues alone exceed the memory budget");
  }
  hipError_t e = hipStreamCreateWithFlags(&sv->fin_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete sv;
    return hip_fail(e, "hipStreamCreate");
  }
  sv->worker = std::thread(serve_loop, sv);
  sv->finisher = std::thread(finish_loop, sv);
  *out = sv;
  return CV_OK;
}

void cv_server_destroy(cv_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->work_cv.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->fin_stop = true;
  }
  sv->fin_cv.notify_all();
  if (sv->finisher.joinable()) sv->finisher.join();
  for (void* p : sv->to_free) (void)hipFree(p);
  if (sv->fin_stream) (void)hipStreamDestroy(sv->fin_stream);
  delete sv;
}

cv_status cv_server_add_scene(cv_server* sv, cv_scene* scene, uint32_t* id) {
  if (!sv || !scene || !id) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  if (scene->ndirect < 8)
    return fail(CV_ERR_LIGHT_COUNT, "the scene has fewer than 8 direct-sample triangles");
  std::lock_guard<std::mutex> lk(sv->mu);
  cv_status ws = cvh::ensure_wide(scene, sv->cfg.width);
  if (ws != CV_OK) return ws;
  const size_t b = scene_bytes(scene);
  if (sv->fixed_bytes + sv->used_bytes + b > sv->budget)
    return fail(CV_ERR_OUT_OF_MEMORY, "the scene does not fit the memory budget");
  sv->fixed_bytes += b;
  sv->scenes.push_back(scene);
  *id = (uint32_t)(sv->scenes.size() - 1);
  return CV_OK;
}

cv_status cv_server_set_denoiser(cv_server* sv, cv_denoiser* d) {
  if (!sv) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->denoiser = d;
  return CV_OK;
}

cv_status cv_server_set_tuning(cv_server* sv, const cv_tuning* t) {
  if (!sv || !t) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  cv_status v = cvh::validate_tuning(*t);
  if (v != CV_OK) return v;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (t->bvh_width)
    for (cv_scene* sc : sv->scenes) {
      const size_t before = scene_bytes(sc);
      cv_status ws = cvh::ensure_wide(sc, t->bvh_width);
      if (ws != CV_OK) return ws;
      sv->fixed_bytes += scene_bytes(sc) - before;
    }
  cvh::apply_tuning(sv->cfg, *t);
  return CV_OK;
}

cv_status cv_server_submit(cv_server* sv, const cv_request* req, uint64_t* ticket) {
  if (!sv || !req || !ticket || !req->hdr) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  if (req->scene >= sv->scenes.size()) return fail(CV_ERR_OUT_OF_RANGE, "unknown scene");
  cv_status cs = check_frame(sv->scenes[req->scene], req->width, req->height);
  if (cs != CV_OK) return cs;
  if (req->spp == 0) return fail(CV_ERR_INVALID_ARGUMENT, "spp must be at least 1");
  if (req->denoise) {
    if (!sv->denoiser) return fail(CV_ERR_INVALID_ARGUMENT, "no denoiser set");
    if (req->width % 64) return fail(CV_ERR_INVALID_ARGUMENT, "denoised widths are multiples of 64");
  }
  if ((size_t)req->width * req->height > sv->max_paths)
    return fail(CV_ERR_OUT_OF_RANGE, "request larger than the step capacity");
  auto r = std::make_shared<ServeReq>();
  r->ticket = sv->next_ticket++;
  r->req = *req;
  r->bytes = request_bytes(*req);
  r->t_submit = now_ms();
  sv->pending.push_back(r);
  sv->tickets[r->ticket] = r;
  sv->submitted++;
  *ticket = r->ticket;
  lk.unlock();
  sv->work_cv.notify_all();
  return CV_OK;
}

cv_status cv_server_poll(cv_server* sv, uint64_t ticket, int* done) {
  if (!sv || !done) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->tickets.find(ticket);
  if (it == sv->tickets.end()) return fail(CV_ERR_OUT_OF_RANGE, "unknown ticket");
  *done = it->second->done ? 1 : 0;
  return CV_OK;
}

cv_status cv_server_wait(cv_server* sv, uint64_t ticket, double* latency_ms) {
  if (!sv) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->tickets.find(ticket);
  if (it == sv->tickets.end()) return fail(CV_ERR_OUT_OF_RANGE, "unknown ticket");
  auto r = it->second;
  sv->done_cv.wait(lk, [&] { return r->done; });
  sv->tickets.erase(ticket);
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  if (r->status != CV_OK) return fail(r->status, r->error);
  return CV_OK;
}

cv_status cv_server_get_stats(cv_server* sv, cv_server_stats* out) {
  if (!sv || !out) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  cv_server_stats s{};
  s.submitted = sv->submitted;
  s.completed = sv->completed;
  s.steps = sv->steps;
  s.busy_ms = sv->busy_ms;
  s.peak_bytes = sv->peak_bytes;
  s.resident_bytes = sv->fixed_bytes;
  std::vector<double> l = sv->latencies;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      size_t k = (size_t)std::min<double>((double)l.size() - 1, q * (double)(l.size() - 1) + 0.5);
      return l[k];
    };
    s.p50_ms = pct(0.50);
    s.p99_ms = pct(0.99);
    double sum = 0;
    for (double v : l) sum += v;
    s.mean_ms = sum / (double)l.size();
  }
  *out = s;
  return CV_OK;
}

cv_status cv_server_reset_stats(cv_server* sv) {
  if (!sv) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->submitted = sv->completed = sv->steps = 0;
  sv->busy_ms = 0;
  sv->latencies.clear();
  sv->peak_bytes = sv->fixed_bytes + sv->used_bytes;
  return CV_OK;
}

}  // extern "C"

// -------------------------------------------------------------------- trace

extern "C" {

cv_status cv_trace_enable(size_t limit) {
  cv_status cs = cvh::ensure_device();
  if (cs != CV_OK) return cs;
  if (limit == 0) return fail(CV_ERR_INVALID_ARGUMENT, "limit must be positive");
  std::lock_guard<std::mutex> lk(g_trace_mu);
  if (!g_trace_epoch) HIPTRY(hipEventCreate(&g_trace_epoch), "event");
  HIPTRY(hipEventRecord(g_trace_epoch, nullptr), "event");
  HIPTRY(hipEventSynchronize(g_trace_epoch), "event");
  g_trace.clear();
  g_trace_limit = limit;
  g_trace_on = true;
  return CV_OK;
}

cv_status cv_trace_disable(void) {
  std::lock_guard<std::mutex> lk(g_trace_mu);
  g_trace_on = false;
  return CV_OK;
}

cv_status cv_trace_export(const char* path) {
  if (!path) return fail(CV_ERR_INVALID_ARGUMENT, "null path");
  std::lock_guard<std::mutex> lk(g_trace_mu);
  FILE* f = fopen(path, "w");
  if (!f) return fail(CV_ERR_IO, std::string("cannot open ") + path);
  static const char* names[4] = {"generate", "extend", "shade", "sort"};
  fputs("{\"traceEvents\":[", f);
  bool first = true;
  for (const TraceSpan& e : g_trace) {
    fprintf(f,
            "%s{\"name\":\"%s\",\"cat\":\"\",\"ph\":\"X\",\"ts\":%.3f,\"dur\":%.3f,\"pid\":0,"
            "\"tid\":%d,\"args\":{\"frame\":%u,\"id\":%u}}",
            first ? "" : ",", names[e.kind], e.ts_us, e.dur_us, e.kind, e.frame, e.id);
    first = false;
  }
  fputs("],\"displayTimeUnit\":\"ms\"}", f);
  const bool ok = fclose(f) == 0;
  return ok ? CV_OK : fail(CV_ERR_IO, std::string("cannot write ") + path);
}

}  // extern "C"
