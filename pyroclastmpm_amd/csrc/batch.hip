// batch.hip -- many independent scenes stepped together.
//
// Each scene is one pmpm_step (its own context, grid, buffers). A batch owns a
// few streams; a batched step forks them from the caller's stream with an
// event, places every scene on one of them by a longest-processing-time split
// of the scenes' weights (particle counts, or measured step times after
// autotuning), and joins them back with events. Small scenes then run side by
// side and fill the GPU that one of them alone would leave mostly idle. The
// same fork/join sequence can be captured into a HIP graph (one per buffer
// parity) and replayed, which removes the host's per-launch cost (about 25
// launches per scene and step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

#include "context.hpp"

using pmpm::hip_ok;

struct pmpm_batch {
  int device = 0;
  int streams = 1;               // streams owned (lanes available)
  hipStream_t lane[PMPM_MAX_LANES] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[PMPM_MAX_LANES] = {};
  pmpm_batch_tuning tuning{1, 0, 0.f};
  std::vector<float> weight;     // per scene (empty: particle counts)
  hipGraphExec_t graph[2] = {nullptr, nullptr};
};

namespace {

void destroy_graphs(pmpm_batch *b) {
  for (auto &g : b->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

int scene_ok(const pmpm_scene &s) {
  return s.ctx && s.set[0].n == s.set[1].n && s.set[0].x && s.set[1].x && s.set[0].x != s.set[1].x;
}

// Per-scene weights for the lane split: the batch's own (measured step
// times, after autotuning) when it holds one per scene, else particle counts.
template <class Count>
std::vector<float> weights_of(const pmpm_batch *b, int count, Count particles) {
  std::vector<float> w(count);
  for (int i = 0; i < count; ++i)
    w[i] = (int)b->weight.size() == count ? b->weight[i] : (float)particles(i);
  return w;
}

// Forks `lanes` streams from `origin` (lane 0 is `origin` itself), enqueues
// scene i with enqueue(i, stream of its lane), and joins the lanes back into
// `origin` with events, also when an enqueue failed (the first failure is
// returned and the scenes after it are not enqueued).
template <class Enqueue>
int fork_join(pmpm_batch *b, const std::vector<float> &w, int lanes, hipStream_t origin, Enqueue enqueue) {
  const int count = (int)w.size();
  std::vector<int32_t> lane_of(count);
  pmpm_batch_assign(w.data(), count, lanes, lane_of.data());
  hipStream_t s[PMPM_MAX_LANES];
  s[0] = origin;
  for (int l = 1; l < lanes; ++l) s[l] = b->lane[l];
  if (lanes > 1) {
    PMPM_HIP_CHECK(hipEventRecord(b->fork, origin));
    for (int l = 1; l < lanes; ++l) PMPM_HIP_CHECK(hipStreamWaitEvent(s[l], b->fork, 0));
  }
  int rc = PMPM_OK;
  for (int i = 0; i < count && rc == PMPM_OK; ++i) rc = enqueue(i, s[lane_of[i]]);
  if (lanes > 1) {
    for (int l = 1; l < lanes; ++l) {
      PMPM_HIP_CHECK(hipEventRecord(b->join[l], s[l]));
      PMPM_HIP_CHECK(hipStreamWaitEvent(origin, b->join[l], 0));
    }
  }
  return rc;
}

// Enqueues one step of every scene on the batch's lanes.
int enqueue_step(pmpm_batch *b, const pmpm_scene *scenes, int count, int cur, int lanes, hipStream_t origin) {
  const std::vector<float> w = weights_of(b, count, [&](int i) { return scenes[i].set[0].n; });
  return fork_join(b, w, lanes, origin, [&](int i, hipStream_t s) {
    const pmpm_scene &sc = scenes[i];
    pmpm_spatial part = sc.part;
    pmpm_nodes nodes = sc.nodes;
    pmpm_particles out = sc.set[cur ^ 1];
    return pmpm_step(sc.ctx, &sc.prm, &sc.grid, sc.musl, &sc.set[cur], &out, &part, &nodes, (pmpm_stream_t)s);
  });
}

int capture(pmpm_batch *b, const pmpm_scene *scenes, int count, int cur, hipStream_t origin) {
  PMPM_HIP_CHECK(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
  const int rc = enqueue_step(b, scenes, count, cur, b->tuning.lanes, origin);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(origin, &g);  // always end the capture
  if (rc != PMPM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (!hip_ok(e, "hipStreamEndCapture") || !g) return PMPM_ERR_HIP;
  if (b->graph[cur]) (void)hipGraphExecDestroy(b->graph[cur]);
  b->graph[cur] = nullptr;
  const hipError_t ei = hipGraphInstantiate(&b->graph[cur], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return hip_ok(ei, "hipGraphInstantiate") ? PMPM_OK : PMPM_ERR_HIP;
}

// Times `steps` batched steps of the current setting (capturing first when it
// replays graphs); returns ms per step or < 0 on error.
float time_steps(pmpm_batch *b, const pmpm_scene *scenes, int count, int cur, int steps, hipStream_t origin) {
  if (b->tuning.use_graph) {
    destroy_graphs(b);
    if (capture(b, scenes, count, cur, origin) != PMPM_OK) return -1.f;
    if (capture(b, scenes, count, cur ^ 1, origin) != PMPM_OK) return -1.f;
  }
  hipEvent_t t0, t1;
  if (!hip_ok(hipEventCreate(&t0), "hipEventCreate")) return -1.f;
  if (!hip_ok(hipEventCreate(&t1), "hipEventCreate")) { (void)hipEventDestroy(t0); return -1.f; }
  float ms = -1.f;
  int rc = hip_ok(hipEventRecord(t0, origin), "hipEventRecord") ? PMPM_OK : PMPM_ERR_HIP;
  for (int k = 0; k < steps && rc == PMPM_OK; ++k) {
    const int c = cur ^ (k & 1);
    rc = b->tuning.use_graph ? pmpm_batch_replay(b, c, (pmpm_stream_t)origin)
                             : enqueue_step(b, scenes, count, c, b->tuning.lanes, origin);
  }
  if (rc == PMPM_OK && hip_ok(hipEventRecord(t1, origin), "hipEventRecord") &&
      hip_ok(hipEventSynchronize(t1), "hipEventSynchronize") && hip_ok(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime"))
    ms /= (float)steps;
  else
    ms = -1.f;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return ms;
}

}  // namespace

extern "C" int pmpm_batch_create(pmpm_batch **batch, int32_t lanes) {
  if (!batch || lanes < 1 || lanes > PMPM_MAX_LANES) return PMPM_ERR_INVALID;
  pmpm_batch *b = new (std::nothrow) pmpm_batch();
  if (!b) return PMPM_ERR_INVALID;
  b->streams = lanes;
  b->tuning.lanes = lanes;
  int rc = hip_ok(hipGetDevice(&b->device), "hipGetDevice") ? PMPM_OK : PMPM_ERR_HIP;
  if (rc == PMPM_OK && !hip_ok(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreateWithFlags"))
    rc = PMPM_ERR_HIP;
  for (int l = 1; l < lanes && rc == PMPM_OK; ++l) {
    if (!hip_ok(hipStreamCreateWithFlags(&b->lane[l], hipStreamNonBlocking), "hipStreamCreateWithFlags"))
      rc = PMPM_ERR_HIP;
    else if (!hip_ok(hipEventCreateWithFlags(&b->join[l], hipEventDisableTiming), "hipEventCreateWithFlags"))
      rc = PMPM_ERR_HIP;
  }
  if (rc != PMPM_OK) {
    pmpm_batch_destroy(b);
    return rc;
  }
  *batch = b;
  return PMPM_OK;
}

extern "C" int pmpm_batch_destroy(pmpm_batch *b) {
  if (!b) return PMPM_OK;
  destroy_graphs(b);
  for (int l = 0; l < PMPM_MAX_LANES; ++l) {
    if (b->lane[l]) (void)hipStreamDestroy(b->lane[l]);
    if (b->join[l]) (void)hipEventDestroy(b->join[l]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  delete b;
  return PMPM_OK;
}

extern "C" int pmpm_batch_assign(const float *weight, int32_t count, int32_t lanes, int32_t *lane_of) {
  if (count < 0 || lanes < 1 || lanes > PMPM_MAX_LANES || (count > 0 && (!weight || !lane_of)))
    return PMPM_ERR_INVALID;
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return weight[a] > weight[b]; });
  double load[PMPM_MAX_LANES] = {};
  for (int i : order) {
    int best = 0;
    for (int l = 1; l < lanes; ++l)
      if (load[l] < load[best]) best = l;
    lane_of[i] = best;
    load[best] += weight[i];
  }
  return PMPM_OK;
}

extern "C" int pmpm_batch_set_weights(pmpm_batch *b, const float *weight, int32_t count) {
  if (!b || count < 0 || (count > 0 && !weight)) return PMPM_ERR_INVALID;
  b->weight.assign(weight, weight + count);
  destroy_graphs(b);  // the placement changed
  return PMPM_OK;
}

extern "C" int pmpm_batch_step(pmpm_batch *b, const pmpm_scene *scenes, int32_t count, int32_t cur,
                               pmpm_stream_t stream) {
  if (!b || count < 0 || (count > 0 && !scenes) || (cur != 0 && cur != 1)) return PMPM_ERR_INVALID;
  for (int i = 0; i < count; ++i)
    if (!scene_ok(scenes[i])) return PMPM_ERR_INVALID;
  return enqueue_step(b, scenes, count, cur, b->tuning.lanes, (hipStream_t)stream);
}

extern "C" int pmpm_batch_capture(pmpm_batch *b, const pmpm_scene *scenes, int32_t count, int32_t cur,
                                  pmpm_stream_t stream) {
  if (!b || count < 0 || (count > 0 && !scenes) || (cur != 0 && cur != 1)) return PMPM_ERR_INVALID;
  for (int i = 0; i < count; ++i)
    if (!scene_ok(scenes[i])) return PMPM_ERR_INVALID;
  return capture(b, scenes, count, cur, (hipStream_t)stream);
}

extern "C" int pmpm_batch_replay(pmpm_batch *b, int32_t cur, pmpm_stream_t stream) {
  if (!b || (cur != 0 && cur != 1) || !b->graph[cur]) return PMPM_ERR_INVALID;
  PMPM_HIP_CHECK(hipGraphLaunch(b->graph[cur], (hipStream_t)stream));
  return PMPM_OK;
}

extern "C" int pmpm_batch_adjoint_step(pmpm_batch *b, const pmpm_adjoint_scene *scenes, int32_t count,
                                       pmpm_stream_t stream) {
  if (!b || count < 0 || (count > 0 && !scenes)) return PMPM_ERR_INVALID;
  for (int i = 0; i < count; ++i)
    if (!scenes[i].ctx || !scenes[i].param_bar) return PMPM_ERR_INVALID;
  const std::vector<float> w = weights_of(b, count, [&](int i) { return scenes[i].state.n; });
  return fork_join(b, w, b->tuning.lanes, (hipStream_t)stream, [&](int i, hipStream_t s) {
    const pmpm_adjoint_scene &sc = scenes[i];
    pmpm_adjoint in_bar = sc.in_bar;
    pmpm_spatial part = sc.part;
    pmpm_nodes nodes = sc.nodes;
    return pmpm_adjoint_step(sc.ctx, &sc.prm, &sc.grid, &sc.state, &sc.out_bar, &in_bar, sc.param_bar,
                             sc.recompute, &part, &nodes, (pmpm_stream_t)s);
  });
}

extern "C" int pmpm_batch_get_tuning(const pmpm_batch *b, pmpm_batch_tuning *out) {
  if (!b || !out) return PMPM_ERR_INVALID;
  *out = b->tuning;
  return PMPM_OK;
}

extern "C" int pmpm_batch_autotune(pmpm_batch *b, const pmpm_scene *scenes, int32_t count, int32_t cur,
                                   int32_t steps, pmpm_stream_t stream, pmpm_batch_tuning *out) {
  if (!b || count < 1 || !scenes || (cur != 0 && cur != 1) || steps < 1) return PMPM_ERR_INVALID;
  for (int i = 0; i < count; ++i)
    if (!scene_ok(scenes[i])) return PMPM_ERR_INVALID;
  hipStream_t origin = (hipStream_t)stream;
  steps += steps & 1;
  // 1. each scene's own step time (two steps, so its parity is unchanged):
  //    small scenes cost their launches, not their particles
  std::vector<float> w(count);
  hipEvent_t t0, t1;
  PMPM_HIP_CHECK(hipEventCreate(&t0));
  if (!hip_ok(hipEventCreate(&t1), "hipEventCreate")) { (void)hipEventDestroy(t0); return PMPM_ERR_HIP; }
  int rc = PMPM_OK;
  for (int i = 0; i < count && rc == PMPM_OK; ++i) {
    const pmpm_scene &sc = scenes[i];
    if (!hip_ok(hipEventRecord(t0, origin), "hipEventRecord")) { rc = PMPM_ERR_HIP; break; }
    for (int k = 0; k < 2 && rc == PMPM_OK; ++k) {
      const int c = cur ^ k;
      pmpm_spatial part = sc.part;
      pmpm_nodes nodes = sc.nodes;
      pmpm_particles o = sc.set[c ^ 1];
      rc = pmpm_step(sc.ctx, &sc.prm, &sc.grid, sc.musl, &sc.set[c], &o, &part, &nodes, stream);
    }
    float ms = 0.f;
    if (rc == PMPM_OK && (!hip_ok(hipEventRecord(t1, origin), "hipEventRecord") ||
                          !hip_ok(hipEventSynchronize(t1), "hipEventSynchronize") ||
                          !hip_ok(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime")))
      rc = PMPM_ERR_HIP;
    w[i] = ms * 0.5f;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc != PMPM_OK) return rc;
  b->weight = w;
  // 2. every lane count, eager and replayed
  pmpm_batch_tuning best{1, 0, -1.f};
  for (int lanes = 1; lanes <= b->streams; ++lanes)
    for (int graph = 0; graph < 2; ++graph) {
      b->tuning = pmpm_batch_tuning{lanes, graph, 0.f};
      const float ms = time_steps(b, scenes, count, cur, steps, origin);
      if (ms < 0.f) {
        destroy_graphs(b);
        return PMPM_ERR_HIP;
      }
      if (best.step_ms < 0.f || ms < best.step_ms) best = pmpm_batch_tuning{lanes, graph, ms};
    }
  destroy_graphs(b);  // captured for the last candidate; the caller captures for the kept one
  b->tuning = best;
  if (out) *out = best;
  return PMPM_OK;
}
