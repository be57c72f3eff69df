// Plan object of the persistent sites launch: it owns the site orderings, the per-block field
// workspace, the work counter and one tuned launch captured as a HIP graph.
//
// Replaces the captured, replayed batches of OpticalPlan.prepare_graph / evaluate_graph
// (maomi/backends/cuda/batched.py:103-146) for the sites launch: instead of a graph per
// batch, one graph holds the whole run's launch (every site, mixed field sides and depths),
// and the autotuner (maomi_sites_plan_tune) picks its grid, block shape -- which sets the
// LDS tile height of the line transforms: 512 threads transform 32-line tiles at N = 256,
// 256 threads 16-line tiles --, site ordering and schedule by timing captured candidates.
// The schedule matters both ways: the work counter balances sites of unequal cost (the mixed
// 4225-site launch: 103 -> 80 ms), while sites of one size run faster with static striding
// (the uniform 4225-site launch: 114 ms with the counter, 103 ms without).  The caller
// keeps owning every buffer the descriptor names; the plan reads them at each replay.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "maomi_amd.h"

struct maomi_sites_plan_t {
  maomi_sites_desc desc{};          // the caller's descriptor (its buffers stay caller-owned)
  int64_t P = 0, nmax = 0;
  int32_t* orders[3] = {};          // device orderings, indexed by MAOMI_ORDER_*
  void* work = nullptr;             // grid * nmax^2 complex64
  int64_t work_grid = 0;
  int32_t* queue = nullptr;         // dynamic site counter (reset inside the graph)
  hipStream_t capture = nullptr;    // stream the launch is captured on (default priority)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  maomi_sites_choice choice{};
};

namespace {

void drop_graph(hipGraph_t& g, hipGraphExec_t& e) {
  if (e) (void)hipGraphExecDestroy(e);
  if (g) (void)hipGraphDestroy(g);
  e = nullptr;
  g = nullptr;
}

int fail(const char* what, hipError_t e) {
  maomi_set_error("%s: %s", what, hipGetErrorString(e));
  return (int)e;
}

// Grow the workspace to `grid` blocks; graphs captured on the old one must be dropped first.
int ensure_work(maomi_sites_plan_t* p, int64_t grid) {
  if (p->work && p->work_grid >= grid) return MAOMI_OK;
  if (p->work) {
    (void)hipDeviceSynchronize();  // no replay may still read the old workspace
    (void)hipFree(p->work);
  }
  p->work = nullptr;
  p->work_grid = 0;
  hipError_t e = hipMalloc(&p->work, (size_t)maomi_sites_workspace_bytes(grid, p->nmax));
  if (e != hipSuccess) return fail("maomi_sites_plan: workspace allocation", e);
  p->work_grid = grid;
  return MAOMI_OK;
}

// Capture one launch of the given choice into (graph, exec).
int capture(maomi_sites_plan_t* p, int64_t grid, int64_t threads, int64_t order,
            int64_t schedule, hipGraph_t* graph, hipGraphExec_t* exec) {
  maomi_sites_desc d = p->desc;
  d.grid = grid;
  d.block_threads = threads;
  d.order = p->orders[order];
  d.workspace = p->work;
  d.scratch_stride = p->nmax * p->nmax;
  d.queue = schedule == MAOMI_SCHEDULE_QUEUE ? p->queue : nullptr;
  hipError_t e = hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail("maomi_sites_plan: begin capture", e);
  const int rc = maomi_persistent_sites(&d, p->capture);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(p->capture, &g);
  if (rc != MAOMI_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;  // message set by maomi_persistent_sites
  }
  if (e != hipSuccess) return fail("maomi_sites_plan: end capture", e);
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail("maomi_sites_plan: instantiate", e);
  }
  *graph = g;
  *exec = x;
  return MAOMI_OK;
}

}  // namespace

MAOMI_API int maomi_sites_plan_create(const maomi_sites_desc* d, maomi_sites_plan* out) {
  if (!d || !out) {
    maomi_set_error("maomi_sites_plan_create: null descriptor or output handle");
    return MAOMI_EINVAL;
  }
  *out = nullptr;
  if (d->P <= 0 || !d->site_n || !d->site_s) {
    maomi_set_error("maomi_sites_plan_create: need P > 0 and the site_n / site_s arrays");
    return MAOMI_EINVAL;
  }
  if (d->psf_out || d->phase16 || d->screen) {
    maomi_set_error("maomi_sites_plan_create: plans hold FP32-phase launches without "
                    "focal-plane output");
    return MAOMI_EUNSUPPORTED;
  }
  const int64_t P = d->P;
  std::vector<int32_t> n(P), s(P);
  hipError_t e = hipMemcpy(n.data(), d->site_n, P * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(s.data(), d->site_s, P * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail("maomi_sites_plan_create: reading site_n / site_s", e);
  // orderings: natural; longest first by cost S N^2 log2 N (ties keep the natural order);
  // field side first, then depth, both descending
  std::vector<int32_t> ord[3];
  for (auto& o : ord) {
    o.resize(P);
    std::iota(o.begin(), o.end(), 0);
  }
  for (int64_t i = 0; i < P; ++i) {
    const bool side_ok = n[i] == 64 || n[i] == 128 || n[i] == 256 || n[i] == 512;
    if (!side_ok || (d->uniform_n && n[i] != d->uniform_n) || s[i] < 0 || s[i] > d->Z) {
      maomi_set_error("maomi_sites_plan_create: site %lld has field side %d and %d screens "
                      "(sides 64/128/256/512, matching uniform_n if set; 0 <= S <= Z)",
                      (long long)i, n[i], s[i]);
      return MAOMI_EINVAL;
    }
  }
  std::vector<double> cost(P);
  int64_t nmax = 0;
  for (int64_t i = 0; i < P; ++i) {
    cost[i] = (double)s[i] * n[i] * n[i] * std::log2((double)std::max(n[i], 2));
    nmax = std::max<int64_t>(nmax, n[i]);
  }
  std::stable_sort(ord[MAOMI_ORDER_LPT].begin(), ord[MAOMI_ORDER_LPT].end(),
                   [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
  std::stable_sort(ord[MAOMI_ORDER_CLASS].begin(), ord[MAOMI_ORDER_CLASS].end(),
                   [&](int32_t a, int32_t b) { return n[a] != n[b] ? n[a] > n[b] : s[a] > s[b]; });
  auto* p = new maomi_sites_plan_t;
  p->desc = *d;
  p->P = P;
  p->nmax = nmax;
  for (int k = 0; k < 3 && e == hipSuccess; ++k) {
    e = hipMalloc(&p->orders[k], P * sizeof(int32_t));
    if (e == hipSuccess)
      e = hipMemcpy(p->orders[k], ord[k].data(), P * sizeof(int32_t), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMalloc(&p->queue, sizeof(int32_t));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking);
  if (e != hipSuccess) {
    const int rc = fail("maomi_sites_plan_create", e);
    maomi_sites_plan_destroy(p);
    return rc;
  }
  *out = p;
  return MAOMI_OK;
}

MAOMI_API int maomi_sites_plan_capture(maomi_sites_plan p, int64_t grid, int64_t threads,
                                       int64_t order, int64_t schedule) {
  if (!p || grid < 1 || (threads != 256 && threads != 512) || order < 0 || order > 2 ||
      (schedule != MAOMI_SCHEDULE_STATIC && schedule != MAOMI_SCHEDULE_QUEUE)) {
    maomi_set_error("maomi_sites_plan_capture: bad arguments (grid >= 1, threads 256 or 512, "
                    "order 0..2, schedule 0 or 1)");
    return MAOMI_EINVAL;
  }
  grid = std::min(grid, p->P);
  drop_graph(p->graph, p->exec);
  int rc = ensure_work(p, grid);
  if (rc == MAOMI_OK) rc = capture(p, grid, threads, order, schedule, &p->graph, &p->exec);
  if (rc == MAOMI_OK) p->choice = maomi_sites_choice{grid, threads, order, schedule, -1.0};
  return rc;
}

MAOMI_API int maomi_sites_plan_tune(maomi_sites_plan p, const int64_t* grids, int n_grids,
                                    const int64_t* threads, int n_threads, const int64_t* orders,
                                    int n_orders, const int64_t* schedules, int n_schedules,
                                    int reps, maomi_stream_t stream, maomi_sites_choice* results,
                                    maomi_sites_choice* best) {
  if (!p || !grids || !threads || !orders || !schedules || n_grids < 1 || n_threads < 1 ||
      n_orders < 1 || n_schedules < 1 || reps < 1) {
    maomi_set_error("maomi_sites_plan_tune: bad arguments");
    return MAOMI_EINVAL;
  }
  int64_t gmax = 0;
  for (int i = 0; i < n_grids; ++i) {
    if (grids[i] < 1) {
      maomi_set_error("maomi_sites_plan_tune: grid sizes must be >= 1");
      return MAOMI_EINVAL;
    }
    gmax = std::max(gmax, std::min(grids[i], p->P));
  }
  for (int i = 0; i < n_threads; ++i)
    if (threads[i] != 256 && threads[i] != 512) {
      maomi_set_error("maomi_sites_plan_tune: block_threads must be 256 or 512");
      return MAOMI_EINVAL;
    }
  for (int i = 0; i < n_orders; ++i)
    if (orders[i] < 0 || orders[i] > 2) {
      maomi_set_error("maomi_sites_plan_tune: orders must be 0, 1 or 2");
      return MAOMI_EINVAL;
    }
  for (int i = 0; i < n_schedules; ++i)
    if (schedules[i] != MAOMI_SCHEDULE_STATIC && schedules[i] != MAOMI_SCHEDULE_QUEUE) {
      maomi_set_error("maomi_sites_plan_tune: schedules must be 0 (static) or 1 (queue)");
      return MAOMI_EINVAL;
    }
  drop_graph(p->graph, p->exec);
  int rc = ensure_work(p, gmax);  // one workspace for every candidate
  if (rc != MAOMI_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t ev[2] = {};
  hipError_t e = hipEventCreate(&ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&ev[1]);
  if (e != hipSuccess) return fail("maomi_sites_plan_tune: events", e);
  int k = 0;
  maomi_sites_choice top{0, 0, 0, 0, -1.0};
  for (int gi = 0; gi < n_grids && rc == MAOMI_OK; ++gi)
    for (int ti = 0; ti < n_threads && rc == MAOMI_OK; ++ti)
      for (int oi = 0; oi < n_orders && rc == MAOMI_OK; ++oi)
        for (int si = 0; si < n_schedules && rc == MAOMI_OK; ++si, ++k) {
          const int64_t g = std::min(grids[gi], p->P);
          hipGraph_t graph = nullptr;
          hipGraphExec_t exec = nullptr;
          rc = capture(p, g, threads[ti], orders[oi], schedules[si], &graph, &exec);
          if (rc != MAOMI_OK) break;
          float best_ms = -1.0f;
          e = hipGraphLaunch(exec, st);  // warm-up (first use loads the code object)
          for (int r = 0; r < reps && e == hipSuccess; ++r) {
            e = hipEventRecord(ev[0], st);
            if (e == hipSuccess) e = hipGraphLaunch(exec, st);
            if (e == hipSuccess) e = hipEventRecord(ev[1], st);
            if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
            if (e == hipSuccess && (best_ms < 0.0f || ms < best_ms)) best_ms = ms;
          }
          if (e != hipSuccess) {
            drop_graph(graph, exec);
            rc = fail("maomi_sites_plan_tune: timed replay", e);
            break;
          }
          const maomi_sites_choice c{g, threads[ti], orders[oi], schedules[si], (double)best_ms};
          if (results) results[k] = c;
          if (top.ms < 0.0 || c.ms < top.ms) {
            drop_graph(p->graph, p->exec);
            p->graph = graph;
            p->exec = exec;
            top = c;
          } else {
            drop_graph(graph, exec);
          }
        }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc != MAOMI_OK) {
    drop_graph(p->graph, p->exec);
    return rc;
  }
  p->choice = top;
  if (best) *best = top;
  return MAOMI_OK;
}

MAOMI_API int maomi_sites_plan_replay(maomi_sites_plan p, maomi_stream_t stream) {
  if (!p || !p->exec) {
    maomi_set_error("maomi_sites_plan_replay: no captured launch (tune or capture first)");
    return MAOMI_EINVAL;
  }
  const hipError_t e = hipGraphLaunch(p->exec, (hipStream_t)stream);
  if (e != hipSuccess) return fail("maomi_sites_plan_replay", e);
  return MAOMI_OK;
}

MAOMI_API int maomi_sites_plan_choice(maomi_sites_plan p, maomi_sites_choice* out) {
  if (!p || !out || !p->exec) {
    maomi_set_error("maomi_sites_plan_choice: no captured launch");
    return MAOMI_EINVAL;
  }
  *out = p->choice;
  return MAOMI_OK;
}

MAOMI_API int maomi_sites_plan_destroy(maomi_sites_plan p) {
  if (!p) return MAOMI_OK;
  (void)hipDeviceSynchronize();  // replays queued on any stream finish before the frees
  drop_graph(p->graph, p->exec);
  for (auto& o : p->orders)
    if (o) (void)hipFree(o);
  if (p->work) (void)hipFree(p->work);
  if (p->queue) (void)hipFree(p->queue);
  if (p->capture) (void)hipStreamDestroy(p->capture);
  delete p;
  return MAOMI_OK;
}
