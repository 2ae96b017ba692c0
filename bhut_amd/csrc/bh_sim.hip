// bh_sim.hip -- kick-drift-kick leapfrog driver with graph-captured steps.
//
// bhut has no integrator of its own; its galaxy example rebuilds a Tree on
// every right-hand-side call of scipy's solve_ivp
// (examples/galaxy/galaxy_simulation.py:277-284).  This driver advances
// positions and velocities resident in device memory with KDK leapfrog,
// refitting the tree every step (bhut/tree/refit.py:17-109) and rebuilding it
// on a fixed schedule (Tree.rebuild, bhut/api.py:564-606; the topology is a
// function of N and leaf_size, so a rebuild re-sorts in place).
//
// One step, as enqueued (and captured into a HIP graph after its first run):
//   clear the status word; v += a dt/2, x += v dt; refit or rebuild;
//   a = F(x) (plain FP64 walk or FP32 grouped walk); v += a dt/2.
// The FP32 walk runs with device-scanned list offsets into lists of planned
// capacity.  A list that would overflow sets a status bit instead: its fill,
// its evaluation and the last kick are skipped, and the host redoes the
// force and the kick eagerly with grown capacity and recaptures the graph.
#include "bh_common.h"

namespace bh {

__global__ __launch_bounds__(256) void k_kick_drift(int64_t nd, double* __restrict__ vel,
                                                    double* __restrict__ pos, const double* __restrict__ acc,
                                                    double hdt, double dt) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nd) return;
  const double v = vel[i] + acc[i] * hdt;
  vel[i] = v;
  pos[i] = pos[i] + v * dt;
}

__global__ __launch_bounds__(256) void k_kick(int64_t nd, double* __restrict__ vel, const double* __restrict__ acc,
                                              double hdt, const uint32_t* __restrict__ status) {
  if (*status != 0u) return;  // the force of this step was not evaluated
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nd) return;
  vel[i] = vel[i] + acc[i] * hdt;
}

}  // namespace bh

using namespace bh;

struct bh_sim {
  bh_ctx* ctx = nullptr;
  bh_tree* tree = nullptr;
  double* pos = nullptr;  // caller's buffers, original particle order
  double* vel = nullptr;
  const double* mass = nullptr;
  double* acc = nullptr;  // owned
  int64_t n = 0;
  int dim = 3;
  double theta = 0.5, eps = 0.0, G = 1.0;
  int precision = 32;
  bh_f32_opts opts{};
  hipStream_t stream = nullptr;
  uint32_t* status = nullptr;  // device: status word, then need[2] (int64) at byte 8
  int64_t* need = nullptr;
  int64_t cap_region = 1024, cap_items = 4096;  // FP32 lists of captured steps: entries per group, work items
  // graphs of a refit step (0) and a rebuild step (1)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  double graph_dt[2] = {0.0, 0.0};
  std::vector<void*> retained[2];
  int64_t step_index = 0;
  int64_t steps = 0, rebuilds = 0, refits = 0, launches = 0, captures = 0, regrows = 0;
};

namespace {

// Enqueues the FP32 or FP64 force into s->acc on the context stream.
int enqueue_force(bh_sim* s) {
  bh_ctx* c = s->ctx;
  if (s->precision == 64)
    return walk_mode(c, s->tree, WALK_ACC, nullptr, s->n, s->theta, s->eps, s->G, s->acc, nullptr, nullptr, nullptr);
  F32Plan plan;
  plan.cap_region = s->cap_region;
  plan.cap_items = s->cap_items;
  plan.err = s->status;
  plan.need = s->need;
  return walk_f32(c, s->tree, nullptr, s->n, s->theta, s->eps, s->G, &s->opts, s->acc, nullptr, &plan);
}

int enqueue_step(bh_sim* s, int kind, double dt) {
  bh_ctx* c = s->ctx;
  const int64_t nd = s->n * s->dim;
  const unsigned nb = (unsigned)cdiv(nd, 256);
  BH_TRY(fill_u32(c, s->status, 0u, 1));
  hipLaunchKernelGGL(k_kick_drift, dim3(nb), dim3(256), 0, c->stream, nd, s->vel, s->pos, s->acc, 0.5 * dt, dt);
  BH_LAUNCH_CHECK();
  if (kind == 1) {
    BH_TRY(rebuild_in_place(c, s->tree, s->pos, s->mass));
  } else {
    BH_TRY(bh_tree_refit_dev(s->tree, s->pos, s->mass));
  }
  BH_TRY(enqueue_force(s));
  hipLaunchKernelGGL(k_kick, dim3(nb), dim3(256), 0, c->stream, nd, s->vel, s->acc, 0.5 * dt, s->status);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

void drop_graph(bh_sim* s, int kind) {
  if (s->exec[kind]) (void)hipGraphExecDestroy(s->exec[kind]);
  if (s->graph[kind]) (void)hipGraphDestroy(s->graph[kind]);
  s->exec[kind] = nullptr;
  s->graph[kind] = nullptr;
  for (void* p : s->retained[kind]) dfree(s->ctx, p);
  s->retained[kind].clear();
}

int capture(bh_sim* s, int kind, double dt) {
  bh_ctx* c = s->ctx;
  drop_graph(s, kind);
  BH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  c->retain = &s->retained[kind];
  c->capturing = true;
  int rc = enqueue_step(s, kind, dt);
  c->capturing = false;
  c->retain = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc == BH_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  if (rc == BH_OK) {
    e = hipGraphInstantiate(&s->exec[kind], g, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  }
  s->graph[kind] = g;
  if (rc != BH_OK) {
    drop_graph(s, kind);
    return rc;
  }
  s->graph_dt[kind] = dt;
  s->captures++;
  return BH_OK;
}

// Reads the status word; on a list or walk-stack overflow grows the
// capacity, redoes the force and the final kick eagerly and drops the
// captured graphs.
int settle(bh_sim* s, double dt) {
  bh_ctx* c = s->ctx;
  for (int attempt = 0;; attempt++) {
    uint64_t h[3] = {0, 0, 0};
    BH_HIP(hipMemcpyAsync(h, s->status, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    BH_HIP(hipStreamSynchronize(c->stream));
    const uint32_t st = (uint32_t)h[0];
    if (st == 0) return BH_OK;
    if (attempt >= 8) {
      set_error("interaction lists or the walk stack still overflow after regrowing");
      return BH_EOVERFLOW;
    }
    if (st & 1u) {
      // the LDS stack of the group walk: twice as many entries
      if (s->opts.stack_cap >= 8192) {
        set_error("group walk stack overflow (stack_cap=%d)", s->opts.stack_cap);
        return BH_EOVERFLOW;
      }
      s->opts.stack_cap = std::min(8192, 2 * s->opts.stack_cap);
    } else {
      // the longest list and the work items, with a quarter to spare
      const int64_t nl = (int64_t)h[1], ni = (int64_t)h[2];
      s->cap_region = std::max(s->cap_region, nl + nl / 4 + 1024);
      s->cap_items = std::max(s->cap_items, ni + ni / 4 + 1024);
    }
    s->regrows++;
    drop_graph(s, 0);
    drop_graph(s, 1);
    BH_TRY(fill_u32(c, s->status, 0u, 1));
    BH_TRY(enqueue_force(s));
    if (dt != 0.0) {
      const int64_t nd = s->n * s->dim;
      hipLaunchKernelGGL(k_kick, dim3((unsigned)cdiv(nd, 256)), dim3(256), 0, c->stream, nd, s->vel, s->acc,
                         0.5 * dt, s->status);
      BH_LAUNCH_CHECK();
    }
  }
}

// Runs the body with the simulation's own stream as the context stream.
struct StreamSwap {
  bh_ctx* c;
  hipStream_t prev;
  StreamSwap(bh_ctx* ctx, hipStream_t s) : c(ctx), prev(ctx->stream) { c->stream = s; }
  ~StreamSwap() { c->stream = prev; }
};

}  // namespace

extern "C" {

int bh_sim_create(bh_ctx* c, double* d_pos, double* d_vel, const double* d_mass, int64_t n, int32_t dim,
                  int32_t leaf_size, double theta, double softening, double G, int32_t precision,
                  const bh_f32_opts* opts, bh_sim** out) {
  BH_REQ(c && out && d_pos && d_vel && d_mass, "NULL argument");
  *out = nullptr;
  BH_REQ(precision == 32 || precision == 64, "precision must be 32 or 64, got %d", precision);
  BH_REQ(theta >= 0 && softening >= 0, "theta and softening must be non-negative");
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  if (hipSetDevice(c->device) != hipSuccess) {
    set_error("cannot select device %d", c->device);
    return BH_EHIP;
  }
  bh_sim* s = new bh_sim();
  s->ctx = c;
  s->pos = d_pos;
  s->vel = d_vel;
  s->mass = d_mass;
  s->n = n;
  s->dim = dim;
  s->theta = theta;
  s->eps = softening;
  s->G = G;
  s->precision = precision;
  if (opts) s->opts = *opts;
  else s->opts.kahan = 1;
  if (!s->opts.group_size) s->opts.group_size = 64;
  if (!s->opts.stack_cap) s->opts.stack_cap = 1024;
  if (!s->opts.multipole) s->opts.multipole = 2;
  int rc = bh_tree_build_dev(c, d_pos, d_mass, n, dim, leaf_size, &s->tree);
  // the simulation's own stream takes over from here
  if (rc == BH_OK && hipStreamSynchronize(c->stream) != hipSuccess)
    rc = hip_fail(hipGetLastError(), "build", __FILE__, __LINE__);
  if (rc == BH_OK && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    rc = hip_fail(hipGetLastError(), "hipStreamCreateWithFlags", __FILE__, __LINE__);
  if (rc == BH_OK) {
    s->acc = dalloc_t<double>(c, (size_t)n * dim);
    s->status = dalloc_t<uint32_t>(c, 6);  // word, pad, need[2]
    if (!s->acc || !s->status) rc = BH_ENOMEM;
    else s->need = reinterpret_cast<int64_t*>(s->status + 2);
  }
  if (rc == BH_OK) {
    // initial force; its list sizes set the first capacities
    StreamSwap sw(c, s->stream);
    rc = fill_u32(c, s->status, 0u, 1);
    if (rc == BH_OK) rc = enqueue_force(s);
    if (rc == BH_OK) rc = settle(s, 0.0);
  }
  if (rc != BH_OK) {
    bh_sim_destroy(s);
    return rc;
  }
  *out = s;
  return BH_OK;
}

int bh_sim_run(bh_sim* s, double dt, int64_t steps, int32_t rebuild_every, int32_t use_graph) {
  BH_REQ(s, "simulation is NULL");
  BH_REQ(steps >= 0 && rebuild_every >= 0, "steps and rebuild_every must be non-negative");
  bh_ctx* c = s->ctx;
  if (hipSetDevice(c->device) != hipSuccess) {
    set_error("cannot select device %d", c->device);
    return BH_EHIP;
  }
  BH_HIP(hipStreamSynchronize(c->stream));  // the caller's writes to x, v
  StreamSwap sw(c, s->stream);
  for (int64_t k = 0; k < steps; k++) {
    const int kind = (rebuild_every > 0 && (s->step_index + 1) % rebuild_every == 0) ? 1 : 0;
    if (use_graph) {
      // a step kind runs eagerly the first time (sizes the caches and the
      // lists), then is captured and replayed
      if (!s->exec[kind] || s->graph_dt[kind] != dt) {
        const bool seen = kind == 0 ? s->refits > 0 : s->rebuilds > 0;
        if (seen) BH_TRY(capture(s, kind, dt));
      }
    }
    if (use_graph && s->exec[kind] && s->graph_dt[kind] == dt) {
      BH_HIP(hipGraphLaunch(s->exec[kind], c->stream));
      s->launches++;
    } else {
      BH_TRY(enqueue_step(s, kind, dt));
    }
    BH_TRY(settle(s, dt));
    s->step_index++;
    s->steps++;
    if (kind) s->rebuilds++;
    else s->refits++;
  }
  return BH_OK;
}

int bh_sim_accelerations(bh_sim* s, double* d_acc) {
  BH_REQ(s && d_acc, "NULL argument");
  BH_HIP(hipMemcpy(d_acc, s->acc, (size_t)s->n * s->dim * sizeof(double), hipMemcpyDeviceToDevice));
  return BH_OK;
}

int bh_sim_set_list_capacity(bh_sim* s, int64_t region, int64_t items) {
  BH_REQ(s, "simulation is NULL");
  BH_REQ(region >= 1 && items >= 1, "capacities must be positive");
  BH_HIP(hipStreamSynchronize(s->stream));
  s->cap_region = region;
  s->cap_items = items;
  drop_graph(s, 0);
  drop_graph(s, 1);
  return BH_OK;
}

int bh_sim_info(const bh_sim* s, int64_t* out) {
  BH_REQ(s && out, "NULL argument");
  const int64_t v[8] = {s->steps, s->rebuilds, s->refits, s->launches, s->captures, s->regrows, s->cap_region,
                        s->cap_items};
  for (int i = 0; i < 8; i++) out[i] = v[i];
  return BH_OK;
}

void bh_sim_destroy(bh_sim* s) {
  if (!s) return;
  bh_ctx* c = s->ctx;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graph(s, 0);
  drop_graph(s, 1);
  if (s->tree) bh_tree_free(s->tree);
  dfree(c, s->acc);
  dfree(c, s->status);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // extern "C"
