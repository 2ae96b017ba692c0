// bh_stream.hip -- accelerations at host-resident targets streamed through
// the device in chunks.
//
// Reference: the Dask target chunks of bhut (bhut/api.py:28-78 and
// Tree._accelerations_dask_targets api.py:690-731, map_blocks over target
// chunks in bhut/backends/dask_.py:173-228): the tree is built once and each
// chunk of targets is evaluated against it on its own.
//
// Here the chunks flow through the device on two streams.  Each stream
// belongs to a lane context of its own (own stream, own allocation cache, so
// scratch released by one stream's walk is never handed to the other's
// while it runs) and is driven by a host thread of its own: lane 0 takes
// chunks 0, 2, 4, ..., lane 1 chunks 1, 3, 5, ...  A lane gathers its chunk
// into pinned staging memory, copies it to the device, walks it, copies the
// result back and scatters it into the caller's array; while one lane's host
// thread gathers or scatters, the other lane's copies and kernels keep the
// device busy.  Device memory holds two chunks, whatever the target count.
//
// The targets are first put in the tree's Morton order on the host, exactly
// the order a single call gives them on the device (host_target_order), and
// the chunks are runs of that order whose length is a multiple of 64.  Each
// chunk therefore holds exactly the groups of 64 targets a single call would
// form, and every target's result, in the plain walk and in the FP32 form
// alike, is bit for bit the one a single call on all targets gives.
#include "bh_common.h"

#include <cstring>
#include <string>
#include <thread>

namespace bh {

// Pinned host staging of a lane context: chunk rows in and out.
static int lane_staging(bh_ctx* L, size_t bytes, double** in, double** out) {
  if (L->stage_bytes < 2 * bytes) {
    if (L->stage) (void)hipHostFree(L->stage);
    L->stage = nullptr;
    L->stage_bytes = 0;
    BH_HIP(hipHostMalloc(&L->stage, 2 * bytes, hipHostMallocDefault));
    L->stage_bytes = 2 * bytes;
  }
  *in = static_cast<double*>(L->stage);
  *out = reinterpret_cast<double*>(static_cast<char*>(L->stage) + bytes);
  return BH_OK;
}

struct StreamJob {
  bh_tree* t;
  const double* tgt;
  const uint32_t* order;
  int64_t m, chunk;
  double theta, eps, G;
  int precision;
  const bh_f32_opts* opts;
  double* out;
  std::atomic<bool> failed{false};
};

// One lane's share of the chunks (lane li of 2).  Returns a status and the
// message of a failure (errors are recorded per thread).
static int lane_run(bh_ctx* L, int li, StreamJob* J, std::string* msg) {
  DeviceGuard g(L->device);
  const int dim = J->t->dim;
  const int64_t nchunks = cdiv(J->m, J->chunk);
  const size_t bytes = (size_t)J->chunk * dim * sizeof(double);
  double *pin_in, *pin_out;
  int rc = lane_staging(L, bytes, &pin_in, &pin_out);
  Scratch s(L);
  double* d_in = rc ? nullptr : s.get<double>((size_t)J->chunk * dim);
  double* d_out = rc ? nullptr : s.get<double>((size_t)J->chunk * dim);
  if (!rc && (!d_in || !d_out)) rc = BH_ENOMEM;
  for (int64_t k = li; !rc && k < nchunks && !J->failed.load(); k += 2) {
    const int64_t a = k * J->chunk, mk = std::min(J->chunk, J->m - a);
    const uint32_t* ord = J->order + a;
    for (int64_t r = 0; r < mk; r++) std::memcpy(pin_in + r * dim, J->tgt + (int64_t)ord[r] * dim, dim * 8);
    hipError_t e = hipMemcpyAsync(d_in, pin_in, mk * dim * 8, hipMemcpyHostToDevice, L->stream);
    if (e != hipSuccess) {
      rc = hip_fail(e, "streamed chunk upload", __FILE__, __LINE__);
      break;
    }
    rc = J->precision == 64
             ? walk_mode(L, J->t, WALK_ACC, d_in, mk, J->theta, J->eps, J->G, d_out, nullptr, nullptr, nullptr)
             : walk_f32(L, J->t, d_in, mk, J->theta, J->eps, J->G, J->opts, d_out, nullptr);
    if (rc) break;
    e = hipMemcpyAsync(pin_out, d_out, mk * dim * 8, hipMemcpyDeviceToHost, L->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
    if (e != hipSuccess) {
      rc = hip_fail(e, "streamed chunk download", __FILE__, __LINE__);
      break;
    }
    for (int64_t r = 0; r < mk; r++) std::memcpy(J->out + (int64_t)ord[r] * dim, pin_out + r * dim, dim * 8);
  }
  if (rc) {
    J->failed.store(true);
    *msg = bh_last_error();
  }
  return rc;
}

}  // namespace bh

using namespace bh;

extern "C" int bh_accel_targets_streamed(bh_tree* t, const double* tgt, int64_t m, int64_t chunk, double theta,
                                         double eps, double G, int32_t precision, const bh_f32_opts* opts,
                                         double* out) {
  BH_REQ(t && (m == 0 || (tgt && out)), "NULL argument");
  BH_REQ(m >= 0 && m < 0xffffffffll, "bad target count");
  BH_REQ(t->nsys == 1, "external targets need a single tree, not a forest");
  BH_REQ(precision == 64 || precision == 32, "precision must be 64 or 32, got %d", precision);
  if (m == 0) return BH_OK;
  if (chunk <= 0) chunk = 1 << 20;
  BH_REQ(chunk < 0x7fffffffll, "chunk of %lld targets too large", (long long)chunk);
  chunk = std::min(chunk, m);
  chunk = (chunk + WAVE - 1) / WAVE * WAVE;  // whole groups of 64 targets
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  const int dim = t->dim;
  // the tree's box orders the targets; FP32 source copies and moments are
  // made once here, before two threads walk the tree
  double box[6];
  BH_HIP(hipMemcpyAsync(box, t->gbox, 2 * dim * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (precision == 32 && !t->f32_ready) BH_TRY(f32_prepare(c, t));
  BH_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 2; i++)
    if (!c->lanes[i]) BH_TRY(bh_ctx_create(c->device, &c->lanes[i]));
  std::vector<uint32_t> order(m);
  host_target_order(tgt, m, dim, box, order.data());
  StreamJob J;
  J.t = t;
  J.tgt = tgt;
  J.order = order.data();
  J.m = m;
  J.chunk = chunk;
  J.theta = theta;
  J.eps = eps;
  J.G = G;
  J.precision = precision;
  J.opts = opts;
  J.out = out;
  int rc1 = BH_OK;
  std::string msg0, msg1;
  std::thread second([&] { rc1 = lane_run(c->lanes[1], 1, &J, &msg1); });
  const int rc0 = lane_run(c->lanes[0], 0, &J, &msg0);
  second.join();
  if (rc0 || rc1) {
    set_error("%s", (rc0 ? msg0 : msg1).c_str());
    return rc0 ? rc0 : rc1;
  }
  return BH_OK;
}
