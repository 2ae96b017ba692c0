// bh_api.hip -- context, caching allocator and the C ABI of libbhut_hip.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <vector>

#include "bh_common.h"

namespace bh {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return e == hipErrorOutOfMemory ? BH_ENOMEM : BH_EHIP;
}

static size_t round_size(size_t b) {
  if (b < (1u << 20)) return (b + 255) & ~size_t(255);
  return (b + (size_t(1) << 16) - 1) & ~((size_t(1) << 16) - 1);
}

static void trim_locked(bh_ctx* c) {
  for (auto& kv : c->free_blocks) (void)hipFree(kv.second);
  c->free_blocks.clear();
  c->cached_bytes = 0;
}

// Device allocations are cached per context and reused by size: every
// kernel runs on the context stream, so a block released by one call and
// handed to the next is ordered behind the kernels that used it.
void* dalloc(bh_ctx* c, size_t bytes) {
  const size_t rs = round_size(bytes);
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->free_blocks.lower_bound(rs);
  if (it != c->free_blocks.end() && it->first <= rs + rs / 4) {
    void* p = it->second;
    c->live[p] = it->first;
    c->cached_bytes -= it->first;
    c->free_blocks.erase(it);
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, rs);
  if (e != hipSuccess && !c->capturing) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(c->stream);
    trim_locked(c);
    e = hipMalloc(&p, rs);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed", rs);
      return nullptr;
    }
  } else if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation of %zu bytes failed during graph capture", rs);
    return nullptr;
  }
  c->live[p] = rs;
  return p;
}

void dfree(bh_ctx* c, void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->live.find(p);
  if (it == c->live.end()) return;
  if (c->retain) {  // a graph being captured still refers to it
    c->retain->push_back(p);
    return;
  }
  c->free_blocks.emplace(it->second, p);
  c->cached_bytes += it->second;
  c->live.erase(it);
}

static hipEvent_t prof_event(bh_ctx* c) {
  if (!c->prof_pool.empty()) {
    hipEvent_t e = c->prof_pool.back();
    c->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

ProfScope::ProfScope(bh_ctx* ctx, const char* name) : c(ctx) {
  if (!c->prof || c->capturing) return;
  bh_ctx::ProfRec r{name, prof_event(c), prof_event(c)};
  (void)hipEventRecord(r.a, c->stream);
  idx = c->prof_recs.size();
  c->prof_recs.push_back(r);
}

ProfScope::~ProfScope() {
  if (idx != (size_t)-1) (void)hipEventRecord(c->prof_recs[idx].b, c->stream);
}

static void prof_reset(bh_ctx* c) {
  for (auto& r : c->prof_recs) {
    c->prof_pool.push_back(r.a);
    c->prof_pool.push_back(r.b);
  }
  c->prof_recs.clear();
}


static void free_tree_arrays(bh_tree* t) {
  bh_ctx* c = t->ctx;
  void* ps[] = {t->spart, t->perm,  t->start, t->count, t->first_child, t->parent, t->skip, t->ndepth,
                t->leaf_order, t->center, t->half, t->mass, t->com, t->bmin, t->bmax, t->gbox,
                t->nodes, t->spart32, t->pleaf, t->quad, t->quad32, t->d_sys_poff, t->d_sys_noff, t->d_gtab,
                t->slot_node, t->d_lvl_sys, t->d_slot_off};
  for (void* p : ps) dfree(c, p);
}

template <class T>
static bool alloc_into(bh_ctx* c, T*& p, size_t n) {
  p = dalloc_t<T>(c, n);
  return p != nullptr;
}

static int alloc_topology(bh_tree* t) {
  bh_ctx* c = t->ctx;
  const size_t M = (size_t)t->M;
  if (!alloc_into(c, t->start, M) || !alloc_into(c, t->count, M) ||
      !alloc_into(c, t->first_child, M) || !alloc_into(c, t->parent, M) ||
      !alloc_into(c, t->skip, M) || !alloc_into(c, t->ndepth, M) ||
      !alloc_into(c, t->leaf_order, (size_t)t->nleaves) ||
      !alloc_into(c, t->center, M * t->dim) || !alloc_into(c, t->half, M) ||
      !alloc_into(c, t->mass, M) || !alloc_into(c, t->com, M * t->dim) ||
      !alloc_into(c, t->bmin, M * t->dim) || !alloc_into(c, t->bmax, M * t->dim) ||
      !alloc_into(c, t->nodes, M))
    return BH_ENOMEM;
  return BH_OK;
}

// The device system offsets of a single tree: {0, n} and {0, M}.
__global__ void k_sys1(int32_t* poff, int32_t* noff, int32_t n, int32_t M) {
  if (threadIdx.x == 0) {
    poff[0] = 0;
    poff[1] = n;
    noff[0] = 0;
    noff[1] = M;
  }
}

// System offsets of a tree (nsys = 1) or forest, host and device copies.
static int set_systems(bh_tree* t, const int64_t* counts, int64_t nsys) {
  bh_ctx* c = t->ctx;
  t->nsys = nsys;
  t->sys_poff.assign(nsys + 1, 0);
  for (int64_t s = 0; s < nsys; s++) t->sys_poff[s + 1] = t->sys_poff[s] + counts[s];
  BH_TRY(forest_plan(counts, nsys, t->leaf, t->sys_noff, t->sys_loff, &t->M, &t->nleaves, &t->depth));
  std::vector<int32_t> p32(nsys + 1), n32(nsys + 1);
  for (int64_t s = 0; s <= nsys; s++) {
    p32[s] = (int32_t)t->sys_poff[s];
    n32[s] = (int32_t)t->sys_noff[s];
  }
  if (!alloc_into(c, t->d_sys_poff, (size_t)nsys + 1) || !alloc_into(c, t->d_sys_noff, (size_t)nsys + 1))
    return BH_ENOMEM;
  if (nsys == 1) {
    // no host copy (and no synchronisation) for a single tree
    hipLaunchKernelGGL(k_sys1, dim3(1), dim3(64), 0, c->stream, t->d_sys_poff, t->d_sys_noff, p32[1], n32[1]);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  BH_HIP(hipMemcpyAsync(t->d_sys_poff, p32.data(), (nsys + 1) * 4, hipMemcpyHostToDevice, c->stream));
  BH_HIP(hipMemcpyAsync(t->d_sys_noff, n32.data(), (nsys + 1) * 4, hipMemcpyHostToDevice, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));  // the host vectors are temporaries
  return BH_OK;
}

// build_tree for one system, or for a forest of independent systems whose
// particles are consecutive runs of counts[s] (one segmented sort, one
// topology launch, one upward pass for all of them).
static int build_dev(bh_ctx* c, const double* pos, const double* mass, const int64_t* counts, int64_t nsys,
                     int dim, int leaf, bh_tree* t) {
  t->ctx = c;
  t->dim = dim;
  t->leaf = leaf;
  t->max_leaf = leaf;  // a split node has more than `leaf` particles, a leaf at most `leaf`
  int64_t n = 0;
  for (int64_t s = 0; s < nsys; s++) n += counts[s];
  t->n = n;
  if (!alloc_into(c, t->gbox, 2 * (size_t)dim * nsys) || !alloc_into(c, t->spart, (size_t)n) ||
      !alloc_into(c, t->perm, (size_t)n))
    return BH_ENOMEM;
  // the (N, leaf) tables give M before the node arrays are allocated
  BH_TRY(set_systems(t, counts, nsys));
  t->binary = true;
  BH_TRY(alloc_topology(t));
  {
    ProfScope p(c, "build_topology");
    BH_TRY(topology(c, t));
  }
  return rebuild_in_place(c, t, pos, mass);
}

// Everything of a build after the topology: boxes, Morton codes, the stable
// (system, Morton) sort, the gather into sorted order and the upward pass.
// No host synchronisation, so it may be captured into a graph.
int rebuild_in_place(bh_ctx* c, bh_tree* t, const double* pos, const double* mass) {
  const int64_t n = t->n, nsys = t->nsys;
  const int dim = t->dim;
  Scratch s(c);
  uint64_t* k0 = s.get<uint64_t>(n);
  uint64_t* k1 = s.get<uint64_t>(n);
  uint32_t* v1 = s.get<uint32_t>(n);
  if (!k0 || !k1 || !v1) return BH_ENOMEM;
  uint32_t* v0 = reinterpret_cast<uint32_t*>(t->perm);
  {
    ProfScope p(c, "build_bbox_morton");
    if (nsys == 1) {
      BH_TRY(bbox(c, pos, n, dim, t->gbox));
      BH_TRY(morton_iota(c, pos, n, dim, t->gbox, 0, k0, v0));
    } else {
      BH_TRY(bbox_seg(c, pos, t->d_sys_poff, nsys, dim, t->gbox));
      BH_TRY(morton_seg(c, pos, n, dim, t->d_sys_poff, nsys, t->gbox, k0, v0));
    }
  }
  {
    ProfScope p(c, "build_radix_sort");
    uint64_t* ko;
    uint32_t* vo;
    BH_TRY(radix_sort(c, k0, v0, n, 21 * dim, &ko, &vo, k1, v1));
    if (vo != v0) BH_HIP(hipMemcpyAsync(v0, vo, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    if (nsys > 1) {
      // stable passes on the system id: (system, Morton) order
      BH_TRY(system_keys(c, v0, n, t->d_sys_poff, nsys, k0));
      int bits = 0;
      while ((1ll << bits) < nsys) bits++;
      BH_TRY(radix_sort(c, k0, v0, n, bits, &ko, &vo, k1, v1));
      if (vo != v0) BH_HIP(hipMemcpyAsync(v0, vo, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    }
  }
  {
    ProfScope p(c, "build_gather");
    BH_TRY(gather(c, pos, mass, v0, n, dim, t->spart));
  }
  {
    ProfScope p(c, "build_upward");
    BH_TRY(upward(c, t, 0));
    BH_TRY(pack_nodes(c, t));
  }
  t->f32_ready = false;
  return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

const char* bh_last_error(void) { return g_err.c_str(); }

int bh_version(void) { return 1; }

int bh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int bh_ctx_create(int device, bh_ctx** out) {
  BH_REQ(out, "out is NULL");
  *out = nullptr;
  int n = bh_device_count();
  if (n <= 0) {
    set_error("no HIP device is visible");
    return BH_ENODEV;
  }
  BH_REQ(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
  DeviceGuard g(device);
  bh_ctx* c = new bh_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
  }
  c->stream = c->own_stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->num_cus = prop.multiProcessorCount;
    c->mem_total = prop.totalGlobalMem;
  }
  // the no-allocation entry point's workspace on this device (looked up
  // once per device, with the device current)
  const int rc = noalloc_setup(device);
  if (rc != BH_OK) {
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return rc;
  }
  *out = c;
  return BH_OK;
}

void bh_ctx_destroy(bh_ctx* c) {
  if (!c) return;
  for (bh_ctx*& L : c->lanes) {
    bh_ctx_destroy(L);
    L = nullptr;
  }
  DeviceGuard g(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->stage) (void)hipHostFree(c->stage);
  {
    std::lock_guard<std::mutex> lk(c->mu);
    trim_locked(c);
    for (auto& kv : c->live) (void)hipFree(kv.first);
    c->live.clear();
  }
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

int bh_ctx_synchronize(bh_ctx* c) {
  BH_REQ(c, "ctx is NULL");
  DeviceGuard g(c->device);
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

void* bh_ctx_stream(bh_ctx* c) { return c ? (void*)c->stream : nullptr; }

int bh_ctx_set_list_budget(bh_ctx* c, int64_t bytes) {
  BH_REQ(c, "ctx is NULL");
  BH_REQ(bytes >= 0, "budget must be non-negative");
  c->list_budget = (size_t)bytes;
  return BH_OK;
}

int bh_ctx_set_stream(bh_ctx* c, void* stream) {
  BH_REQ(c, "ctx is NULL");
  DeviceGuard g(c->device);
  BH_HIP(hipStreamSynchronize(c->stream));
  c->stream = stream ? (hipStream_t)stream : c->own_stream;
  return BH_OK;
}

int bh_ctx_profile(bh_ctx* c, int32_t enable) {
  BH_REQ(c, "ctx is NULL");
  DeviceGuard g(c->device);
  BH_HIP(hipStreamSynchronize(c->stream));
  prof_reset(c);
  c->prof = enable != 0;
  return BH_OK;
}

int bh_ctx_profile_summary(bh_ctx* c, char* buf, int64_t cap) {
  BH_REQ(c && buf && cap > 0, "NULL argument");
  DeviceGuard g(c->device);
  BH_HIP(hipStreamSynchronize(c->stream));
  std::map<std::string, std::pair<int64_t, double>> agg;
  for (auto& r : c->prof_recs) {
    float ms = 0.f;
    BH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    auto& e = agg[r.name];
    e.first += 1;
    e.second += ms;
  }
  std::string s;
  char line[256];
  for (auto& kv : agg) {
    snprintf(line, sizeof(line), "%s %lld %.6f\n", kv.first.c_str(), (long long)kv.second.first,
             kv.second.second);
    s += line;
  }
  snprintf(buf, (size_t)cap, "%s", s.c_str());
  return (int64_t)s.size() < cap ? BH_OK : BH_EOVERFLOW;
}

int bh_ctx_trim(bh_ctx* c) {
  BH_REQ(c, "ctx is NULL");
  for (bh_ctx* L : c->lanes)
    if (L) BH_TRY(bh_ctx_trim(L));
  DeviceGuard g(c->device);
  BH_HIP(hipStreamSynchronize(c->stream));
  std::lock_guard<std::mutex> lk(c->mu);
  trim_locked(c);
  return BH_OK;
}

// ------------------------------------------------------------------ build
int bh_tree_build_dev(bh_ctx* c, const double* d_pos, const double* d_mass, int64_t n, int32_t dim,
                      int32_t leaf_size, bh_tree** out) {
  BH_REQ(c && out, "NULL argument");
  *out = nullptr;
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  BH_REQ(leaf_size > 0, "leaf_size must be positive, got %d", leaf_size);
  if (n <= 0) {
    set_error("Cannot build tree with zero particles");
    return BH_EEMPTY;
  }
  BH_REQ(n < 0x7fffffffll, "n = %lld exceeds the int32 index range", (long long)n);
  BH_REQ(d_pos && d_mass, "NULL positions or masses");
  DeviceGuard g(c->device);
  bh_tree* t = new bh_tree();
  const int64_t counts[1] = {n};
  // enqueued on the context stream without waiting for it (like every _dev
  // entry point): the work that uses the tree follows on the same stream
  int rc = build_dev(c, d_pos, d_mass, counts, 1, dim, leaf_size, t);
  if (rc != BH_OK) {
    free_tree_arrays(t);
    delete t;
    return rc;
  }
  *out = t;
  return BH_OK;
}

int bh_tree_build(bh_ctx* c, const double* pos, const double* mass, int64_t n, int32_t dim,
                  int32_t leaf_size, bh_tree** out) {
  BH_REQ(c && out, "NULL argument");
  *out = nullptr;
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  if (n <= 0) {
    set_error("Cannot build tree with zero particles");
    return BH_EEMPTY;
  }
  BH_REQ(pos && mass, "NULL positions or masses");
  DeviceGuard g(c->device);
  Scratch s(c);
  double* dp = s.get<double>((size_t)n * dim);
  double* dm = s.get<double>((size_t)n);
  if (!dp || !dm) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(dp, pos, (size_t)n * dim * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BH_HIP(hipMemcpyAsync(dm, mass, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BH_TRY(bh_tree_build_dev(c, dp, dm, n, dim, leaf_size, out));
  // host entry points return with their work done
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    bh_tree_free(*out);
    *out = nullptr;
    return hip_fail(e, "build", __FILE__, __LINE__);
  }
  return BH_OK;
}

int bh_forest_build_dev(bh_ctx* c, const double* d_pos, const double* d_mass, const int64_t* counts,
                        int64_t nsys, int32_t dim, int32_t leaf_size, bh_tree** out) {
  BH_REQ(c && out && counts, "NULL argument");
  *out = nullptr;
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  BH_REQ(leaf_size > 0, "leaf_size must be positive, got %d", leaf_size);
  BH_REQ(nsys >= 1 && nsys <= (1 << 24), "number of systems must be in [1, 2^24], got %lld", (long long)nsys);
  int64_t n = 0;
  for (int64_t s = 0; s < nsys; s++) {
    if (counts[s] <= 0) {
      set_error("Cannot build tree with zero particles (system %lld)", (long long)s);
      return BH_EEMPTY;
    }
    n += counts[s];
  }
  BH_REQ(n < 0x7fffffffll, "n = %lld exceeds the int32 index range", (long long)n);
  BH_REQ(d_pos && d_mass, "NULL positions or masses");
  DeviceGuard g(c->device);
  bh_tree* t = new bh_tree();
  int rc = build_dev(c, d_pos, d_mass, counts, nsys, dim, leaf_size, t);
  if (rc != BH_OK) {
    free_tree_arrays(t);
    delete t;
    return rc;
  }
  *out = t;
  return BH_OK;
}

int bh_forest_build(bh_ctx* c, const double* pos, const double* mass, const int64_t* counts, int64_t nsys,
                    int32_t dim, int32_t leaf_size, bh_tree** out) {
  BH_REQ(c && out && counts && pos && mass, "NULL argument");
  *out = nullptr;
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  int64_t n = 0;
  for (int64_t s = 0; s < nsys; s++) n += counts[s] > 0 ? counts[s] : 0;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* dp = s.get<double>((size_t)n * dim);
  double* dm = s.get<double>((size_t)n);
  if (!dp || !dm) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(dp, pos, (size_t)n * dim * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BH_HIP(hipMemcpyAsync(dm, mass, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BH_TRY(bh_forest_build_dev(c, dp, dm, counts, nsys, dim, leaf_size, out));
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    bh_tree_free(*out);
    *out = nullptr;
    return hip_fail(e, "forest build", __FILE__, __LINE__);
  }
  return BH_OK;
}

int bh_tree_systems(const bh_tree* t, int64_t* nsys, int64_t* particle_offsets, int64_t* node_offsets) {
  BH_REQ(t, "tree is NULL");
  if (nsys) *nsys = t->nsys;
  for (int64_t s = 0; s <= t->nsys; s++) {
    if (particle_offsets) particle_offsets[s] = t->sys_poff[s];
    if (node_offsets) node_offsets[s] = t->sys_noff[s];
  }
  return BH_OK;
}

void bh_tree_free(bh_tree* t) {
  if (!t) return;
  DeviceGuard g(t->ctx->device);
  (void)hipStreamSynchronize(t->ctx->stream);
  free_tree_arrays(t);
  delete t;
}

int bh_tree_shape(const bh_tree* t, int64_t* n, int64_t* num_nodes, int32_t* dim, int32_t* leaf_size,
                  int32_t* depth) {
  BH_REQ(t, "tree is NULL");
  if (n) *n = t->n;
  if (num_nodes) *num_nodes = t->M;
  if (dim) *dim = t->dim;
  if (leaf_size) *leaf_size = t->leaf;
  if (depth) *depth = t->depth;
  return BH_OK;
}

int bh_tree_f32_stack(const bh_tree* t, int32_t* out) {
  BH_REQ(t && out, "tree and out must not be NULL");
  out[0] = t->stack_hint[0].load();
  out[1] = t->stack_hint[1].load();
  return BH_OK;
}

int bh_tree_export(const bh_tree* t, double* center, double* half_size, double* mass, double* com,
                   int64_t* first_child, int64_t* child_count, int64_t* start, int64_t* count,
                   uint8_t* is_leaf, int64_t* perm) {
  BH_REQ(t, "tree is NULL");
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  const size_t M = (size_t)t->M, n = (size_t)t->n, D = (size_t)t->dim;
  if (center) BH_HIP(hipMemcpyAsync(center, t->center, M * D * 8, hipMemcpyDeviceToHost, c->stream));
  if (half_size) BH_HIP(hipMemcpyAsync(half_size, t->half, M * 8, hipMemcpyDeviceToHost, c->stream));
  if (mass) BH_HIP(hipMemcpyAsync(mass, t->mass, M * 8, hipMemcpyDeviceToHost, c->stream));
  if (com) BH_HIP(hipMemcpyAsync(com, t->com, M * D * 8, hipMemcpyDeviceToHost, c->stream));
  std::vector<int32_t> fc, st, cn, pm;
  if (first_child || child_count || is_leaf) {
    fc.resize(M);
    BH_HIP(hipMemcpyAsync(fc.data(), t->first_child, M * 4, hipMemcpyDeviceToHost, c->stream));
  }
  if (start) {
    st.resize(M);
    BH_HIP(hipMemcpyAsync(st.data(), t->start, M * 4, hipMemcpyDeviceToHost, c->stream));
  }
  if (count) {
    cn.resize(M);
    BH_HIP(hipMemcpyAsync(cn.data(), t->count, M * 4, hipMemcpyDeviceToHost, c->stream));
  }
  if (perm) {
    pm.resize(n);
    BH_HIP(hipMemcpyAsync(pm.data(), t->perm, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  BH_HIP(hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < M; k++) {
    if (first_child) first_child[k] = fc[k];
    if (child_count) child_count[k] = fc[k] >= 0 ? 2 : 0;
    if (is_leaf) is_leaf[k] = fc[k] < 0;
    if (start) start[k] = st[k];
    if (count) count[k] = cn[k];
  }
  if (perm)
    for (size_t i = 0; i < n; i++) perm[i] = pm[i];
  return BH_OK;
}

int bh_tree_import(bh_ctx* c, int64_t n, int64_t M, int32_t dim, int32_t leaf_size,
                   const double* center, const double* half_size, const double* mass,
                   const double* com, const int64_t* first_child, const int64_t* child_count,
                   const int64_t* start, const int64_t* count, const int64_t* perm,
                   const double* src_pos, const double* src_mass, bh_tree** out) {
  BH_REQ(c && out, "NULL argument");
  *out = nullptr;
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  BH_REQ(n > 0 && M > 0 && n < 0x7fffffffll && M < 0x7fffffffll, "bad tree shape n=%lld M=%lld",
         (long long)n, (long long)M);
  BH_REQ(center && half_size && mass && com && first_child && child_count && start && count && perm &&
             src_pos && src_mass,
         "NULL array");
  // every index the kernels will follow is checked here
  for (int64_t i = 0; i < n; i++) BH_REQ(perm[i] >= 0 && perm[i] < n, "perm[%lld] out of range", (long long)i);
  std::vector<int32_t> skip(M, -1), par(M, -1), opento(M, -1), depth(M, 0);
  std::vector<NodeD> nodes(M);
  bool binary = true;
  for (int64_t k = 0; k < M; k++) {
    BH_REQ(start[k] >= 0 && count[k] >= 0 && start[k] + count[k] <= n, "node %lld range out of bounds",
           (long long)k);
    if (first_child[k] != -1) {
      BH_REQ(first_child[k] > k && child_count[k] >= 0 && first_child[k] + child_count[k] <= M,
             "node %lld children out of bounds", (long long)k);
      if (child_count[k] != 2) binary = false;
    }
  }
  for (int64_t k = 0; k < M; k++) {  // parents before children (first_child > k)
    if (first_child[k] == -1) continue;
    const int64_t fc = first_child[k], cc = child_count[k];
    for (int64_t i = 0; i < cc; i++) {
      skip[fc + i] = i > 0 ? (int32_t)(fc + i - 1) : skip[k];
      par[fc + i] = (int32_t)k;
      depth[fc + i] = depth[k] + 1;
    }
    opento[k] = cc > 0 ? (int32_t)(fc + cc - 1) : skip[k];
  }
  std::vector<int32_t> leaves;
  for (int64_t k = 0; k < M; k++) {
    NodeD& r = nodes[k];
    r.cx = com[k * dim];
    r.cy = com[k * dim + 1];
    r.cz = dim == 3 ? com[k * dim + 2] : 0.0;
    r.m = mass[k];
    r.size = 2 * half_size[k];
    r.fc = (int32_t)first_child[k];
    r.skip = skip[k];
    r.start = (int32_t)start[k];
    r.count = (int32_t)count[k];
    r.pad[0] = opento[k];
    r.pad[1] = depth[k];
    if (first_child[k] == -1) leaves.push_back((int32_t)k);
  }
  std::sort(leaves.begin(), leaves.end(), [&](int32_t a, int32_t b) { return start[a] < start[b]; });
  DeviceGuard g(c->device);
  bh_tree* t = new bh_tree();
  t->ctx = c;
  t->n = n;
  t->dim = dim;
  t->leaf = leaf_size;
  t->max_leaf = 0;
  for (int32_t v : leaves) t->max_leaf = std::max<int64_t>(t->max_leaf, count[v]);
  t->M = M;
  t->nleaves = (int64_t)leaves.size();
  t->depth = 0;
  t->imported = true;
  t->binary = binary;
  t->nsys = 1;
  t->sys_poff = {0, n};
  t->sys_noff = {0, M};
  t->sys_loff = {0, (int64_t)leaves.size()};
  int rc = BH_OK;
  const int32_t p32[2] = {0, (int32_t)n}, n32[2] = {0, (int32_t)M};
  std::vector<int32_t> st(M), cn(M), fc(M), pm(n);
  for (int64_t k = 0; k < M; k++) {
    st[k] = (int32_t)start[k];
    cn[k] = (int32_t)count[k];
    fc[k] = (int32_t)first_child[k];
  }
  for (int64_t i = 0; i < n; i++) pm[i] = (int32_t)perm[i];
  Scratch s(c);
  double* dp = s.get<double>((size_t)n * dim);
  double* dm = s.get<double>((size_t)n);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (rc == BH_OK) {
      hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) rc = hip_fail(e, "import copy", __FILE__, __LINE__);
    }
  };
  if (!dp || !dm || !alloc_into(c, t->gbox, 2 * (size_t)dim) || !alloc_into(c, t->spart, (size_t)n) ||
      !alloc_into(c, t->perm, (size_t)n) || alloc_topology(t) != BH_OK || !alloc_into(c, t->d_sys_poff, 2) ||
      !alloc_into(c, t->d_sys_noff, 2))
    rc = BH_ENOMEM;
  up(t->d_sys_poff, p32, 8);
  up(t->d_sys_noff, n32, 8);
  up(dp, src_pos, (size_t)n * dim * 8);
  up(dm, src_mass, (size_t)n * 8);
  up(t->perm, pm.data(), (size_t)n * 4);
  up(t->start, st.data(), (size_t)M * 4);
  up(t->count, cn.data(), (size_t)M * 4);
  up(t->first_child, fc.data(), (size_t)M * 4);
  up(t->parent, par.data(), (size_t)M * 4);
  up(t->skip, skip.data(), (size_t)M * 4);
  up(t->ndepth, depth.data(), (size_t)M * 4);
  up(t->leaf_order, leaves.data(), leaves.size() * 4);
  up(t->center, center, (size_t)M * dim * 8);
  up(t->half, half_size, (size_t)M * 8);
  up(t->mass, mass, (size_t)M * 8);
  up(t->com, com, (size_t)M * dim * 8);
  up(t->nodes, nodes.data(), (size_t)M * sizeof(NodeD));
  if (rc == BH_OK) rc = bbox(c, dp, n, dim, t->gbox);
  if (rc == BH_OK) rc = gather(c, dp, dm, reinterpret_cast<uint32_t*>(t->perm), n, dim, t->spart);
  if (rc == BH_OK) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "import", __FILE__, __LINE__);
  }
  if (rc != BH_OK) {
    free_tree_arrays(t);
    delete t;
    return rc;
  }
  *out = t;
  return BH_OK;
}

int bh_tree_clone(const bh_tree* src, bh_tree** out) {
  BH_REQ(src && out, "NULL argument");
  *out = nullptr;
  bh_ctx* c = src->ctx;
  DeviceGuard g(c->device);
  bh_tree* t = new bh_tree();
  t->ctx = c;
  t->n = src->n;
  t->dim = src->dim;
  t->leaf = src->leaf;
  t->max_leaf = src->max_leaf;
  t->M = src->M;
  t->nleaves = src->nleaves;
  t->depth = src->depth;
  t->imported = src->imported;
  t->binary = src->binary;
  int rc = BH_OK;
  if (!alloc_into(c, t->gbox, 2 * (size_t)t->dim) || !alloc_into(c, t->spart, (size_t)t->n) ||
      !alloc_into(c, t->perm, (size_t)t->n) || alloc_topology(t) != BH_OK)
    rc = BH_ENOMEM;
  auto cp = [&](void* dst, const void* s, size_t bytes) {
    if (rc == BH_OK && bytes) {
      hipError_t e = hipMemcpyAsync(dst, s, bytes, hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) rc = hip_fail(e, "clone copy", __FILE__, __LINE__);
    }
  };
  const size_t M = (size_t)t->M, n = (size_t)t->n, D = (size_t)t->dim;
  cp(t->gbox, src->gbox, 2 * D * 8);
  cp(t->spart, src->spart, n * sizeof(double4));
  cp(t->perm, src->perm, n * 4);
  cp(t->start, src->start, M * 4);
  cp(t->count, src->count, M * 4);
  cp(t->first_child, src->first_child, M * 4);
  cp(t->parent, src->parent, M * 4);
  cp(t->skip, src->skip, M * 4);
  cp(t->ndepth, src->ndepth, M * 4);
  cp(t->leaf_order, src->leaf_order, (size_t)t->nleaves * 4);
  cp(t->center, src->center, M * D * 8);
  cp(t->half, src->half, M * 8);
  cp(t->mass, src->mass, M * 8);
  cp(t->com, src->com, M * D * 8);
  cp(t->bmin, src->bmin, M * D * 8);
  cp(t->bmax, src->bmax, M * D * 8);
  cp(t->nodes, src->nodes, M * sizeof(NodeD));
  t->nsys = src->nsys;
  t->sys_poff = src->sys_poff;
  t->sys_noff = src->sys_noff;
  t->sys_loff = src->sys_loff;
  if (rc == BH_OK && (!alloc_into(c, t->d_sys_poff, (size_t)t->nsys + 1) ||
                      !alloc_into(c, t->d_sys_noff, (size_t)t->nsys + 1)))
    rc = BH_ENOMEM;
  cp(t->d_sys_poff, src->d_sys_poff, ((size_t)t->nsys + 1) * 4);
  cp(t->d_sys_noff, src->d_sys_noff, ((size_t)t->nsys + 1) * 4);
  if (rc == BH_OK) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "clone", __FILE__, __LINE__);
  }
  if (rc != BH_OK) {
    free_tree_arrays(t);
    delete t;
    return rc;
  }
  *out = t;
  return BH_OK;
}

// ------------------------------------------------------------------ refit
int bh_tree_should_refit_dev(bh_tree* t, const double* d_pos, int64_t n, int32_t dim,
                             int32_t* refit_ok) {
  BH_REQ(t && refit_ok, "NULL argument");
  BH_REQ(t->nsys == 1, "the refit-or-rebuild test applies to a single tree, not a forest");
  if (n != t->n || dim != t->dim) {
    *refit_ok = 0;
    return BH_OK;
  }
  DeviceGuard g(t->ctx->device);
  int ok = 0;
  BH_TRY(refit_check(t->ctx, t, d_pos, n, &ok));
  *refit_ok = ok;
  return BH_OK;
}

int bh_tree_should_refit(bh_tree* t, const double* pos, int64_t n, int32_t dim, int32_t* refit_ok) {
  BH_REQ(t && refit_ok, "NULL argument");
  if (n != t->n || dim != t->dim) {
    *refit_ok = 0;
    return BH_OK;
  }
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* dp = s.get<double>((size_t)n * dim);
  if (!dp) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(dp, pos, (size_t)n * dim * 8, hipMemcpyHostToDevice, c->stream));
  return bh_tree_should_refit_dev(t, dp, n, dim, refit_ok);
}

int bh_tree_refit_dev(bh_tree* t, const double* d_pos, const double* d_mass) {
  BH_REQ(t && d_pos, "NULL argument");
  BH_REQ(t->binary, "refit needs a tree whose internal nodes have two children");
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  BH_TRY(gather(c, d_pos, d_mass, reinterpret_cast<uint32_t*>(t->perm), t->n, t->dim, t->spart));
  BH_TRY(upward(c, t, 1));
  BH_TRY(pack_nodes(c, t));
  t->f32_ready = false;
  return BH_OK;
}

int bh_tree_refit(bh_tree* t, const double* pos, const double* mass) {
  BH_REQ(t && pos, "NULL argument");
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* dp = s.get<double>((size_t)t->n * t->dim);
  double* dm = mass ? s.get<double>((size_t)t->n) : nullptr;
  if (!dp || (mass && !dm)) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(dp, pos, (size_t)t->n * t->dim * 8, hipMemcpyHostToDevice, c->stream));
  if (mass) BH_HIP(hipMemcpyAsync(dm, mass, (size_t)t->n * 8, hipMemcpyHostToDevice, c->stream));
  BH_TRY(bh_tree_refit_dev(t, dp, dm));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

// ------------------------------------------------------------------ walks
int bh_accel_self_dev(bh_tree* t, double theta, double eps, double G, double* d_out) {
  BH_REQ(t && d_out, "NULL argument");
  DeviceGuard g(t->ctx->device);
  return walk_mode(t->ctx, t, WALK_ACC, nullptr, t->n, theta, eps, G, d_out, nullptr, nullptr,
                   nullptr);
}

int bh_accel_self(bh_tree* t, double theta, double eps, double G, double* out) {
  BH_REQ(t && out, "NULL argument");
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* d = s.get<double>((size_t)t->n * t->dim);
  if (!d) return BH_ENOMEM;
  BH_TRY(bh_accel_self_dev(t, theta, eps, G, d));
  BH_HIP(hipMemcpyAsync(out, d, (size_t)t->n * t->dim * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_accel_targets_dev(bh_tree* t, const double* d_tgt, int64_t m, double theta, double eps,
                         double G, double* d_out) {
  BH_REQ(t && (m == 0 || (d_tgt && d_out)), "NULL argument");
  BH_REQ(m >= 0 && m < 0x7fffffffll, "bad target count");
  BH_REQ(t->nsys == 1, "external targets need a single tree, not a forest");
  DeviceGuard g(t->ctx->device);
  return walk_mode(t->ctx, t, WALK_ACC, d_tgt, m, theta, eps, G, d_out, nullptr, nullptr, nullptr);
}

int bh_accel_targets_dev_noalloc(bh_tree* t, const double* d_tgt, int64_t m, double theta, double eps,
                                 double G, double* d_out) {
  BH_REQ(t && (m == 0 || (d_tgt && d_out)), "NULL argument");
  BH_REQ(m >= 0, "bad target count");
  BH_REQ(t->nsys == 1, "external targets need a single tree, not a forest");
  DeviceGuard g(t->ctx->device);
  return walk_targets_noalloc(t->ctx, t, d_tgt, m, theta, eps, G, d_out);
}

int bh_accel_targets(bh_tree* t, const double* tgt, int64_t m, double theta, double eps, double G,
                     double* out) {
  BH_REQ(t && (m == 0 || (tgt && out)), "NULL argument");
  if (m == 0) return BH_OK;
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* dt = s.get<double>((size_t)m * t->dim);
  double* d = s.get<double>((size_t)m * t->dim);
  if (!dt || !d) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(dt, tgt, (size_t)m * t->dim * 8, hipMemcpyHostToDevice, c->stream));
  BH_TRY(bh_accel_targets_dev(t, dt, m, theta, eps, G, d));
  BH_HIP(hipMemcpyAsync(out, d, (size_t)m * t->dim * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

// Softened potential through the plain walk: same node visits and opening
// decisions as the accelerations, -G m / sqrt(r^2 + eps^2) per accepted node
// and per leaf pair (self excluded in self mode).
int bh_potential_dev(bh_tree* t, const double* d_tgt, int64_t m, double theta, double eps, double G,
                     double* d_out) {
  BH_REQ(t && (m == 0 || d_out), "NULL argument");
  BH_REQ(d_tgt || m == t->n, "self mode needs m == n");
  BH_REQ(!d_tgt || t->nsys == 1, "external targets need a single tree, not a forest");
  BH_REQ(m >= 0 && m < 0x7fffffffll, "bad target count");
  DeviceGuard g(t->ctx->device);
  return walk_mode(t->ctx, t, WALK_POT, d_tgt, m, theta, eps, G, d_out, nullptr, nullptr, nullptr);
}

int bh_potential(bh_tree* t, const double* tgt, int64_t m, double theta, double eps, double G, double* out) {
  BH_REQ(t && (m == 0 || out), "NULL argument");
  BH_REQ(tgt || m == t->n, "self mode needs m == n");
  if (m == 0) return BH_OK;
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* dt = tgt ? s.get<double>((size_t)m * t->dim) : nullptr;
  double* d = s.get<double>((size_t)m);
  if ((tgt && !dt) || !d) return BH_ENOMEM;
  if (tgt) BH_HIP(hipMemcpyAsync(dt, tgt, (size_t)m * t->dim * 8, hipMemcpyHostToDevice, c->stream));
  BH_TRY(bh_potential_dev(t, dt, m, theta, eps, G, d));
  BH_HIP(hipMemcpyAsync(out, d, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_interaction_lists(bh_tree* t, const double* tgt, int64_t m, double theta, int64_t* ptr,
                         int32_t* ids, int64_t cap, int64_t* total) {
  BH_REQ(t && ptr, "NULL argument");
  BH_REQ(tgt || m == t->n, "self mode needs m == n");
  BH_REQ(!tgt || t->nsys == 1, "external targets need a single tree, not a forest");
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  const double* dt = nullptr;
  if (tgt && m > 0) {
    double* d = s.get<double>((size_t)m * t->dim);
    if (!d) return BH_ENOMEM;
    BH_HIP(hipMemcpyAsync(d, tgt, (size_t)m * t->dim * 8, hipMemcpyHostToDevice, c->stream));
    dt = d;
  }
  int64_t* dcnt = s.get<int64_t>((size_t)m + 1);
  if (!dcnt) return BH_ENOMEM;
  BH_TRY(walk_mode(c, t, WALK_LCOUNT, dt, m, theta, 0.0, 1.0, nullptr, nullptr, nullptr, dcnt));
  std::vector<int64_t> cnt(m);
  if (m) BH_HIP(hipMemcpyAsync(cnt.data(), dcnt, m * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  ptr[0] = 0;
  for (int64_t i = 0; i < m; i++) ptr[i + 1] = ptr[i] + cnt[i];
  const int64_t tot = ptr[m];
  if (total) *total = tot;
  if (!ids || cap <= 0 || tot == 0) return BH_OK;
  int64_t* dptr = s.get<int64_t>((size_t)m + 1);
  int32_t* dids = s.get<int32_t>((size_t)tot);
  if (!dptr || !dids) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(dptr, ptr, (m + 1) * 8, hipMemcpyHostToDevice, c->stream));
  BH_TRY(walk_mode(c, t, WALK_LFILL, dt, m, theta, 0.0, 1.0, nullptr, dptr, dids, nullptr));
  BH_HIP(hipMemcpyAsync(ids, dids, (size_t)std::min(tot, cap) * 4, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_counters(bh_tree* t, const double* tgt, int64_t m, double theta, int64_t* out) {
  BH_REQ(t && out, "NULL argument");
  BH_REQ(tgt || m == t->n, "self mode needs m == n");
  BH_REQ(!tgt || t->nsys == 1, "external targets need a single tree, not a forest");
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  const double* dt = nullptr;
  if (tgt && m > 0) {
    double* d = s.get<double>((size_t)m * t->dim);
    if (!d) return BH_ENOMEM;
    BH_HIP(hipMemcpyAsync(d, tgt, (size_t)m * t->dim * 8, hipMemcpyHostToDevice, c->stream));
    dt = d;
  }
  int64_t* dout = s.get<int64_t>((size_t)m * 5);
  if (!dout) return BH_ENOMEM;
  BH_TRY(walk_mode(c, t, WALK_STATS, dt, m, theta, 0.0, 1.0, nullptr, nullptr, nullptr, dout));
  if (m) BH_HIP(hipMemcpyAsync(out, dout, (size_t)m * 5 * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

// ------------------------------------------------------------------ FP32 form
int bh_accel_f32_self_dev(bh_tree* t, double theta, double eps, double G, const bh_f32_opts* o,
                          double* d_out, int64_t* stats) {
  BH_REQ(t && d_out, "NULL argument");
  DeviceGuard g(t->ctx->device);
  return walk_f32(t->ctx, t, nullptr, t->n, theta, eps, G, o, d_out, stats);
}

int bh_accel_f32_self(bh_tree* t, double theta, double eps, double G, const bh_f32_opts* o,
                      double* out, int64_t* stats) {
  BH_REQ(t && out, "NULL argument");
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* d = s.get<double>((size_t)t->n * t->dim);
  if (!d) return BH_ENOMEM;
  BH_TRY(bh_accel_f32_self_dev(t, theta, eps, G, o, d, stats));
  BH_HIP(hipMemcpyAsync(out, d, (size_t)t->n * t->dim * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_accel_f32_targets_dev(bh_tree* t, const double* d_tgt, int64_t m, double theta, double eps,
                             double G, const bh_f32_opts* o, double* d_out, int64_t* stats) {
  BH_REQ(t && (m == 0 || (d_tgt && d_out)), "NULL argument");
  BH_REQ(m >= 0 && m < 0x7fffffffll, "bad target count");
  BH_REQ(t->nsys == 1, "external targets need a single tree, not a forest");
  DeviceGuard g(t->ctx->device);
  return walk_f32(t->ctx, t, d_tgt, m, theta, eps, G, o, d_out, stats);
}

int bh_accel_f32_targets(bh_tree* t, const double* tgt, int64_t m, double theta, double eps,
                         double G, const bh_f32_opts* o, double* out, int64_t* stats) {
  BH_REQ(t && (m == 0 || (tgt && out)), "NULL argument");
  if (m == 0) return BH_OK;
  bh_ctx* c = t->ctx;
  DeviceGuard g(c->device);
  Scratch s(c);
  double* dt = s.get<double>((size_t)m * t->dim);
  double* d = s.get<double>((size_t)m * t->dim);
  if (!dt || !d) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(dt, tgt, (size_t)m * t->dim * 8, hipMemcpyHostToDevice, c->stream));
  BH_TRY(bh_accel_f32_targets_dev(t, dt, m, theta, eps, G, o, d, stats));
  BH_HIP(hipMemcpyAsync(out, d, (size_t)m * t->dim * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

// ------------------------------------------------------------------ direct sum
int bh_direct_dev(bh_ctx* c, const double* d_src, const double* d_mass, int64_t n,
                  const double* d_tgt, int64_t m, int32_t dim, int32_t self_mode, double eps,
                  double G, double* d_acc, double* d_pot) {
  BH_REQ(c, "ctx is NULL");
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3");
  BH_REQ(!self_mode || m == n, "self mode needs m == n");
  BH_REQ(n >= 0 && m >= 0, "negative size");
  DeviceGuard g(c->device);
  return direct(c, d_src, d_mass, n, self_mode ? d_src : d_tgt, m, dim, self_mode, eps, G, d_acc,
                d_pot);
}

int bh_direct(bh_ctx* c, const double* src, const double* mass, int64_t n, const double* tgt,
              int64_t m, int32_t dim, int32_t self_mode, double eps, double G, double* acc,
              double* pot) {
  BH_REQ(c, "ctx is NULL");
  BH_REQ(dim == 2 || dim == 3, "dim must be 2 or 3");
  BH_REQ(!self_mode || m == n, "self mode needs m == n");
  DeviceGuard g(c->device);
  Scratch s(c);
  double* ds = s.get<double>((size_t)n * dim);
  double* dm = s.get<double>((size_t)n);
  double* dt = self_mode ? ds : s.get<double>((size_t)m * dim);
  double* da = acc ? s.get<double>((size_t)m * dim) : nullptr;
  double* dp = pot ? s.get<double>((size_t)m) : nullptr;
  if (!ds || !dm || !dt || (acc && !da) || (pot && !dp)) return BH_ENOMEM;
  BH_HIP(hipMemcpyAsync(ds, src, (size_t)n * dim * 8, hipMemcpyHostToDevice, c->stream));
  BH_HIP(hipMemcpyAsync(dm, mass, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  if (!self_mode) BH_HIP(hipMemcpyAsync(dt, tgt, (size_t)m * dim * 8, hipMemcpyHostToDevice, c->stream));
  BH_TRY(direct(c, ds, dm, n, dt, m, dim, self_mode, eps, G, da, dp));
  if (acc) BH_HIP(hipMemcpyAsync(acc, da, (size_t)m * dim * 8, hipMemcpyDeviceToHost, c->stream));
  if (pot) BH_HIP(hipMemcpyAsync(pot, dp, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

}  // extern "C"
