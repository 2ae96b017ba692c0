// bh_common.h -- shared types and device helpers of libbhut_hip.
//
// Written for gfx950 (MI355X, wave64).  Every translation unit is compiled
// with -ffp-contract=off: the plain FP64 form reproduces bhut's rounding
// sequence operation by operation, and the FP32 form spells out its fused
// multiply-adds with fmaf().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <cstddef>
#include <mutex>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bhut_hip.h"

namespace bh {

constexpr int WAVE = 64;
constexpr int MAX_DEPTH = 64;  // topology tables; depth <= log2(N) + 2

// ------------------------------------------------------------ errors
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define BH_HIP(call)                                                       \
  do {                                                                     \
    hipError_t _e = (call);                                                \
    if (_e != hipSuccess) return ::bh::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define BH_LAUNCH_CHECK() BH_HIP(hipGetLastError())

// argument check: records the message and returns BH_EINVAL
#define BH_REQ(cond, ...)            \
  do {                               \
    if (!(cond)) {                   \
      ::bh::set_error(__VA_ARGS__);  \
      return BH_EINVAL;              \
    }                                \
  } while (0)

#define BH_TRY(expr)                  \
  do {                                \
    int _rc = (expr);                 \
    if (_rc != BH_OK) return _rc;     \
  } while (0)

// ------------------------------------------------------------ records
// Traversal record of one node (64 B, one cache line).  `skip` is the next
// node of bhut's pop order once this node's subtree is done; the LIFO stack
// walk of bh.py:506-546 pops the right child first, so the pop order is the
// right-first preorder and `skip` reproduces it without a stack.
struct __attribute__((aligned(64))) NodeD {
  double cx, cy, cz, m;  // centre of mass, mass
  double size;           // 2*half_size (bh.py:589)
  int32_t fc;            // first child (-1 for leaves)
  int32_t skip;          // next node after this subtree (-1: walk ends)
  int32_t start, count;  // particle range in sorted order
  int32_t pad[2];        // [0] node popped first when opened, [1] depth below the system root
};
static_assert(sizeof(NodeD) == 64, "NodeD layout");

// The node of thread t at depth d of the level-by-level upward passes:
// nthreads = 2^d x the number of systems at least d deep; lvl_sys lists those
// systems (NULL: one tree), slot_off gives each system's first slot (NULL:
// one tree), slot_node maps a (depth, position) slot to its node (-1 below a
// leaf).
__device__ __forceinline__ int level_node(int d, int64_t t, const int32_t* __restrict__ lvl_sys,
                                          const int64_t* __restrict__ slot_off,
                                          const int32_t* __restrict__ slot_node) {
  const int64_t sys = lvl_sys ? lvl_sys[t >> d] : 0;
  const int64_t base = slot_off ? slot_off[sys] : 0;
  return slot_node[base + (1ll << d) - 1 + (t & ((1ll << d) - 1))];
}

// ------------------------------------------------------------ context
struct Ctx;

struct Topo {
  int64_t n = 0, leaf = 0, M = 0, nleaves = 0;
  int depth = 0;  // maximum node depth (root = 0)
};

}  // namespace bh

struct bh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // stream all work is enqueued on
  hipStream_t own_stream = nullptr;  // the context's own (default-priority) stream
  int num_cus = 0;
  size_t mem_total = 0;  // device memory, bytes
  // caching device allocator: exact-size free lists
  std::mutex mu;
  std::multimap<size_t, void*> free_blocks;
  std::unordered_map<void*, size_t> live;
  size_t cached_bytes = 0;
  // small pinned staging buffer for flags / counters read back by the host
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  // While a stream capture records work into a graph, blocks released by
  // scoped scratch are kept on this list instead (the graph still uses them)
  // and no host synchronisation or event profiling may happen.
  std::vector<void*>* retain = nullptr;
  bool capturing = false;
  // bytes of FP32 interaction lists materialised at once (0: default)
  size_t list_budget = 0;
  // work items of the previous host-scanned FP32 call, with a margin
  int64_t items_hint = 65536;
  // group order for the next host-scanned FP32 walk (longest lists of the
  // previous call first); valid for walk_order_n groups
  int32_t* walk_order = nullptr;
  int64_t walk_order_cap = 0, walk_order_n = 0;
  // list lengths of the previous host-scanned FP32 call (cells, then leaves;
  // prev_ng groups) and their totals: region sizes for the next one-pass walk
  int64_t* prev_cnt = nullptr;
  int64_t prev_cap = 0, prev_ng = -1, prev_ntc = 0, prev_ntk = 0;
  // optional per-kernel timing with HIP events on the context stream
  bool prof = false;
  struct ProfRec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_pool;
  // streamed targets: two lane contexts on the same device (own stream, own
  // cache), created on first use; a lane's pinned host staging
  bh_ctx* lanes[2] = {nullptr, nullptr};
  void* stage = nullptr;
  size_t stage_bytes = 0;
};

namespace bh {
// Makes `dev` the calling thread's current device for the scope.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Records HIP events around the launches in its scope when profiling is on.
struct ProfScope {
  bh_ctx* c;
  size_t idx = (size_t)-1;
  ProfScope(bh_ctx* ctx, const char* name);
  ~ProfScope();
};
}  // namespace bh

struct bh_tree {
  bh_ctx* ctx = nullptr;
  int64_t n = 0;
  int dim = 3;
  int leaf = 32;
  int64_t max_leaf = 32;  // particles in the largest leaf (at most `leaf` for built trees)
  int64_t M = 0;
  int64_t nleaves = 0;
  int depth = 0;
  bool imported = false;  // built from host TreeData arrays
  bool binary = true;     // every internal node has two children (refit needs it)
  // forest of independent systems (nsys = 1 for a single tree): system s
  // owns sorted particles [sys_poff[s], sys_poff[s+1]), nodes from
  // sys_noff[s] (its root) and leaves from sys_loff[s] in leaf_order
  int64_t nsys = 1;
  std::vector<int64_t> sys_poff{0, 0}, sys_noff{0, 0}, sys_loff{0, 0};
  int32_t* d_sys_poff = nullptr;  // nsys + 1
  int32_t* d_sys_noff = nullptr;  // nsys + 1
  // particles in Morton order
  double4* spart = nullptr;   // n: x, y, z (0 in 2-D), m
  int32_t* perm = nullptr;    // n: sorted index -> original index
  // topology (pure function of n, leaf)
  int32_t* start = nullptr;
  int32_t* count = nullptr;
  int32_t* first_child = nullptr;
  int32_t* parent = nullptr;
  int32_t* skip = nullptr;
  int32_t* ndepth = nullptr;      // depth below the node's system root (root 0)
  int32_t* leaf_order = nullptr;  // leaves left to right
  // depth tables of a built tree (NULL slot_node: imported or cloned): slot
  // (system, depth d, position) -> node or -1; per depth the number of
  // systems deep enough, their ids (forest) and each system's first slot
  int32_t* slot_node = nullptr;
  int dmax = 0;
  std::vector<int64_t> lvl_nsys, h_lvl_off;
  int32_t* d_lvl_sys = nullptr;
  int64_t* d_slot_off = nullptr;
  // geometry and monopoles in TreeData layout
  double* center = nullptr;   // M*dim
  double* half = nullptr;     // M
  double* mass = nullptr;     // M
  double* com = nullptr;      // M*dim
  double* bmin = nullptr;     // M*dim node box (build only)
  double* bmax = nullptr;
  double* gbox = nullptr;     // 2*dim global pmin/pmax
  bh::NodeD* nodes = nullptr; // packed traversal records
  // FP32 form: particles relative to their leaf centre, each particle's
  // leaf, and the nodes' second moments about their centre of mass
  float4* spart32 = nullptr;
  int32_t* pleaf = nullptr;
  double* quad = nullptr;  // M*6: xx, yy, zz, xy, xz, yz
  float4* quad32 = nullptr;  // M*2: (xx, yy, zz, trace), (xy, xz, yz, 0) in FP32
  bool f32_ready = false;
  bool quad_ready = false;
  // forest group table of the FP32 form for group size gtab_gs: ng + 1
  // first targets, then ng roots (host copy kept for statistics)
  int32_t* d_gtab = nullptr;
  int gtab_gs = 0;
  int64_t gtab_ng = 0;
  std::vector<int32_t> h_gtab;
  // largest group-walk stack an earlier call on this tree had to grow to,
  // for self (0) and external (1) targets; later calls start there (the
  // walk's lists do not depend on the stack size)
  // (atomic: the streamed target path walks one tree from two host threads)
  std::atomic<int> stack_hint[2]{{0}, {0}};
};

namespace bh {

// ------------------------------------------------------------ memory
void* dalloc(bh_ctx* c, size_t bytes);
void dfree(bh_ctx* c, void* p);
template <class T>
T* dalloc_t(bh_ctx* c, size_t n) {
  return static_cast<T*>(dalloc(c, n * sizeof(T) + (n == 0 ? 16 : 0)));
}

// Scoped scratch allocations returned to the cache on exit.
struct Scratch {
  bh_ctx* c;
  std::vector<void*> ps;
  explicit Scratch(bh_ctx* ctx) : c(ctx) {}
  template <class T>
  T* get(size_t n) {
    T* p = dalloc_t<T>(c, n);
    if (p) ps.push_back(p);
    return p;
  }
  ~Scratch() {
    for (void* p : ps) dfree(c, p);
  }
};

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------ host entry points per TU
// bh_sort.hip
int bbox(bh_ctx* c, const double* pos, int64_t n, int dim, double* d_box);
// per-system boxes (nsys x 2*dim) and Morton codes relative to each box
int bbox_seg(bh_ctx* c, const double* pos, const int32_t* d_poff, int64_t nsys, int dim, double* d_boxes);
int morton_seg(bh_ctx* c, const double* pos, int64_t n, int dim, const int32_t* d_poff, int64_t nsys,
               const double* d_boxes, uint64_t* keys, uint32_t* vals);
// keys[i] = system of original particle vals[i]
int system_keys(bh_ctx* c, const uint32_t* vals, int64_t n, const int32_t* d_poff, int64_t nsys, uint64_t* keys);
int morton_iota(bh_ctx* c, const double* pos, int64_t n, int dim, const double* d_box, int clamp,
                uint64_t* keys, uint32_t* vals);
int radix_sort(bh_ctx* c, uint64_t* keys, uint32_t* vals, int64_t n, int bits, uint64_t** keys_out,
               uint32_t** vals_out, uint64_t* keys_alt, uint32_t* vals_alt);
// the same sort in a caller's workspace: counts holds 256 * radix_tiles(n)
// entries, totals 256 (both written before they are read)
constexpr int RADIX_TILE = 4096;  // keys per tile of the digit histograms
inline int64_t radix_tiles(int64_t n) { return (n + RADIX_TILE - 1) / RADIX_TILE; }
int radix_sort_ws(bh_ctx* c, uint64_t* keys, uint32_t* vals, int64_t n, int bits, uint64_t** keys_out,
                  uint32_t** vals_out, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t* counts,
                  uint32_t* totals);
int iota(bh_ctx* c, uint32_t* v, int64_t n);
int gather(bh_ctx* c, const double* pos, const double* mass, const uint32_t* perm, int64_t n,
           int dim, double4* spart);
// bh_build.hip
int topology_plan(int64_t N, int64_t L, int64_t* M, int64_t* nleaves, int* depth);
int forest_plan(const int64_t* counts, int64_t nsys, int64_t L, std::vector<int64_t>& noff,
                std::vector<int64_t>& loff, int64_t* M, int64_t* nleaves, int* depth);
int topology(bh_ctx* c, bh_tree* t);
int upward(bh_ctx* c, bh_tree* t, int refit);
int pack_nodes(bh_ctx* c, bh_tree* t);
int refit_check(bh_ctx* c, const bh_tree* t, const double* pos, int64_t n, int* ok);
int fill_u32(bh_ctx* c, uint32_t* p, uint32_t v, int64_t n);
// bh_walk.hip
enum { WALK_ACC = 0, WALK_LCOUNT = 1, WALK_LFILL = 2, WALK_STATS = 3, WALK_POT = 4 };
int order_targets(bh_ctx* c, const bh_tree* t, const double* tgt, int64_t m, uint32_t* torder);
// device scratch of the no-allocation entry point, looked up once per device
int noalloc_setup(int device);
int walk_targets_noalloc(bh_ctx* c, const bh_tree* t, const double* tgt, int64_t m, double theta, double eps,
                         double G, double* out);
int walk_mode(bh_ctx* c, const bh_tree* t, int mode, const double* tgt, int64_t m, double theta,
              double eps, double G, double* out, const int64_t* lptr, int32_t* lids, int64_t* stats);
// bh_group.hip
int f32_prepare(bh_ctx* c, bh_tree* t);
// Capture-safe variant of the FP32 self walk: list offsets are scanned on the
// device into lists of fixed capacity (cap_cells, cap_chunks); no host
// synchronisation.  When a list would exceed its capacity, bit 1 of *err is
// set, the lists and the evaluation are skipped, and need[0..1] receive the
// required capacities.  *err must be zero on entry.
// A planned FP32 walk (no host synchronisation, e.g. under graph capture):
// lists in regions of cap_region entries per group, at most cap_items work
// items; an overflow sets bit 2 of *err and need[] = (longest list, items).
struct F32Plan {
  int64_t cap_region = 0, cap_items = 0;
  uint32_t* err = nullptr;    // device
  int64_t* need = nullptr;    // device, 2 entries
};
int walk_f32(bh_ctx* c, bh_tree* t, const double* tgt, int64_t m, double theta, double eps,
             double G, const bh_f32_opts* o, double* out, int64_t* stats, const F32Plan* plan = nullptr);
// bh_sort.hip: the Morton order of external targets computed on the host,
// bit for bit the order order_targets gives them on the device
void host_target_order(const double* tgt, int64_t m, int dim, const double* box, uint32_t* order);
// bh_api.hip: a new Morton order, particle order and monopoles for a tree
// whose topology (a function of the system sizes) is kept
int rebuild_in_place(bh_ctx* c, bh_tree* t, const double* pos, const double* mass);
// bh_direct.hip
int direct(bh_ctx* c, const double* src, const double* mass, int64_t n, const double* tgt,
           int64_t m, int dim, int self_mode, double eps, double G, double* acc, double* pot);

// ------------------------------------------------------------ device helpers
// numpy.minimum / numpy.maximum: a NaN in either operand wins.
__device__ __forceinline__ double np_min(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ __forceinline__ double np_max(double a, double b) { return (a > b || a != a) ? a : b; }

// System owning particle i of a forest: the last s with poff[s] <= i.
// Systems are consecutive runs both in input order and in sorted order.
// Each wavefront touches only LDS it owns (its stack, its tile), and a
// wavefront's LDS operations complete in issue order, so a compiler-level
// wave barrier orders its own writes and reads; no workgroup barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ int64_t find_sys(const int32_t* __restrict__ poff, int64_t nsys, int64_t i) {
  int64_t lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (poff[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace bh
