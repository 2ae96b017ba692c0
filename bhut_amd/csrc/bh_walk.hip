// bh_walk.hip -- the plain FP64 per-target walk (bit-faithful form).
//
// Reference: _barnes_hut_single_target (bhut/traverse/bh.py:485-548) and
// _barnes_hut_single_target_external (bh.py:83-146), _should_open_node
// (bh.py:576-601), acc_monopole_node/acc_monopole (bhut/traverse/kernels.py:
// 14-69,136-168), the Numba leaf loops (bh.py:174-280) with their pure-Python
// fallback (bh.py:307-311,339-343,384-425).  WALK_POT evaluates the softened
// potential over the same visits (bhut documents `bhut.potential`,
// docs/getting-started.md:36, without implementing it).
//
// One lane per target, targets in Morton order (self targets in sorted order,
// external targets sorted on their Morton codes), 64 consecutive targets per
// wavefront.  The reference's LIFO stack pops the right child first, i.e.
// visits nodes in right-first preorder; each record carries the node that
// follows its subtree in that order (`skip`), the child popped first when it
// is opened (`pad[0]`) and its depth (`pad[1]`).
//
// The wavefront walks the UNION of its targets' walks in lockstep: one node
// at a time, in that preorder.  A target visits a node exactly when it opened
// every ancestor of the node, so every target's own visits are a subsequence
// of the union and each lane still tests, accepts and sums in its own pop
// order, bit for bit.  Per lane, `follow` is the depth down to which the
// lane opened the current path (-1: not walking this system): at a node of
// depth d it is min(follow, d), and the lane visits the node iff follow == d;
// opening the node sets follow = d + 1.  The wavefront goes down into the
// children iff some lane opened the node, else on to `skip`.  Node records
// are read once per wavefront (uniform scalar loads) and leaf particles once
// per wavefront through its LDS tile, and broadcast to the lanes; a leaf's
// particle pairs run on every visiting lane at once, where a per-lane walk of
// the same lists would run only the lanes that happened to be at a leaf.
#include "bh_common.h"
#include "bh_fp64.h"

namespace bh {

// The Numba leaf formula (bh.py:249-278) for one leaf with the middle-range
// square root and reciprocal: r^2 + eps^2 >= eps^2 >= 2^-700 holds by the
// caller's choice of this path, so only the upper end is checked; returns
// whether some pair of this lane left the range (the caller then recomputes
// the leaf with sqrt() and '/').  G1: G == 1, where G * m == m exactly.
// The leaf's sources pass through the wavefront's LDS tile `wt`, 64 at a time
// (one coalesced load per chunk instead of one scalar load, and its wait,
// per source), and a few of them are read from the tile before their pairs
// are computed; every lane still adds its pairs in source order.
// External targets compute every pair on every lane, with a zero magnitude
// on lanes that do not visit the leaf (whose partial sums are discarded),
// which lets the compiler interleave consecutive pairs; self targets keep the
// per-pair branch (and skip their own particle), which measured faster there
// (the walk of external targets touches more scattered leaves and gains from
// the overlap).
template <bool SELF>
constexpr int leaf_reads() { return SELF ? 4 : 2; }

template <int DIM, bool SELF, bool G1>
__device__ __forceinline__ bool leaf_mid(double4* __restrict__ wt, const double4* __restrict__ spart, int j0,
                                         int j1, int64_t t, bool mine, double tx, double ty, double tz, double eps2,
                                         double G, double& px, double& py, double& pz) {
  constexpr int LEAF_READS = leaf_reads<SELF>();
  bool bad = false;
  auto pair = [&](const double4& q, int j) {
    const bool use = mine && !(SELF && j == t);
    if (SELF && !use) return;
    const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
    double rsq = dx * dx;
    rsq = rsq + dy * dy;
    if (DIM == 3) rsq = rsq + dz * dz;
    rsq = rsq + eps2;
    bad |= use && !(rsq <= 0x1p700);
    const double rinv = recip_mid(sqrt_mid(rsq));
    const double ric = rinv * rinv * rinv;
    const double mag = use ? (G1 ? q.w : G * q.w) * ric : 0.0;
    px = px + mag * dx;
    py = py + mag * dy;
    if (DIM == 3) pz = pz + mag * dz;
  };
  const int lane = threadIdx.x % WAVE;
  for (int c0 = j0; c0 < j1; c0 += WAVE) {
    const int cn = j1 - c0 < WAVE ? j1 - c0 : WAVE;
    wave_sync();  // every read of the previous chunk precedes these writes
    if (lane < cn) wt[lane] = spart[c0 + lane];
    wave_sync();
    int k = 0;
    for (; k + LEAF_READS <= cn; k += LEAF_READS) {
      double4 q[LEAF_READS];
#pragma unroll
      for (int u = 0; u < LEAF_READS; u++) q[u] = wt[k + u];
#pragma unroll
      for (int u = 0; u < LEAF_READS; u++) pair(q[u], c0 + k + u);
    }
    for (; k < cn; k++) pair(wt[k], c0 + k);
  }
  return bad;
}

// At most 74 SGPRs (80 with VCC and the reserved ones): a gfx950 SIMD holds
// eight wavefronts of a kernel with 80 SGPRs but only seven with 82 and up
// (measured), and the compiler's own choice for this kernel was 86.
template <int DIM, bool SELF, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_num_sgpr(74))) void k_walk(const NodeD* __restrict__ nodes,
                                              const double4* __restrict__ spart,
                                              const int32_t* __restrict__ perm, int64_t m,
                                              const double* __restrict__ tgt,
                                              const uint32_t* __restrict__ torder, double theta,
                                              double eps2, double G, double* __restrict__ out,
                                              const int64_t* __restrict__ lptr,
                                              int32_t* __restrict__ lids,
                                              int64_t* __restrict__ stats,
                                              const int32_t* __restrict__ sys_poff,
                                              const int32_t* __restrict__ sys_noff, int64_t nsys) {
  __shared__ double4 ltile[256 / WAVE][WAVE];  // each wavefront's leaf sources (accelerations)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  // the whole wavefront past the end (its first target index is uniform)
  if (blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~(WAVE - 1)) >= m) return;
  const bool live = t < m;
  int64_t orig = 0;
  double tx = 0.0, ty = 0.0, tz = 0.0;
  if (live) {
    if (SELF) {
      orig = perm[t];
      const double4 p = spart[t];
      tx = p.x;
      ty = p.y;
      tz = p.z;
    } else {
      orig = torder[t];
      tx = tgt[orig * DIM + 0];
      ty = tgt[orig * DIM + 1];
      if (DIM == 3) tz = tgt[orig * DIM + 2];
    }
  }
  double ax = 0.0, ay = 0.0, az = 0.0, phi = 0.0;
  int64_t nl = 0, wpos = 0;
  int64_t c_pop = 0, c_test = 0, c_mono = 0, c_leaf = 0, c_pair = 0;
  if (MODE == WALK_LFILL && live) wpos = lptr[orig];
  // eps^2 in [2^-700, 2^690]: the leaf pairs take the middle-range square
  // root and reciprocal (every r^2 + eps^2 is at least eps^2)
  const bool fast_eps = eps2 >= 0x1p-700 && eps2 <= 0x1p690;
  const bool g_one = G == 1.0;
  // a forest walks each system alone: the lanes of one system at a time
  const int64_t sys = (SELF && nsys > 1 && live) ? find_sys(sys_poff, nsys, t) : 0;
  uint64_t todo = __ballot(live);
  while (todo) {
    const int lead = __ffsll((unsigned long long)todo) - 1;
    const int64_t s = __shfl(sys, lead);
    const bool part = live && sys == s;
    todo &= ~__ballot(part);
    int follow = part ? 0 : -1;
    int v = (SELF && nsys > 1) ? sys_noff[s] : 0;
    while (v >= 0) {
      v = __builtin_amdgcn_readfirstlane(v);
      const NodeD nd = nodes[v];
      const int dv = nd.pad[1];
      follow = min(follow, dv);
      const bool mine = follow == dv;
      if (MODE == WALK_STATS && mine) c_pop++;
      if (nd.m == 0.0) {  // bh.py:512-513
        v = nd.skip;
        continue;
      }
      if (nd.fc < 0) {  // leaf: direct particle-particle sum (bh.py:518-522)
        if (mine) {
          if (MODE == WALK_LCOUNT) nl++;
          if (MODE == WALK_LFILL) lids[wpos++] = v;
          if (MODE == WALK_STATS) {
            c_leaf++;
            c_pair += nd.count;
            if (SELF && t >= nd.start && t < nd.start + nd.count) c_pair--;
          }
        }
        const int j0 = nd.start, j1 = nd.start + nd.count;
        if (MODE == WALK_ACC) {
          double px = 0.0, py = 0.0, pz = 0.0;
          bool redo = true;
          if (fast_eps) {
            double4* wt = ltile[threadIdx.x / WAVE];
            const bool bad =
                g_one ? leaf_mid<DIM, SELF, true>(wt, spart, j0, j1, t, mine, tx, ty, tz, eps2, G, px, py, pz)
                      : leaf_mid<DIM, SELF, false>(wt, spart, j0, j1, t, mine, tx, ty, tz, eps2, G, px, py, pz);
            redo = __ballot(bad) != 0;
          }
          bool zdiv = false;
          // Numba loop (bh.py:249-278): r = sqrt(r2 + eps^2); 1/r; r^-3.
          // The sources are the same for every lane (scalar loads).
          if (redo) {
            px = py = pz = 0.0;
            for (int j = j0; j < j1; j++) {
              const double4 q = spart[j];
              if (mine && !(SELF && j == t)) {
                const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
                double rsq = dx * dx;
                rsq = rsq + dy * dy;
                if (DIM == 3) rsq = rsq + dz * dz;
                rsq = rsq + eps2;
                const double r = sqrt(rsq);
                zdiv |= r == 0.0;  // Numba raises ZeroDivisionError here
                const double rinv = 1.0 / r;
                const double ric = rinv * rinv * rinv;
                const double mag = G * q.w * ric;
                px = px + mag * dx;
                py = py + mag * dy;
                if (DIM == 3) pz = pz + mag * dz;
              }
            }
          }
          if (zdiv) {
            // the whole leaf again with the pure-Python loop (bh.py:384-425):
            // G m / (r2 + eps^2)^1.5, zero where r2 + eps^2 == 0
            px = py = pz = 0.0;
            for (int j = j0; j < j1; j++) {
              if (SELF && j == t) continue;
              const double4 q = spart[j];
              const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
              double r2 = dx * dx;
              r2 = r2 + dy * dy;
              if (DIM == 3) r2 = r2 + dz * dz;
              const double r2s = r2 + eps2;
              if (r2s == 0.0) {
                px = px + 0.0;
                py = py + 0.0;
                pz = pz + 0.0;
                continue;
              }
              const double f = (G * q.w) / (r2s * sqrt(r2s));
              px = px + f * dx;
              py = py + f * dy;
              if (DIM == 3) pz = pz + f * dz;
            }
          }
          if (mine) {
            ax = ax + px;
            ay = ay + py;
            if (DIM == 3) az = az + pz;
          }
        }
        if (MODE == WALK_POT) {
          // softened pair potential -G m / sqrt(r^2 + eps^2), zero where the
          // softened distance is zero; leaf partial added to the total
          double pp = 0.0;
          for (int j = j0; j < j1; j++) {
            const double4 q = spart[j];
            if (mine && !(SELF && j == t)) {
              const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
              double r2 = dx * dx;
              r2 = r2 + dy * dy;
              if (DIM == 3) r2 = r2 + dz * dz;
              const double r2s = r2 + eps2;
              if (r2s != 0.0) pp = pp - (G * q.w) / sqrt(r2s);
            }
          }
          if (mine) phi = phi + pp;
        }
        v = nd.skip;
        continue;
      }
      // opening test (bh.py:586-601): d = ||t - com|| as OpenBLAS ddot forms
      // it (x*x, then fma for y and z), open when d == 0 or size/d >= theta;
      // the middle-range square root and division when the node's size and
      // every visiting lane's d^2 allow them (then d > 0)
      double d2 = 0.0;
      if (mine) {
        const double ux = tx - nd.cx, uy = ty - nd.cy, uz = tz - nd.cz;
        d2 = ux * ux;
        d2 = fma(uy, uy, d2);
        if (DIM == 3) d2 = fma(uz, uz, d2);
      }
      const bool mid = nd.size >= 0x1p-300 && nd.size <= 0x1p300 && d2 >= 0x1p-700 && d2 <= 0x1p700;
      bool open = false;
      if (__ballot(mine && !mid) == 0) {
        if (mine) open = div_mid(nd.size, sqrt_mid(d2)) >= theta;
      } else if (mine) {
        const double dist = sqrt(d2);
        open = (dist == 0.0) || (nd.size / dist >= theta);
      }
      if (mine) {
        if (MODE == WALK_STATS) c_test++;
        if (open) {
          follow = dv + 1;
        } else {
          // monopole (kernels.py:57-69 via acc_monopole_node)
          if (MODE == WALK_LCOUNT) nl++;
          if (MODE == WALK_LFILL) lids[wpos++] = v;
          if (MODE == WALK_STATS) c_mono++;
          if (MODE == WALK_POT) {
            const double dx = nd.cx - tx, dy = nd.cy - ty, dz = nd.cz - tz;
            double r2 = dx * dx;
            r2 = r2 + dy * dy;
            if (DIM == 3) r2 = r2 + dz * dz;
            const double r2s = r2 + eps2;
            if (r2s != 0.0) phi = phi - (G * nd.m) / sqrt(r2s);
          }
          if (MODE == WALK_ACC) {
            const double dx = nd.cx - tx, dy = nd.cy - ty, dz = nd.cz - tz;
            double r2 = dx * dx;
            r2 = r2 + dy * dy;
            if (DIM == 3) r2 = r2 + dz * dz;
            const double r2s = r2 + eps2;
            if (r2s != 0.0) {
              // the middle-range square root and division where r2s and G m
              // allow them (then r2s * sqrt(r2s) lies in [2^-350, 2^350])
              const double gm = G * nd.m;
              double f;
              if (r2s >= 0x1p-233 && r2s <= 0x1p233 && gm >= 0x1p-300 && gm <= 0x1p300) {
                f = div_mid(gm, r2s * sqrt_mid(r2s));
              } else {
                f = gm / (r2s * sqrt(r2s));
              }
              ax = ax + f * dx;
              ay = ay + f * dy;
              if (DIM == 3) az = az + f * dz;
            } else {
              ax = ax + 0.0;
              ay = ay + 0.0;
              az = az + 0.0;
            }
          }
        }
      }
      v = __ballot(open) ? nd.pad[0] : nd.skip;
    }
  }
  if (!live) return;
  if (MODE == WALK_ACC) {
    out[orig * DIM + 0] = ax;
    out[orig * DIM + 1] = ay;
    if (DIM == 3) out[orig * DIM + 2] = az;
  }
  if (MODE == WALK_POT) out[orig] = phi;
  if (MODE == WALK_LCOUNT) stats[orig] = nl;
  if (MODE == WALK_STATS) {
    int64_t* o = stats + orig * 5;
    o[0] = c_pop;
    o[1] = c_test;
    o[2] = c_mono;
    o[3] = c_leaf;
    o[4] = c_pair;
  }
}

template <int DIM, bool SELF>
static void launch_walk(bh_ctx* c, int mode, const bh_tree* t, int64_t m, const double* tgt,
                        const uint32_t* torder, double theta, double eps2, double G, double* out,
                        const int64_t* lptr, int32_t* lids, int64_t* stats) {
  const dim3 grid((unsigned)cdiv(m, 256)), block(256);
#define BH_W(MODE)                                                                             \
  hipLaunchKernelGGL((k_walk<DIM, SELF, MODE>), grid, block, 0, c->stream, t->nodes, t->spart,  \
                     t->perm, m, tgt, torder, theta, eps2, G, out, lptr, lids, stats,  \
                     t->d_sys_poff, t->d_sys_noff, t->nsys)
  switch (mode) {
    case WALK_ACC: BH_W(WALK_ACC); break;
    case WALK_LCOUNT: BH_W(WALK_LCOUNT); break;
    case WALK_LFILL: BH_W(WALK_LFILL); break;
    case WALK_POT: BH_W(WALK_POT); break;
    default: BH_W(WALK_STATS); break;
  }
#undef BH_W
}

// Orders external targets along the tree's Morton curve (codes relative to
// the source box, clamped) so each wavefront walks nearby targets.  The
// workspace holds two key and one value buffer of m entries and the radix
// sort's counts (radix_tiles(m) * 256) and totals (256).
static int order_targets_ws(bh_ctx* c, const bh_tree* t, const double* tgt, int64_t m, uint32_t* torder,
                            uint64_t* k0, uint64_t* k1, uint32_t* v1, uint32_t* counts, uint32_t* totals) {
  int rc = morton_iota(c, tgt, m, t->dim, t->gbox, 1, k0, torder);
  if (rc) return rc;
  uint64_t* ko;
  uint32_t* vo;
  rc = radix_sort_ws(c, k0, torder, m, 21 * t->dim, &ko, &vo, k1, v1, counts, totals);
  if (rc) return rc;
  if (vo != torder) BH_HIP(hipMemcpyAsync(torder, vo, m * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
  return BH_OK;
}

int order_targets(bh_ctx* c, const bh_tree* t, const double* tgt, int64_t m, uint32_t* torder) {
  if (m == 0) return BH_OK;
  Scratch s(c);
  uint64_t* k0 = s.get<uint64_t>(m);
  uint64_t* k1 = s.get<uint64_t>(m);
  uint32_t* v1 = s.get<uint32_t>(m);
  uint32_t* counts = s.get<uint32_t>((size_t)256 * radix_tiles(m));
  uint32_t* totals = s.get<uint32_t>(256);
  if (!k0 || !k1 || !v1 || !counts || !totals) return BH_ENOMEM;
  return order_targets_ws(c, t, tgt, m, torder, k0, k1, v1, counts, totals);
}

static void launch_mode(bh_ctx* c, const bh_tree* t, int mode, const double* tgt, const uint32_t* torder, int64_t m,
                        double theta, double eps2, double G, double* out, const int64_t* lptr, int32_t* lids,
                        int64_t* stats) {
  if (t->dim == 3) {
    if (tgt) launch_walk<3, false>(c, mode, t, m, tgt, torder, theta, eps2, G, out, lptr, lids, stats);
    else launch_walk<3, true>(c, mode, t, m, tgt, torder, theta, eps2, G, out, lptr, lids, stats);
  } else {
    if (tgt) launch_walk<2, false>(c, mode, t, m, tgt, torder, theta, eps2, G, out, lptr, lids, stats);
    else launch_walk<2, true>(c, mode, t, m, tgt, torder, theta, eps2, G, out, lptr, lids, stats);
  }
}

// Runs one walk mode.  tgt == nullptr: self mode (m == n).  For list fill,
// lptr holds per-target offsets (original order); for list count and for
// counters, `stats` receives per-target values.
int walk_mode(bh_ctx* c, const bh_tree* t, int mode, const double* tgt, int64_t m, double theta,
              double eps, double G, double* out, const int64_t* lptr, int32_t* lids, int64_t* stats) {
  if (m == 0) return BH_OK;
  Scratch s(c);
  uint32_t* torder = nullptr;
  if (tgt) {
    torder = s.get<uint32_t>(m);
    if (!torder) return BH_ENOMEM;
    int rc = order_targets(c, t, tgt, m, torder);
    if (rc) return rc;
  }
  static const char* names[] = {"k_walk_f64", "k_walk_f64_lcount", "k_walk_f64_lfill", "k_walk_f64_stats",
                                "k_walk_f64_pot"};
  ProfScope ps(c, names[mode]);
  launch_mode(c, t, mode, tgt, torder, m, theta, eps * eps, G, out, lptr, lids, stats);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

// ---------------------------------------------------------------- no allocation
// Workspace of walk_targets_noalloc in one file-scope device array: the
// Morton ordering of at most NOALLOC_TARGETS external targets (two key
// buffers, the order and its radix partner, the radix sort's per-tile digit
// counts and digit totals), 24 B per target + 1 KiB per 4096 targets + 1 KiB
// = 24.25 MiB, held by every device the library runs on until the process
// ends.  Every region is written before it is read, so nothing needs
// clearing between calls; calls that overlap on one device share it and race.
constexpr int64_t NOALLOC_TARGETS = 1 << 20;
constexpr size_t NA_KEYS0 = 0;
constexpr size_t NA_KEYS1 = NA_KEYS0 + 8 * NOALLOC_TARGETS;
constexpr size_t NA_ORDER = NA_KEYS1 + 8 * NOALLOC_TARGETS;
constexpr size_t NA_VALS1 = NA_ORDER + 4 * NOALLOC_TARGETS;
constexpr size_t NA_COUNTS = NA_VALS1 + 4 * NOALLOC_TARGETS;
constexpr size_t NA_TOTALS = NA_COUNTS + 4 * 256 * (NOALLOC_TARGETS / RADIX_TILE);
constexpr size_t NA_BYTES = NA_TOTALS + 4 * 256;
static_assert(NOALLOC_TARGETS % RADIX_TILE == 0 && NA_TOTALS % 256 == 0, "workspace layout");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static unsigned char* g_na_ptr[NA_MAX_DEVICES] = {};

// Device address of `scratch` on `device`, which must be current: each
// device holds its own copy of the array at its own address, and host code
// passes this address, never the array's name, to kernels and copies.
int noalloc_setup(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) {
    set_error("device %d beyond the %d devices the no-allocation workspace supports", device, NA_MAX_DEVICES);
    return BH_ENODEV;
  }
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_ptr[device]) return BH_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
  if (!p) {
    set_error("hipGetSymbolAddress(scratch) returned NULL on device %d", device);
    return BH_EHIP;
  }
  g_na_ptr[device] = static_cast<unsigned char*>(p);
  return BH_OK;
}

// The plain external-target walk of walk_mode(WALK_ACC) with its workspace
// in `scratch`: no device allocation, no profiling events.
int walk_targets_noalloc(bh_ctx* c, const bh_tree* t, const double* tgt, int64_t m, double theta, double eps,
                         double G, double* out) {
  if (m > NOALLOC_TARGETS) {
    set_error("the no-allocation entry point takes at most %lld targets per call, got %lld",
              (long long)NOALLOC_TARGETS, (long long)m);
    return BH_EINVAL;
  }
  if (m == 0) return BH_OK;
  unsigned char* base;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    base = c->device < NA_MAX_DEVICES ? g_na_ptr[c->device] : nullptr;
  }
  if (!base) {
    set_error("no-allocation workspace not set up on device %d", c->device);
    return BH_EHIP;
  }
  uint64_t* k0 = reinterpret_cast<uint64_t*>(base + NA_KEYS0);
  uint64_t* k1 = reinterpret_cast<uint64_t*>(base + NA_KEYS1);
  uint32_t* torder = reinterpret_cast<uint32_t*>(base + NA_ORDER);
  uint32_t* v1 = reinterpret_cast<uint32_t*>(base + NA_VALS1);
  uint32_t* counts = reinterpret_cast<uint32_t*>(base + NA_COUNTS);
  uint32_t* totals = reinterpret_cast<uint32_t*>(base + NA_TOTALS);
  BH_TRY(order_targets_ws(c, t, tgt, m, torder, k0, k1, v1, counts, totals));
  launch_mode(c, t, WALK_ACC, tgt, torder, m, theta, eps * eps, G, out, nullptr, nullptr, nullptr);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

}  // namespace bh
