// bh_build.hip -- tree topology, upward pass (geometry, mass, centre of mass:
// one launch per depth for trees built here, an atomic-flag climb for
// imported ones), traversal records, refit check.
//
// Reference: _build_tree_iterative (bhut/tree/build.py:110-212),
// _compute_node_geometry (build.py:215-242), _compute_mass_and_com
// (build.py:245-295), refit_tree (bhut/tree/refit.py:17-192),
// should_refit_vs_rebuild (refit.py:195-251).
#include <cstdarg>
#include <cstdio>

#include "bh_common.h"

namespace bh {

// ---------------------------------------------------------------- topology
// The median-split tree depends only on (N, leaf).  Every node at depth d
// holds floor(N/2^d) or floor(N/2^d)+1 particles, so the number of leaves
// (and internal nodes) under any node is a table lookup by (depth, count).
// Node numbering follows the LIFO stack of build.py:138-186: the children of
// the r-th internal node in pop order are 1+2r (left) and 2+2r (right), and
// the pop order is the right-first preorder, so
//   r(right child) = r(v) + 1,   r(left child) = r(v) + 1 + internal(right).
// Each thread evaluates one (depth, position) slot by walking its root path.
struct TopoTables {
  int64_t q[MAX_DEPTH + 1];           // floor(N / 2^d)
  int64_t leaves[MAX_DEPTH + 1][2];   // leaves under a node of count q[d] + {0,1}
};

__device__ __forceinline__ int64_t tb_leaves(const TopoTables* tb, int d, int64_t c) {
  return tb->leaves[d][c - tb->q[d]];
}

// One tree of a forest: its particle count and where its slots, nodes,
// particles and leaves start in the flat arrays.  A single tree is the
// one-system forest with zero offsets.
struct SysInfo {
  int64_t N, slot_off, noff, poff, loff;
};

__global__ __launch_bounds__(256) void k_topology(int64_t nsys, int64_t L, int64_t nslots,
                                                  const SysInfo* __restrict__ sys,
                                                  const TopoTables* __restrict__ tbs,
                                                  int32_t* __restrict__ start,
                                                  int32_t* __restrict__ count,
                                                  int32_t* __restrict__ first_child,
                                                  int32_t* __restrict__ parent,
                                                  int32_t* __restrict__ skip,
                                                  int32_t* __restrict__ ndepth,
                                                  int32_t* __restrict__ leaf_order,
                                                  int32_t* __restrict__ slot_node) {
  for (int64_t gslot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; gslot < nslots;
       gslot += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = nsys - 1;  // last system whose slots start at or before gslot
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) / 2;
      if (sys[mid].slot_off <= gslot) lo = mid;
      else hi = mid - 1;
    }
    const SysInfo S = sys[lo];
    const TopoTables* tb = tbs + lo;
    const int64_t slot = gslot - S.slot_off;
    const int64_t N = S.N;
    const int k = 63 - __clzll((unsigned long long)(slot + 1));  // depth
    const int64_t j = slot + 1 - (1ll << k);                     // position at depth k
    int64_t c = N, s = 0, r = 0, lrank = 0;
    int64_t idx = 0, par = -1, sk = -1;
    bool ok = true;
    for (int t = 1; t <= k; t++) {
      if (c <= L) {
        ok = false;
        break;
      }
      const int64_t half = c / 2;
      const int bit = (int)((j >> (k - t)) & 1);
      int64_t nidx;
      if (bit == 0) {  // left child: popped after the whole right subtree
        nidx = 1 + 2 * r;
        r = r + 1 + (tb_leaves(tb, t, c - half) - 1);
        c = half;
        // skip of a left child is its parent's skip (unchanged)
      } else {  // right child: popped right after its parent
        nidx = 2 + 2 * r;
        lrank += tb_leaves(tb, t, half);
        r = r + 1;
        s += half;
        c -= half;
        sk = nidx - 1;  // then the left sibling
      }
      par = idx;
      idx = nidx;
    }
    if (!ok) {
      slot_node[gslot] = -1;  // below a leaf
      continue;
    }
    const int64_t gi = idx + S.noff;
    slot_node[gslot] = (int32_t)gi;
    start[gi] = (int32_t)(s + S.poff);
    count[gi] = (int32_t)c;
    parent[gi] = par >= 0 ? (int32_t)(par + S.noff) : -1;
    skip[gi] = sk >= 0 ? (int32_t)(sk + S.noff) : -1;
    ndepth[gi] = k;
    if (c > L) {
      first_child[gi] = (int32_t)(1 + 2 * r + S.noff);
    } else {
      first_child[gi] = -1;
      leaf_order[S.loff + lrank] = (int32_t)gi;
    }
  }
}

static int topo_tables(int64_t N, int64_t L, TopoTables& tb, int& D) {
  D = 0;
  tb.q[0] = N;
  while (tb.q[D] + 1 > L) {  // some node at depth D may still split
    if (D + 1 > MAX_DEPTH) {
      set_error("tree deeper than %d levels", MAX_DEPTH);
      return BH_EOVERFLOW;
    }
    tb.q[D + 1] = tb.q[D] / 2;
    D++;
  }
  for (int d = D; d >= 0; d--) {
    for (int e = 0; e < 2; e++) {
      const int64_t v = tb.q[d] + e;
      if (v <= L || d == D) {
        tb.leaves[d][e] = 1;
      } else {
        const int64_t a = v / 2, b = v - v / 2;
        tb.leaves[d][e] = tb.leaves[d + 1][a - tb.q[d + 1]] + tb.leaves[d + 1][b - tb.q[d + 1]];
      }
    }
  }
  for (int d = D + 1; d <= MAX_DEPTH; d++) {
    tb.q[d] = 0;
    tb.leaves[d][0] = tb.leaves[d][1] = 1;
  }
  return BH_OK;
}

// Node and leaf counts of the (N, L) tree; node budget of build.py:124 is
// never reached by this split (leaves hold more than L/2 particles).
int topology_plan(int64_t N, int64_t L, int64_t* M, int64_t* nleaves, int* depth) {
  TopoTables tb;
  int D;
  int rc = topo_tables(N, L, tb, D);
  if (rc) return rc;
  const int64_t nl = N <= L ? 1 : tb.leaves[0][0];
  *nleaves = nl;
  *M = 2 * nl - 1;
  *depth = D;
  const int64_t budget = std::max<int64_t>(4 * N / L, 10);
  if (*M > budget || *M > 0x7fffffffll) {
    set_error("node count %lld exceeds the node budget %lld", (long long)*M, (long long)budget);
    return BH_EOVERFLOW;
  }
  return BH_OK;
}

// Plan of a forest whose systems hold counts[0..nsys): per-system node and
// leaf offsets (host) and totals.  A single tree is nsys = 1.
int forest_plan(const int64_t* counts, int64_t nsys, int64_t L, std::vector<int64_t>& noff,
                std::vector<int64_t>& loff, int64_t* M, int64_t* nleaves, int* depth) {
  noff.assign(nsys + 1, 0);
  loff.assign(nsys + 1, 0);
  int dmax = 0;
  for (int64_t s = 0; s < nsys; s++) {
    int64_t Ms, ls;
    int D;
    int rc = topology_plan(counts[s], L, &Ms, &ls, &D);
    if (rc) return rc;
    noff[s + 1] = noff[s] + Ms;
    loff[s + 1] = loff[s] + ls;
    dmax = std::max(dmax, D);
  }
  if (noff[nsys] > 0x7fffffffll) {
    set_error("forest of %lld nodes exceeds the int32 index range", (long long)noff[nsys]);
    return BH_EOVERFLOW;
  }
  *M = noff[nsys];
  *nleaves = loff[nsys];
  *depth = dmax;
  return BH_OK;
}

// A single tree's tables, stored from the kernel arguments (no host copy to
// wait for).
__global__ void k_put_tables(TopoTables tb, SysInfo si, TopoTables* __restrict__ d_tb, SysInfo* __restrict__ d_si) {
  if (threadIdx.x == 0) {
    *d_tb = tb;
    *d_si = si;
  }
}

int topology(bh_ctx* c, bh_tree* t) {
  const int64_t L = t->leaf, nsys = t->nsys;
  std::vector<TopoTables> tbs(nsys);
  std::vector<SysInfo> info(nsys);
  std::vector<int> depth(nsys);
  int64_t slots = 0;
  for (int64_t s = 0; s < nsys; s++) {
    const int64_t N = t->sys_poff[s + 1] - t->sys_poff[s];
    int D = 0;
    int rc = topo_tables(N, L, tbs[s], D);
    if (rc) return rc;
    depth[s] = D;
    info[s].N = N;
    info[s].slot_off = slots;
    info[s].noff = t->sys_noff[s];
    info[s].poff = t->sys_poff[s];
    info[s].loff = t->sys_loff[s];
    slots += (2ll << D) - 1;
  }
  // depth tables of the level-by-level upward passes: per depth, the systems
  // deep enough (and, for a forest, each system's first slot)
  dfree(c, t->slot_node);
  dfree(c, t->d_lvl_sys);
  dfree(c, t->d_slot_off);
  t->slot_node = nullptr;
  t->d_lvl_sys = nullptr;
  t->d_slot_off = nullptr;
  t->dmax = 0;
  for (int64_t s = 0; s < nsys; s++) t->dmax = std::max(t->dmax, depth[s]);
  t->lvl_nsys.assign(t->dmax + 1, 0);
  t->h_lvl_off.assign(t->dmax + 2, 0);
  std::vector<int32_t> lvl;
  std::vector<int64_t> soff(nsys);
  for (int d = 0; d <= t->dmax; d++) {
    t->h_lvl_off[d] = (int64_t)lvl.size();
    for (int64_t s = 0; s < nsys; s++)
      if (depth[s] >= d) {
        lvl.push_back((int32_t)s);
        t->lvl_nsys[d]++;
      }
  }
  t->h_lvl_off[t->dmax + 1] = (int64_t)lvl.size();
  for (int64_t s = 0; s < nsys; s++) soff[s] = info[s].slot_off;
  t->slot_node = dalloc_t<int32_t>(c, (size_t)slots);
  if (!t->slot_node) return BH_ENOMEM;
  if (nsys > 1) {
    t->d_lvl_sys = dalloc_t<int32_t>(c, lvl.size());
    t->d_slot_off = dalloc_t<int64_t>(c, (size_t)nsys);
    if (!t->d_lvl_sys || !t->d_slot_off) return BH_ENOMEM;
    BH_HIP(hipMemcpyAsync(t->d_lvl_sys, lvl.data(), lvl.size() * 4, hipMemcpyHostToDevice, c->stream));
    BH_HIP(hipMemcpyAsync(t->d_slot_off, soff.data(), nsys * 8, hipMemcpyHostToDevice, c->stream));
  }
  Scratch sc(c);
  TopoTables* d_tb = sc.get<TopoTables>(nsys);
  SysInfo* d_info = sc.get<SysInfo>(nsys);
  if (!d_tb || !d_info) return BH_ENOMEM;
  if (nsys == 1) {
    hipLaunchKernelGGL(k_put_tables, dim3(1), dim3(64), 0, c->stream, tbs[0], info[0], d_tb, d_info);
    BH_LAUNCH_CHECK();
  } else {
    BH_HIP(hipMemcpyAsync(d_tb, tbs.data(), nsys * sizeof(TopoTables), hipMemcpyHostToDevice, c->stream));
    BH_HIP(hipMemcpyAsync(d_info, info.data(), nsys * sizeof(SysInfo), hipMemcpyHostToDevice, c->stream));
  }
  const int nb = (int)std::min<int64_t>(cdiv(slots, 256), 65536);
  hipLaunchKernelGGL(k_topology, dim3(nb), dim3(256), 0, c->stream, nsys, L, slots, d_info, d_tb, t->start,
                     t->count, t->first_child, t->parent, t->skip, t->ndepth, t->leaf_order, t->slot_node);
  BH_LAUNCH_CHECK();
  // the host tables of a forest must outlive the asynchronous copies
  if (nsys > 1) BH_HIP(hipStreamSynchronize(c->stream));
  return BH_OK;
}

// ---------------------------------------------------------------- NumPy order
// np.sum of a contiguous float64 run (build.py:263, refit.py:140): the
// reduction adds the pairwise sum to the +0.0 identity.  Pairwise: from -0.0
// sequentially below 8 elements, 8 interleaved accumulators up to 128, then
// halving on multiples of 8.
__device__ double pw_block(const double4* __restrict__ p, int64_t n) {
  if (n < 8) {
    double r = -0.0;
    for (int64_t i = 0; i < n; i++) r += p[i].w;
    return r;
  }
  double r0 = p[0].w, r1 = p[1].w, r2 = p[2].w, r3 = p[3].w;
  double r4 = p[4].w, r5 = p[5].w, r6 = p[6].w, r7 = p[7].w;
  int64_t i = 8;
  const int64_t lim = n - (n % 8);
  for (; i < lim; i += 8) {
    r0 += p[i].w;
    r1 += p[i + 1].w;
    r2 += p[i + 2].w;
    r3 += p[i + 3].w;
    r4 += p[i + 4].w;
    r5 += p[i + 5].w;
    r6 += p[i + 6].w;
    r7 += p[i + 7].w;
  }
  double res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (; i < n; i++) res += p[i].w;
  return res;
}

__device__ double np_pairwise_mass(const double4* __restrict__ p, int64_t n) {
  if (n <= 128) return pw_block(p, n);
  // the recursive split, evaluated with an explicit stack
  int64_t off[48], len[48];
  double left[48];
  int state[48];
  int sp = 0;
  off[0] = 0;
  len[0] = n;
  state[0] = 0;
  double ret = 0.0;
  while (sp >= 0) {
    if (state[sp] == 0) {
      if (len[sp] <= 128) {
        ret = pw_block(p + off[sp], len[sp]);
        sp--;
        continue;
      }
      int64_t n2 = len[sp] / 2;
      n2 -= n2 % 8;
      state[sp] = 1;
      off[sp + 1] = off[sp];
      len[sp + 1] = n2;
      state[sp + 1] = 0;
      sp++;
    } else if (state[sp] == 1) {
      left[sp] = ret;
      state[sp] = 2;
      int64_t n2 = len[sp] / 2;
      n2 -= n2 % 8;
      off[sp + 1] = off[sp] + n2;
      len[sp + 1] = len[sp] - n2;
      state[sp + 1] = 0;
      sp++;
    } else {
      ret = left[sp] + ret;
      sp--;
    }
  }
  return ret;
}

__device__ __forceinline__ double comp(const double4& p, int d) {
  return d == 0 ? p.x : (d == 1 ? p.y : p.z);
}

// Release this thread's node writes and count the arrival at `node`; true
// for the second arriver, which then sees both children's data.
__device__ __forceinline__ bool arrive(uint32_t* flags, int node) {
  const uint32_t old =
      __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  return old != 0;
}

// Leaf v: tight box (build.py:232-242), mass and centre of mass in NumPy
// order (build.py:258-272); REFIT keeps the box and centre.  BIG: the tree
// has leaves of more than 128 particles, whose masses take NumPy's
// recursive pairwise split (its explicit stack is per-lane private memory,
// so trees without such leaves use kernels that do not carry it).
template <int DIM, bool REFIT, bool BIG>
__device__ __forceinline__ void up_leaf(int v, const int32_t* __restrict__ start, const int32_t* __restrict__ count,
                                        const double4* __restrict__ spart, double* center, double* half,
                                        double* mass, double* com, double* bmin, double* bmax) {
  const int64_t s = start[v], c = count[v];
  const double4* p = spart + s;
  double ctr[DIM];
  if (!REFIT) {
    double lo[DIM], hi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) lo[d] = hi[d] = comp(p[0], d);
    for (int64_t i = 1; i < c; i++) {
      const double4 q = p[i];
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        lo[d] = np_min(lo[d], comp(q, d));
        hi[d] = np_max(hi[d], comp(q, d));
      }
    }
    double em = hi[0] - lo[0];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      ctr[d] = (lo[d] + hi[d]) / 2;
      if (d > 0) em = np_max(em, hi[d] - lo[d]);
      center[v * DIM + d] = ctr[d];
      bmin[v * DIM + d] = lo[d];
      bmax[v * DIM + d] = hi[d];
    }
    half[v] = em / 2;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; d++) ctr[d] = center[v * DIM + d];
  }
  double total = 0.0;
  if (BIG) {  // np.sum's buffers of 8192 elements, each summed pairwise
    for (int64_t s0 = 0; s0 < c; s0 += 8192) total += np_pairwise_mass(p + s0, c - s0 < 8192 ? c - s0 : 8192);
  } else {
    total = 0.0 + pw_block(p, c);
  }
  mass[v] = total;
  if (total > 0) {
    double sx[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) sx[d] = 0.0;
    for (int64_t i = 0; i < c; i++) {
      const double4 q = p[i];
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const double w = q.w * comp(q, d);
        sx[d] = sx[d] + w;
      }
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) com[v * DIM + d] = sx[d] / total;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; d++) com[v * DIM + d] = ctr[d];
  }
}

// Internal node from its two children (build.py:274-295, or
// refit.py:152-192 when REFIT, which keeps centre/half_size and adds only
// children with mass > 0).
template <int DIM, bool REFIT>
__device__ __forceinline__ void up_node(int node, const int32_t* __restrict__ first_child, double* center,
                                        double* half, double* mass, double* com, double* bmin, double* bmax) {
  const int a = first_child[node], b = a + 1;
  double cc[DIM];
  if (!REFIT) {
    double em = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const double lo = np_min(bmin[a * DIM + d], bmin[b * DIM + d]);
      const double hi = np_max(bmax[a * DIM + d], bmax[b * DIM + d]);
      bmin[node * DIM + d] = lo;
      bmax[node * DIM + d] = hi;
      cc[d] = (lo + hi) / 2;
      center[node * DIM + d] = cc[d];
      em = d == 0 ? hi - lo : np_max(em, hi - lo);
    }
    half[node] = em / 2;
    const double ma = mass[a], mb = mass[b];
    double tsum = -0.0;
    tsum += ma;
    tsum += mb;
    const double tot = 0.0 + tsum;
    mass[node] = tot;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      double sd = 0.0;
      const double wa = ma * com[a * DIM + d];
      sd = sd + wa;
      const double wb = mb * com[b * DIM + d];
      sd = sd + wb;
      com[node * DIM + d] = tot > 0 ? sd / tot : cc[d];
    }
  } else {
#pragma unroll
    for (int d = 0; d < DIM; d++) cc[d] = center[node * DIM + d];
    const double ma = mass[a], mb = mass[b];
    double tot = 0.0, wc[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) wc[d] = 0.0;
    if (ma > 0) {
      tot += ma;
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const double w = ma * com[a * DIM + d];
        wc[d] = wc[d] + w;
      }
    }
    if (mb > 0) {
      tot += mb;
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const double w = mb * com[b * DIM + d];
        wc[d] = wc[d] + w;
      }
    }
    mass[node] = tot;
#pragma unroll
    for (int d = 0; d < DIM; d++) com[node * DIM + d] = tot > 0 ? wc[d] / tot : cc[d];
  }
}

// Trees without depth tables (imported): one thread per leaf, leaves in
// left-to-right (memory) order; after its leaf a thread climbs, and the
// second thread to reach a parent combines its two children.
template <int DIM, bool REFIT, bool BIG>
__global__ __launch_bounds__(256) void k_leaf_up(
    int64_t nleaves, const int32_t* __restrict__ leaf_order, const int32_t* __restrict__ start,
    const int32_t* __restrict__ count, const int32_t* __restrict__ first_child,
    const int32_t* __restrict__ parent, const double4* __restrict__ spart, double* center,
    double* half, double* mass, double* com, double* bmin, double* bmax, uint32_t* flags) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nleaves) return;
  const int v = leaf_order[t];
  up_leaf<DIM, REFIT, BIG>(v, start, count, spart, center, half, mass, com, bmin, bmax);
  int node = parent[v];
  while (node >= 0) {
    if (!arrive(flags, node)) return;
    up_node<DIM, REFIT>(node, first_child, center, half, mass, com, bmin, bmax);
    node = parent[node];
  }
}

// Trees built here: one launch per depth, deepest first, so a node's
// children are complete when it is combined (no atomics, no fences).
template <int DIM, bool REFIT, bool BIG>
__global__ __launch_bounds__(256) void k_up_level(
    int d, int64_t nthreads, const int32_t* __restrict__ lvl_sys, const int64_t* __restrict__ slot_off,
    const int32_t* __restrict__ slot_node, const int32_t* __restrict__ start, const int32_t* __restrict__ count,
    const int32_t* __restrict__ first_child, const double4* __restrict__ spart, double* center, double* half,
    double* mass, double* com, double* bmin, double* bmax) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  const int v = level_node(d, t, lvl_sys, slot_off, slot_node);
  if (v < 0) return;
  if (first_child[v] < 0) up_leaf<DIM, REFIT, BIG>(v, start, count, spart, center, half, mass, com, bmin, bmax);
  else up_node<DIM, REFIT>(v, first_child, center, half, mass, com, bmin, bmax);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int fill_u32(bh_ctx* c, uint32_t* p, uint32_t v, int64_t n) {
  if (n <= 0) return BH_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, c->stream, p, v, n);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

int upward(bh_ctx* c, bh_tree* t, int refit) {
  const bool big = t->max_leaf > 128;
  if (t->slot_node) {
    for (int d = t->dmax; d >= 0; d--) {
      const int64_t nthr = t->lvl_nsys[d] << d;
      const int32_t* ls = t->d_lvl_sys ? t->d_lvl_sys + t->h_lvl_off[d] : nullptr;
      const unsigned nb = (unsigned)cdiv(nthr, 256);
#define BH_UPL(D, R, B)                                                                                     \
  hipLaunchKernelGGL((k_up_level<D, R, B>), dim3(nb), dim3(256), 0, c->stream, d, nthr, ls, t->d_slot_off,  \
                     t->slot_node, t->start, t->count, t->first_child, t->spart, t->center, t->half, t->mass, \
                     t->com, t->bmin, t->bmax)
      if (big) {
        if (t->dim == 3) {
          if (refit) BH_UPL(3, true, true); else BH_UPL(3, false, true);
        } else {
          if (refit) BH_UPL(2, true, true); else BH_UPL(2, false, true);
        }
      } else {
        if (t->dim == 3) {
          if (refit) BH_UPL(3, true, false); else BH_UPL(3, false, false);
        } else {
          if (refit) BH_UPL(2, true, false); else BH_UPL(2, false, false);
        }
      }
#undef BH_UPL
      BH_LAUNCH_CHECK();
    }
    return BH_OK;
  }
  Scratch s(c);
  uint32_t* flags = s.get<uint32_t>((size_t)t->M);
  if (!flags) return BH_ENOMEM;
  int rc = fill_u32(c, flags, 0u, t->M);
  if (rc) return rc;
  const unsigned nb = (unsigned)cdiv(t->nleaves, 256);
#define BH_UP(D, R, B)                                                                         \
  hipLaunchKernelGGL((k_leaf_up<D, R, B>), dim3(nb), dim3(256), 0, c->stream, t->nleaves,       \
                     t->leaf_order, t->start, t->count, t->first_child, t->parent, t->spart,    \
                     t->center, t->half, t->mass, t->com, t->bmin, t->bmax, flags)
  if (big) {
    if (t->dim == 3) {
      if (refit) BH_UP(3, true, true); else BH_UP(3, false, true);
    } else {
      if (refit) BH_UP(2, true, true); else BH_UP(2, false, true);
    }
  } else {
    if (t->dim == 3) {
      if (refit) BH_UP(3, true, false); else BH_UP(3, false, false);
    } else {
      if (refit) BH_UP(2, true, false); else BH_UP(2, false, false);
    }
  }
#undef BH_UP
  BH_LAUNCH_CHECK();
  return BH_OK;
}

// ---------------------------------------------------------------- records
template <int DIM>
__global__ __launch_bounds__(256) void k_pack(int64_t M, const double* __restrict__ half,
                                              const double* __restrict__ mass,
                                              const double* __restrict__ com,
                                              const int32_t* __restrict__ first_child,
                                              const int32_t* __restrict__ skip,
                                              const int32_t* __restrict__ ndepth,
                                              const int32_t* __restrict__ start,
                                              const int32_t* __restrict__ count,
                                              NodeD* __restrict__ nodes) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= M) return;
  NodeD r;
  r.cx = com[k * DIM + 0];
  r.cy = com[k * DIM + 1];
  r.cz = DIM == 3 ? com[k * DIM + 2] : 0.0;
  r.m = mass[k];
  r.size = 2 * half[k];
  r.fc = first_child[k];
  r.skip = skip[k];
  r.start = start[k];
  r.count = count[k];
  r.pad[0] = r.fc >= 0 ? r.fc + 1 : r.skip;  // node popped first when opened
  r.pad[1] = ndepth[k];
  nodes[k] = r;
}

int pack_nodes(bh_ctx* c, bh_tree* t) {
  const unsigned nb = (unsigned)cdiv(t->M, 256);
  if (t->dim == 3)
    hipLaunchKernelGGL(k_pack<3>, dim3(nb), dim3(256), 0, c->stream, t->M, t->half, t->mass, t->com,
                       t->first_child, t->skip, t->ndepth, t->start, t->count, t->nodes);
  else
    hipLaunchKernelGGL(k_pack<2>, dim3(nb), dim3(256), 0, c->stream, t->M, t->half, t->mass, t->com,
                       t->first_child, t->skip, t->ndepth, t->start, t->count, t->nodes);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

// ---------------------------------------------------------------- refit check
// should_refit_vs_rebuild (refit.py:236-249): rebuild when any particle's
// max-norm distance from the stored root centre exceeds 1.1 root half sizes.
template <int DIM>
__global__ __launch_bounds__(256) void k_refit_check(const double* __restrict__ pos, int64_t n,
                                                     const double* __restrict__ center,
                                                     const double* __restrict__ half,
                                                     uint32_t* __restrict__ outside) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool out = false;
  if (i < n) {
    const double lim = 1.1 * half[0];
    double md = fabs(pos[i * DIM] - center[0]);
#pragma unroll
    for (int d = 1; d < DIM; d++) md = np_max(md, fabs(pos[i * DIM + d] - center[d]));
    out = md > lim;
  }
  if (__ballot(out) != 0ull && __lane_id() == 0) atomicOr(outside, 1u);
}

int refit_check(bh_ctx* c, const bh_tree* t, const double* pos, int64_t n, int* ok) {
  Scratch s(c);
  uint32_t* flag = s.get<uint32_t>(1);
  if (!flag) return BH_ENOMEM;
  int rc = fill_u32(c, flag, 0u, 1);
  if (rc) return rc;
  const unsigned nb = (unsigned)cdiv(n, 256);
  if (n > 0) {
    if (t->dim == 3)
      hipLaunchKernelGGL(k_refit_check<3>, dim3(nb), dim3(256), 0, c->stream, pos, n, t->center,
                         t->half, flag);
    else
      hipLaunchKernelGGL(k_refit_check<2>, dim3(nb), dim3(256), 0, c->stream, pos, n, t->center,
                         t->half, flag);
    BH_LAUNCH_CHECK();
  }
  uint32_t h = 0;
  BH_HIP(hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  *ok = h == 0;
  return BH_OK;
}

}  // namespace bh
