// bh_group.hip -- the FP32 grouped form (the chosen, fast form).
//
// Same tree and the same leaf rule as bhut (bhut/traverse/bh.py:485-548:
// leaves are always summed particle by particle, internal nodes are accepted
// or opened), evaluated for groups of GS Morton-contiguous targets at once:
//
//  1. k_group_box: FP64 bounding box of each group's targets.
//  2. k_group_walk: one wavefront per group walks the tree breadth-wise with
//     an LDS stack: up to 64 nodes are tested per step, one per lane, with a
//     conservative group test -- accept a node if size < theta * dmin, dmin
//     the distance from its centre of mass to the group box, or, when that
//     box test opens, if no target of the group lies within size/theta of the
//     centre of mass (exact test: the lanes hold the targets and measure one
//     candidate node at a time).  Either way every target of the group would
//     also accept the node under bhut's own test (bh.py:586-601).  Lanes are compacted
//     with ballots into the cell list, the leaf list and the stack.  One
//     pass writes the lists into per-group regions and counts them; a count
//     pass and a packed fill are the fallback when a region is too small.
//  3. k_group_eval: persistent wavefronts over work items (a group's list
//     cut into pieces of at most 64 tiles), one (or two) targets per lane.
//     Tiles of 64 cells, or of 64 sources of the item's leaf chunks taken as
//     one compacted stream, are loaded cooperatively (coalesced) into one
//     LDS tile per wavefront and read back as wave-wide broadcasts; the next
//     tile's HBM reads are in flight in registers while the current tile is
//     consumed.  k_item_reduce adds each target's item partials in item
//     order.  All interactions are
//     FP32 on coordinates relative to the group centre: sources are stored
//     relative to their leaf's centre of mass and shifted by the FP64-formed
//     leaf-to-group offset, so precision follows the leaf size, not the
//     coordinate size.  A tile whose sources share one mass skips the
//     per-pair mass product and scales its sum by the mass once.  Each tile's
//     partial sum is added to the per-target total with Kahan compensation.
//
// Accepted cells carry their mass, centre of mass and second moments
// (quadrupole, accumulated in FP64 by k_quad_level one depth at a time, or
// k_quad_up for imported trees); the monopole-only variant
// is kept as an option.  FP64 stays in the tree build, the moments, the group
// box and the centre/shift subtractions.
#include "bh_common.h"

namespace bh {

struct GroupBox {
  double cx, cy, cz, hx, hy, hz;
};

// List entries are node indices: an accepted cell (its centre of mass, mass
// and second moments are read from the node record and the FP32 moment
// payload when the cell is evaluated), or a leaf (its particle range and
// centre of mass likewise).  4 B per entry keeps the lists small next to the
// tree, so the walk writes and the evaluation reads little HBM.
using CellEnt = int32_t;
using ChunkEnt = int32_t;

// ---------------------------------------------------------------- prepare
// Sources relative to their leaf's centre of mass (FP32) and each
// particle's leaf.  The centre of mass is the one the node record carries, so
// the walk forms the leaf-to-group shift from the record it already holds.
template <int DIM>
__global__ __launch_bounds__(256) void k_src32(int64_t nleaves, const int32_t* __restrict__ leaf_order,
                                               const int32_t* __restrict__ start,
                                               const int32_t* __restrict__ count,
                                               const double* __restrict__ com,
                                               const double4* __restrict__ spart,
                                               float4* __restrict__ src32, int32_t* __restrict__ pleaf) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nleaves) return;
  const int v = leaf_order[t];
  const double cx = com[v * DIM], cy = com[v * DIM + 1], cz = DIM == 3 ? com[v * DIM + 2] : 0.0;
  const int64_t s = start[v], c = count[v];
  for (int64_t j = s; j < s + c; j++) {
    const double4 p = spart[j];
    src32[j] = make_float4((float)(p.x - cx), (float)(p.y - cy), (float)(p.z - cz), (float)p.w);
    pleaf[j] = v;
  }
}

// Second moments Q = sum m (x - com)(x - com)^T, FP64: leaves from their
// particles, internal nodes by the parallel-axis shift of their children's
// moments.
__device__ __forceinline__ void quad_leaf(int v, const int32_t* __restrict__ start, const int32_t* __restrict__ count,
                                          const double4* spart, const double* com, int dim, double* quad) {
  const double cx = com[v * dim], cy = com[v * dim + 1], cz = dim == 3 ? com[v * dim + 2] : 0.0;
  double q[6] = {0, 0, 0, 0, 0, 0};
  const int64_t s = start[v], c = count[v];
  for (int64_t j = s; j < s + c; j++) {
    const double4 p = spart[j];
    const double dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
    q[0] += p.w * dx * dx;
    q[1] += p.w * dy * dy;
    q[2] += p.w * dz * dz;
    q[3] += p.w * dx * dy;
    q[4] += p.w * dx * dz;
    q[5] += p.w * dy * dz;
  }
  for (int k = 0; k < 6; k++) quad[(int64_t)v * 6 + k] = q[k];
}

__device__ __forceinline__ void quad_node(int node, const int32_t* __restrict__ first_child, const double* com,
                                          int dim, const double* mass, double* quad) {
  const double cx = com[node * dim], cy = com[node * dim + 1], cz = dim == 3 ? com[node * dim + 2] : 0.0;
  double q[6] = {0, 0, 0, 0, 0, 0};
  const int a = first_child[node];
  for (int ch = a; ch < a + 2; ch++) {
    const double m = mass[ch];
    const double dx = com[ch * dim] - cx, dy = com[ch * dim + 1] - cy, dz = dim == 3 ? com[ch * dim + 2] - cz : 0.0;
    const double* qc = quad + (int64_t)ch * 6;
    q[0] += qc[0] + m * dx * dx;
    q[1] += qc[1] + m * dy * dy;
    q[2] += qc[2] + m * dz * dz;
    q[3] += qc[3] + m * dx * dy;
    q[4] += qc[4] + m * dx * dz;
    q[5] += qc[5] + m * dy * dz;
  }
  for (int k = 0; k < 6; k++) quad[(int64_t)node * 6 + k] = q[k];
}

// Trees without depth tables (imported): the atomic-flag climb of the
// build's upward pass.
__device__ __forceinline__ bool arrive_q(uint32_t* flags, int node) {
  return __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__global__ __launch_bounds__(256) void k_quad_up(int64_t nleaves, const int32_t* __restrict__ leaf_order,
                                                 const int32_t* __restrict__ start,
                                                 const int32_t* __restrict__ count,
                                                 const int32_t* __restrict__ first_child,
                                                 const int32_t* __restrict__ parent, const double4* spart,
                                                 const double* com, int dim, const double* mass, double* quad,
                                                 uint32_t* flags) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nleaves) return;
  const int v = leaf_order[t];
  quad_leaf(v, start, count, spart, com, dim, quad);
  int node = parent[v];
  while (node >= 0) {
    if (!arrive_q(flags, node)) return;
    quad_node(node, first_child, com, dim, mass, quad);
    node = parent[node];
  }
}

// Trees built here: one launch per depth, deepest first (the slot tables of
// the build's level-by-level upward pass).
__global__ __launch_bounds__(256) void k_quad_level(int d, int64_t nthreads, const int32_t* __restrict__ lvl_sys,
                                                    const int64_t* __restrict__ slot_off,
                                                    const int32_t* __restrict__ slot_node,
                                                    const int32_t* __restrict__ start,
                                                    const int32_t* __restrict__ count,
                                                    const int32_t* __restrict__ first_child, const double4* spart,
                                                    const double* com, int dim, const double* mass, double* quad) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  const int v = level_node(d, t, lvl_sys, slot_off, slot_node);
  if (v < 0) return;
  if (first_child[v] < 0) quad_leaf(v, start, count, spart, com, dim, quad);
  else quad_node(v, first_child, com, dim, mass, quad);
}

// Accepted-cell payload of every node in the layout the cell list uses:
// (Qxx, Qyy, Qzz, trace) and (Qxy, Qxz, Qyz, 0) in FP32.
__global__ __launch_bounds__(256) void k_quad32(int64_t M, const double* __restrict__ quad, float4* __restrict__ q32) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= M) return;
  const double* q = quad + k * 6;
  q32[2 * k] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)(q[0] + q[1] + q[2]));
  q32[2 * k + 1] = make_float4((float)q[3], (float)q[4], (float)q[5], 0.f);
}

int f32_prepare(bh_ctx* c, bh_tree* t) {
  if (!t->spart32) t->spart32 = dalloc_t<float4>(c, (size_t)t->n);
  if (!t->pleaf) t->pleaf = dalloc_t<int32_t>(c, (size_t)t->n);
  if (!t->quad) t->quad = dalloc_t<double>(c, (size_t)t->M * 6);
  if (!t->quad32) t->quad32 = dalloc_t<float4>(c, (size_t)t->M * 2);
  if (!t->spart32 || !t->pleaf || !t->quad || !t->quad32) return BH_ENOMEM;
  const unsigned nb = (unsigned)cdiv(t->nleaves, 256);
  if (t->dim == 3)
    hipLaunchKernelGGL(k_src32<3>, dim3(nb), dim3(256), 0, c->stream, t->nleaves, t->leaf_order, t->start,
                       t->count, t->com, t->spart, t->spart32, t->pleaf);
  else
    hipLaunchKernelGGL(k_src32<2>, dim3(nb), dim3(256), 0, c->stream, t->nleaves, t->leaf_order, t->start,
                       t->count, t->com, t->spart, t->spart32, t->pleaf);
  BH_LAUNCH_CHECK();
  if (t->binary && t->slot_node) {
    for (int d = t->dmax; d >= 0; d--) {
      const int64_t nthr = t->lvl_nsys[d] << d;
      const int32_t* ls = t->d_lvl_sys ? t->d_lvl_sys + t->h_lvl_off[d] : nullptr;
      hipLaunchKernelGGL(k_quad_level, dim3((unsigned)cdiv(nthr, 256)), dim3(256), 0, c->stream, d, nthr, ls,
                         t->d_slot_off, t->slot_node, t->start, t->count, t->first_child, t->spart, t->com, t->dim,
                         t->mass, t->quad);
    }
    hipLaunchKernelGGL(k_quad32, dim3((unsigned)cdiv(t->M, 256)), dim3(256), 0, c->stream, t->M, t->quad, t->quad32);
    BH_LAUNCH_CHECK();
    t->quad_ready = true;
  } else if (t->binary) {
    Scratch s(c);
    uint32_t* flags = s.get<uint32_t>((size_t)t->M);
    if (!flags) return BH_ENOMEM;
    int rc = fill_u32(c, flags, 0u, t->M);
    if (rc) return rc;
    hipLaunchKernelGGL(k_quad_up, dim3(nb), dim3(256), 0, c->stream, t->nleaves, t->leaf_order, t->start,
                       t->count, t->first_child, t->parent, t->spart, t->com, t->dim, t->mass, t->quad, flags);
    hipLaunchKernelGGL(k_quad32, dim3((unsigned)cdiv(t->M, 256)), dim3(256), 0, c->stream, t->M, t->quad, t->quad32);
    BH_LAUNCH_CHECK();
    t->quad_ready = true;
  } else {
    t->quad_ready = false;  // imported trees with other fan-outs: monopoles only
  }
  t->f32_ready = true;
  return BH_OK;
}

// ---------------------------------------------------------------- groups
// Target range of group g: runs of gs targets, or (forest) the table gt0
// whose groups never straddle two systems.
__device__ __forceinline__ void group_range(int64_t g, int64_t m, int gs, const int32_t* __restrict__ gt0,
                                            int64_t& t0, int64_t& t1) {
  if (gt0) {
    t0 = gt0[g];
    t1 = gt0[g + 1];
  } else {
    t0 = g * gs;
    t1 = t0 + gs < m ? t0 + gs : m;
  }
}

__global__ __launch_bounds__(64) void k_group_box(const double4* __restrict__ tpos, int64_t m, int gs,
                                                  const int32_t* __restrict__ gt0, GroupBox* __restrict__ boxes) {
  const int64_t g = blockIdx.x;
  int64_t t0, t1;
  group_range(g, m, gs, gt0, t0, t1);
  double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
  double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
  for (int64_t i = t0 + threadIdx.x; i < t1; i += 64) {
    const double4 p = tpos[i];
    lo[0] = fmin(lo[0], p.x);
    lo[1] = fmin(lo[1], p.y);
    lo[2] = fmin(lo[2], p.z);
    hi[0] = fmax(hi[0], p.x);
    hi[1] = fmax(hi[1], p.y);
    hi[2] = fmax(hi[2], p.z);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off));
    }
  }
  if (threadIdx.x == 0) {
    GroupBox b;
    b.cx = 0.5 * (lo[0] + hi[0]);
    b.cy = 0.5 * (lo[1] + hi[1]);
    b.cz = 0.5 * (lo[2] + hi[2]);
    // half extents rounded up so the box still contains every target
    b.hx = 0.5 * (hi[0] - lo[0]) * (1 + 1e-12);
    b.hy = 0.5 * (hi[1] - lo[1]) * (1 + 1e-12);
    b.hz = 0.5 * (hi[2] - lo[2]) * (1 + 1e-12);
    boxes[g] = b;
  }
}

__global__ void k_gather_targets(const double* __restrict__ tgt, const uint32_t* __restrict__ order, int64_t m,
                                 int dim, double4* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t o = order[i];
  out[i] = make_double4(tgt[o * dim], tgt[o * dim + 1], dim == 3 ? tgt[o * dim + 2] : 0.0, 0.0);
}

// ---------------------------------------------------------------- walk

constexpr int MAX_STACK = 8192;  // LDS stack entries per wavefront, at most
constexpr int WALK_WAVES = 4;    // groups (wavefronts) per workgroup, each with its own LDS stack
constexpr int CBUF = 16;         // candidate nodes per round of the exact test (LDS, per wavefront)

// Each step pops up to 64 stack entries, one per lane, and tests those nodes
// at once; the output order (step, lane) is the same in the count and fill
// passes.  TPW = group size / 64 targets per lane: the exact test runs
// across the lanes, every lane measuring its own targets against one
// candidate node at a time, the candidates read from LDS as broadcasts.
// Opening radius: a node is accepted for a point at squared distance d2 from
// its centre of mass iff size^2 / theta^2 < d2 (itheta2 = 1 / theta^2; at
// theta = 0 it is +inf and every node opens).
template <int DIM, bool FILL, int TPW>
__global__ __launch_bounds__(64 * WALK_WAVES) void k_group_walk(const NodeD* __restrict__ nodes,
                                                   const GroupBox* __restrict__ boxes, double itheta2,
                                                   int stack_cap, int64_t* __restrict__ ccount,
                                                   int64_t* __restrict__ kcount, int64_t* __restrict__ pcount,
                                                   const int64_t* __restrict__ coff,
                                                   const int64_t* __restrict__ koff,
                                                   CellEnt* __restrict__ clist, ChunkEnt* __restrict__ klist,
                                                   uint32_t* __restrict__ err, const int32_t* __restrict__ groot,
                                                   int64_t g0, int64_t nlaunch, int wpb,
                                                   const double4* __restrict__ tpos, int64_t m,
                                                   const int32_t* __restrict__ gt0,
                                                   const int32_t* __restrict__ order, int64_t cap) {
  extern __shared__ int32_t lds_stacks[];
  // lists of fixed capacity that would overflow are not written at all
  if (FILL && (__builtin_amdgcn_readfirstlane(*err) & 2u)) return;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int64_t gl = (int64_t)blockIdx.x * wpb + wv;
  if (wv >= wpb || gl >= nlaunch) return;  // whole wavefront
  // order: launch slot -> group (relative to g0), longest lists first
  const int64_t g = g0 + (order ? (int64_t)order[gl] : gl);
  const int lane = threadIdx.x % 64;
  int32_t* stack = lds_stacks + wv * stack_cap;
  double4* cbuf = reinterpret_cast<double4*>(lds_stacks + wpb * stack_cap) + wv * CBUF;
  int64_t t0, t1;
  group_range(g, m, 64 * TPW, gt0, t0, t1);
  // this lane's targets; lanes past the group's end repeat its first target,
  // which is measured anyway
  double4 tq[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int64_t i = t0 + lane + 64 * q;
    tq[q] = tpos[i < t1 ? i : t0];
  }
  const double4 T0 = tpos[t0];
  const GroupBox B = boxes[g];
  if (lane == 0) stack[0] = groot ? groot[g] : 0;  // the group's own system
  wave_sync();
  // the targets have arrived before the loop, so no wait inside it (where
  // the list stores of earlier steps are outstanding) is for them
  __builtin_amdgcn_s_waitcnt(0);
  int sp = 1;
  int64_t nc = 0, nk = 0, np = 0;
  // list ranges: from the scanned offsets, or (cap > 0) fixed regions of
  // cap entries per group, with the counts written too
  const int64_t cb = !FILL ? 0 : cap ? (g - g0) * cap : coff[g];
  const int64_t ce = !FILL ? 0 : cap ? cb + cap : coff[g + 1];
  const int64_t kb = !FILL ? 0 : cap ? (g - g0) * cap : koff[g];
  const int64_t ke = !FILL ? 0 : cap ? kb + cap : koff[g + 1];
  const uint64_t lt = lanemask_lt();
  while (sp > 0) {
    const int k = sp < 64 ? sp : 64;
    const int v = lane < k ? stack[sp - k + lane] : -1;
    wave_sync();
    sp -= k;
    NodeD nd;
    if (v >= 0) nd = nodes[v];
    int kind = 0;  // 1 cell, 2 leaf, 3 open
    const double r2 = nd.size * nd.size * itheta2;  // opening radius^2
    if (v >= 0 && nd.m != 0.0) {
      if (nd.fc < 0) {
        kind = 2;
      } else {
        double ex = fabs(nd.cx - B.cx) - B.hx;
        double ey = fabs(nd.cy - B.cy) - B.hy;
        double ez = DIM == 3 ? fabs(nd.cz - B.cz) - B.hz : 0.0;
        ex = ex > 0.0 ? ex : 0.0;
        ey = ey > 0.0 ? ey : 0.0;
        ez = ez > 0.0 ? ez : 0.0;
        const double d2 = fma(ex, ex, fma(ey, ey, ez * ez));
        const bool open = !(d2 > 0.0) || !(r2 < d2);
        kind = open ? 3 : 1;
        if (open) {
          // the box test opens; the exact test -- every target of the
          // group accepts, i.e. size^2 < theta^2 min_t |t - com|^2 -- can
          // still accept when the box is loose around its targets.  Its
          // necessary condition uses the first target.
          const double ux = T0.x - nd.cx, uy = T0.y - nd.cy, uz = DIM == 3 ? T0.z - nd.cz : 0.0;
          if (r2 < fma(ux, ux, fma(uy, uy, uz * uz))) kind = 4;
        }
      }
    }
    // exact test, one candidate node at a time across the lanes: accept iff
    // no target lies within the opening radius of the centre of mass.  The
    // candidates go through LDS in rounds of CBUF, in lane order.
    uint64_t cand = __ballot(kind == 4);
    if (cand) {
      uint64_t nearm = 0;
      do {
        const int r = __popcll(cand & lt);
        if (((cand >> lane) & 1) && r < CBUF) cbuf[r] = make_double4(nd.cx, nd.cy, DIM == 3 ? nd.cz : 0.0, r2);
        wave_sync();
        const int nb = __popcll(cand) < CBUF ? __popcll(cand) : CBUF;
        for (int i = 0; i < nb; i++) {
          const int j = __builtin_ctzll(cand);
          cand &= cand - 1;
          const double4 cn = cbuf[i];
          bool near = false;
#pragma unroll
          for (int q = 0; q < TPW; q++) {
            const double ux = tq[q].x - cn.x, uy = tq[q].y - cn.y, uz = DIM == 3 ? tq[q].z - cn.z : 0.0;
            near = near || !(cn.w < fma(ux, ux, fma(uy, uy, uz * uz)));
          }
          if (__ballot(near)) nearm |= 1ull << j;
        }
        wave_sync();  // every read of this round precedes the next round's writes
      } while (cand);
      if (kind == 4) kind = (nearm >> lane) & 1 ? 3 : 1;
    }
    const uint64_t bc = __ballot(kind == 1), bo = __ballot(kind == 3);
    const uint64_t bk = __ballot(kind == 2);
    const int kx = __popcll(bk & lt), ktot = __popcll(bk);
    if (kind == 3) {
      const int o = sp + 2 * __popcll(bo & lt);
      if (o + 1 < stack_cap) {
        stack[o] = nd.fc;
        stack[o + 1] = nd.fc + 1;
      }
    }
    if (FILL) {
      if (kind == 1) {
        const int64_t o = cb + nc + __popcll(bc & lt);
        if (o < ce) clist[o] = v;
      }
      if (kind == 2) {
        const int64_t o = kb + nk + kx;
        if (o < ke) klist[o] = v;
      }
    }
    if (ccount && kind == 2) {
      np += nd.count;  // this lane's share; summed over the lanes at the end
    }
    nc += __popcll(bc);
    nk += ktot;
    const int push = 2 * __popcll(bo);
    if (sp + push > stack_cap) {
      if (lane == 0) atomicOr(err, 1u);
      return;
    }
    sp += push;
    wave_sync();
  }
  if (ccount) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) np += __shfl_xor(np, off);
    if (lane == 0) {
      ccount[g] = nc;
      kcount[g] = nk;
      pcount[g] = np;
    }
  }
}

// One launch of the walk for the tree's dimension and the group size (64 or
// 128 targets: one or two per lane).
template <bool FILL>
void launch_walk(bh_ctx* c, const bh_tree* t, dim3 grid, size_t lds, int gs, const GroupBox* boxes,
                 double itheta2, int stack_cap, int64_t* ccount, int64_t* kcount, int64_t* pcount,
                 const int64_t* coff, const int64_t* koff, CellEnt* clist, ChunkEnt* klist, uint32_t* err,
                 const int32_t* groot, int64_t g0, int64_t nlaunch, int wpb, const double4* tpos, int64_t m,
                 const int32_t* gt0, const int32_t* order, int64_t cap = 0) {
#define BH_WALK(D, TPW)                                                                                        \
  hipLaunchKernelGGL((k_group_walk<D, FILL, TPW>), grid, dim3(64 * WALK_WAVES), lds, c->stream, t->nodes,     \
                     boxes, itheta2, stack_cap, ccount, kcount, pcount, coff,                       \
                     koff, clist, klist, err, groot, g0, nlaunch, wpb, tpos, m, gt0, order, cap)
  if (t->dim == 3) {
    if (gs == 64) BH_WALK(3, 1);
    else BH_WALK(3, 2);
  } else {
    if (gs == 64) BH_WALK(2, 1);
    else BH_WALK(2, 2);
  }
#undef BH_WALK
}

// ---------------------------------------------------------------- eval
// UNI: every source has the same mass, applied to the tile's sum instead
// (one multiply less per pair).
template <bool ZEPS, bool UNI>
__device__ __forceinline__ void p2p(float4 s, float tx, float ty, float tz, float eps2, float& px, float& py,
                                    float& pz) {
  const float dx = s.x - tx, dy = s.y - ty, dz = s.z - tz;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
  float ri = __builtin_amdgcn_rsqf(r2);
  if (ZEPS) ri = r2 > 0.f ? ri : 0.f;
  const float mg = UNI ? ri * ri * ri : s.w * (ri * ri * ri);
  px = fmaf(mg, dx, px);
  py = fmaf(mg, dy, py);
  pz = fmaf(mg, dz, pz);
}

// a = u (M r^-3 - 1.5 tr r^-5 + 7.5 uQu r^-7) - 3 r^-5 Q u, u = com - target:
// the monopole plus quadrupole field of the cell's second moments about its
// centre of mass (softened radius throughout).
template <bool ZEPS, bool QUAD>
__device__ __forceinline__ void cell(const float4& a, const float4& b, const float4& c, float tx, float ty, float tz,
                                     float eps2, float& px, float& py, float& pz) {
  const float ux = a.x - tx, uy = a.y - ty, uz = a.z - tz;
  const float r2 = fmaf(ux, ux, fmaf(uy, uy, fmaf(uz, uz, eps2)));
  float ri = __builtin_amdgcn_rsqf(r2);
  if (ZEPS) ri = r2 > 0.f ? ri : 0.f;
  const float ri2 = ri * ri;
  const float ri3 = ri2 * ri;
  if (!QUAD) {
    const float mg = a.w * ri3;
    px = fmaf(mg, ux, px);
    py = fmaf(mg, uy, py);
    pz = fmaf(mg, uz, pz);
    return;
  }
  const float ri5 = ri3 * ri2;
  const float ri7 = ri5 * ri2;
  const float qx = fmaf(b.x, ux, fmaf(c.x, uy, c.y * uz));
  const float qy = fmaf(c.x, ux, fmaf(b.y, uy, c.z * uz));
  const float qz = fmaf(c.y, ux, fmaf(c.z, uy, b.z * uz));
  const float uqu = fmaf(ux, qx, fmaf(uy, qy, uz * qz));
  const float cu = fmaf(7.5f * uqu, ri7, fmaf(-1.5f * b.w, ri5, a.w * ri3));
  const float cq = -3.f * ri5;
  px = fmaf(cu, ux, fmaf(cq, qx, px));
  py = fmaf(cu, uy, fmaf(cq, qy, py));
  pz = fmaf(cu, uz, fmaf(cq, qz, pz));
}

__device__ __forceinline__ void kahan_add(float v, float& s, float& c) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// The tile's sums p (times the tile's common factor f) into the running
// sums s (compensated by c when KAHAN).
template <int TPL, bool KAHAN>
__device__ __forceinline__ void flush(float (&px)[TPL], float (&py)[TPL], float (&pz)[TPL], float (&sx)[TPL],
                                      float (&sy)[TPL], float (&sz)[TPL], float (&cx)[TPL], float (&cy)[TPL],
                                      float (&cz)[TPL], float f = 1.f) {
#pragma unroll
  for (int q = 0; q < TPL; q++) {
    if (KAHAN) {
      kahan_add(f * px[q], sx[q], cx[q]);
      kahan_add(f * py[q], sy[q], cy[q]);
      kahan_add(f * pz[q], sz[q], cz[q]);
    } else {
      sx[q] += f * px[q];
      sy[q] += f * py[q];
      sz[q] += f * pz[q];
    }
    px[q] = py[q] = pz[q] = 0.f;
  }
}


constexpr int EV_WAVES = 4;  // groups (wavefronts) per workgroup
constexpr int EV_XCDS = 8;   // item ranges of the evaluation: one per XCD of an MI355X

// Work items.  A group's list is a sequence of tiles: its cell tiles (64
// cells each) then its chunk tiles (64/CH chunks each).  Item j of a group
// covers tiles [j*ITEM_TILES, (j+1)*ITEM_TILES) of that sequence, so every
// item costs at most ITEM_TILES tiles and a wavefront's tail is at most one
// item long.  The split depends only on the group's own list lengths.
constexpr int ITEM_TILES = 64;

__host__ __device__ __forceinline__ int64_t group_items(int64_t ncell, int64_t nchunk, int cpt) {
  const int64_t tiles = (ncell + 63) / 64 + (nchunk + cpt - 1) / cpt;
  const int64_t k = (tiles + ITEM_TILES - 1) / ITEM_TILES;
  return k > 1 ? k : 1;
}

// CH: particles per chunk (64/CH chunks fill one 64-slot tile).
// One LDS tile per wavefront, refilled in place: the wavefront issued every
// read of the old tile before its writes of the new one, and its LDS
// operations execute in order.  The next tile's HBM reads are in flight in
// registers while the current tile is consumed.  The item's compensated
// sums go to part[(q*64 + lane)*3 + d] as FP64 (sum - compensation).
template <int TPL, int CH, bool ZEPS, bool SELF, bool KAHAN, bool QUAD>
__device__ __forceinline__ void eval_item(
    int64_t g, int64_t ta, int64_t tb, int lane, float4 (*tile)[64], int64_t m, int gs,
    const double4* __restrict__ tpos, const float4* __restrict__ src32, const int32_t* __restrict__ pleaf,
    const NodeD* __restrict__ nodes, const float4* __restrict__ quad32,
    const double* __restrict__ com, int dim, const GroupBox* __restrict__ boxes,
    const int64_t* __restrict__ coff, const CellEnt* __restrict__ clist, const int64_t* __restrict__ koff,
    const int64_t* __restrict__ ccnt, const int64_t* __restrict__ kcnt,
    const ChunkEnt* __restrict__ klist, float eps2, const int32_t* __restrict__ gt0, double* __restrict__ part) {
  constexpr int CPT = 64 / CH;
  int64_t t0, t1;
  group_range(g, m, gs, gt0, t0, t1);
  const GroupBox B = boxes[g];
  float tx[TPL], ty[TPL], tz[TPL];
  float sx[TPL], sy[TPL], sz[TPL], cx[TPL], cy[TPL], cz[TPL], px[TPL], py[TPL], pz[TPL];
#pragma unroll
  for (int q = 0; q < TPL; q++) {
    const int64_t ti = t0 + lane + 64 * q;
    const int64_t i = ti < t1 ? ti : t0;
    if (SELF) {
      // built exactly like the source copy of the same particle, so its own
      // pair has dx == 0 and contributes nothing
      const float4 s = src32[i];
      const int v = pleaf[i];
      tx[q] = s.x + (float)(com[(int64_t)v * dim] - B.cx);
      ty[q] = s.y + (float)(com[(int64_t)v * dim + 1] - B.cy);
      tz[q] = s.z + (dim == 3 ? (float)(com[(int64_t)v * dim + 2] - B.cz) : 0.f);
    } else {
      const double4 p = tpos[i];
      tx[q] = (float)(p.x - B.cx);
      ty[q] = (float)(p.y - B.cy);
      tz[q] = (float)(p.z - B.cz);
    }
    sx[q] = sy[q] = sz[q] = cx[q] = cy[q] = cz[q] = px[q] = py[q] = pz[q] = 0.f;
  }
  const int64_t c0 = coff[g], c1 = c0 + ccnt[g];
  const int64_t ctiles = (c1 - c0 + 63) / 64;

  // ---- accepted cells of tiles [ta, min(tb, ctiles)): each lane gathers
  // one cell's node record and moment payload; the indices are loaded two
  // tiles ahead and the payloads one tile ahead of the tile being consumed
  if (ta < ctiles) {
    const int64_t cs = c0 + 64 * ta;
    const int64_t ce = tb < ctiles ? c0 + 64 * tb : c1;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    double4 na = make_double4(0.0, 0.0, 0.0, 0.0);  // centre of mass, mass
    float4 nb = z4, nc = z4;
    auto load_cell = [&](int v) {
      if (v >= 0) {
        const double4* r = reinterpret_cast<const double4*>(nodes + v);
        na = r[0];
        if (QUAD) {
          nb = quad32[2 * (int64_t)v];
          nc = quad32[2 * (int64_t)v + 1];
        }
      }
    };
    int v_nxt = cs + 64 + lane < ce ? clist[cs + 64 + lane] : -1;
    load_cell(cs + lane < ce ? clist[cs + lane] : -1);
    for (int64_t e = cs; e < ce; e += 64) {
      wave_sync();
      tile[0][lane] = make_float4((float)(na.x - B.cx), (float)(na.y - B.cy), (float)(na.z - B.cz), (float)na.w);
      tile[1][lane] = nb;
      tile[2][lane] = nc;
      wave_sync();
      load_cell(v_nxt);
      v_nxt = e + 128 + lane < ce ? clist[e + 128 + lane] : -1;
      const int cnt = (int)(ce - e < 64 ? ce - e : 64);
      for (int k = 0; k < cnt; k++) {
        const float4 a = tile[0][k], b = tile[1][k], c = tile[2][k];
#pragma unroll
        for (int q = 0; q < TPL; q++) cell<ZEPS, QUAD>(a, b, c, tx[q], ty[q], tz[q], eps2, px[q], py[q], pz[q]);
      }
      flush<TPL, KAHAN>(px, py, pz, sx, sy, sz, cx, cy, cz);
    }
  }

  // ---- leaf chunks of tiles [max(ta, ctiles), tb) - ctiles, compacted: the
  // item's chunks form one stream of sources cut into tiles of 64, so a
  // chunk of fewer than CH particles leaves no idle slots (only the item's
  // last tile is partly filled).  Chunks are staged in windows of 64 in the
  // LDS tiles the cell part used: entries (shift xyz, first source) in
  // tile[1], the exclusive prefix of their counts in tile[2]; each lane finds
  // the chunk of its stream position by a binary search of that prefix.
  const int64_t k0 = koff[g], k1 = k0 + kcnt[g];
  const int64_t ka0 = ta > ctiles ? ta - ctiles : 0;
  const int64_t kb0 = tb - ctiles;
  const int64_t ks = k0 + CPT * ka0;
  const int64_t ke = k0 + CPT * kb0 < k1 ? k0 + CPT * kb0 : k1;
  if (kb0 > 0 && ks < ke) {
    float4* went = tile[1];                                    // (shift x, y, z, first source bits)
    int* wpre = reinterpret_cast<int*>(tile[2]);                // [0..64]
    int64_t w = ks;  // first chunk of the window
    int nw = 0;      // chunks in the window
    int wtot = 0;    // sources in the window
    int tp = 0;      // stream position of the next tile, relative to the window
    auto load_window = [&]() {
      const int64_t e = w + lane;
      float4 ex = make_float4(0.f, 0.f, 0.f, 0.f);
      int cnt = 0;
      if (e < ke) {
        const NodeD& r = nodes[klist[e]];
        ex = make_float4((float)(r.cx - B.cx), (float)(r.cy - B.cy), (float)(r.cz - B.cz), __int_as_float(r.start));
        cnt = r.count;
      }
      int inc = cnt;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
      }
      wave_sync();  // every read of the previous window precedes these writes
      went[lane] = ex;
      wpre[lane] = inc - cnt;
      wave_sync();
      nw = ke - w < 64 ? (int)(ke - w) : 64;
      wtot = __builtin_amdgcn_readfirstlane(__shfl(inc, 63));
    };
    // the source at stream position P of the window (0 <= P < wtot): the
    // last chunk i < nw whose prefix is <= P
    auto locate = [&](int P, float4& sh, int& idx) {
      int lo = 0, plo = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        if (lo + step < nw) {
          const int q = wpre[lo + step];
          if (q <= P) {
            lo += step;
            plo = q;
          }
        }
      }
      sh = went[lo];
      idx = __float_as_int(sh.w) + (P - plo);
    };
    // the next tile: each lane's source (load in flight) and shift; returns
    // the number of filled slots (lanes 0 .. n-1)
    auto map_tile = [&](float4& src, float4& sh, bool& valid) -> int {
      if (tp >= wtot && w + nw < ke) {
        tp -= wtot;
        w += nw;
        load_window();
      }
      int P = tp + lane;
      int idx = 0;
      valid = false;
      if (P < wtot) {
        locate(P, sh, idx);
        valid = true;
      }
      if (tp + 64 > wtot && w + nw < ke) {
        // the tile continues in the next window (a window of 64 chunks holds
        // at least 64 sources, so no tile spans three)
        const int old = wtot;
        w += nw;
        load_window();
        tp -= old;
        P -= old;
        if (!valid && P >= 0 && P < wtot) {
          locate(P, sh, idx);
          valid = true;
        }
      }
      tp += 64;
      src = make_float4(0.f, 0.f, 0.f, 0.f);
      if (valid) src = src32[idx];
      return __popcll(__ballot(valid));
    };
    load_window();
    float4 src_cur, sh_cur;
    bool v_cur;
    int n_cur = map_tile(src_cur, sh_cur, v_cur);
    for (;;) {
      wave_sync();
      tile[0][lane] = v_cur ? make_float4(src_cur.x + sh_cur.x, src_cur.y + sh_cur.y, src_cur.z + sh_cur.z,
                                          src_cur.w)
                            : make_float4(1e15f, 1e15f, 1e15f, 0.f);
      wave_sync();
      // a tile whose sources share one finite mass takes the loop without the
      // per-pair mass product; its sum is scaled by that mass once
      const float um = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(src_cur.w)));
      const bool uni = __ballot(v_cur && src_cur.w != um) == 0 && __builtin_isfinite(um);
      const int nslots = n_cur;
      const bool more = !(tp >= wtot && w + nw >= ke);
      if (more) n_cur = map_tile(src_cur, sh_cur, v_cur);  // the next tile's loads are in flight
      if (uni) {
#pragma unroll 8
        for (int k = 0; k < nslots; k++) {
          const float3 v3 = *reinterpret_cast<const float3*>(&tile[0][k]);  // the mass is not read
          const float4 v = make_float4(v3.x, v3.y, v3.z, 0.f);
#pragma unroll
          for (int q = 0; q < TPL; q++) p2p<ZEPS, true>(v, tx[q], ty[q], tz[q], eps2, px[q], py[q], pz[q]);
        }
      } else {
#pragma unroll 8
        for (int k = 0; k < nslots; k++) {
          const float4 v = tile[0][k];
#pragma unroll
          for (int q = 0; q < TPL; q++) p2p<ZEPS, false>(v, tx[q], ty[q], tz[q], eps2, px[q], py[q], pz[q]);
        }
      }
      flush<TPL, KAHAN>(px, py, pz, sx, sy, sz, cx, cy, cz, uni ? um : 1.f);
      if (!more) break;
    }
  }
#pragma unroll
  for (int q = 0; q < TPL; q++) {
    double* o = part + (int64_t)(q * 64 + lane) * 3;
    o[0] = (double)sx[q] - (double)cx[q];
    o[1] = (double)sy[q] - (double)cy[q];
    o[2] = (double)sz[q] - (double)cz[q];
  }
}

// Persistent evaluation: as many workgroups as fit on the device at once;
// each wavefront takes the next work item from a device counter until none
// is left (one counter per item range, below).  Items are numbered group by
// group (ioff[g] = first item of group g, ioff[ng] = number of items).
// next[0 .. EV_XCDS) must be zero at launch.
template <int TPL, int CH, bool ZEPS, bool SELF, bool KAHAN, bool QUAD>
__global__ __launch_bounds__(64 * EV_WAVES, 8) void k_group_eval(
    int64_t g0, int64_t ngroups, const int32_t* __restrict__ ioff, int64_t m, int gs,
    const double4* __restrict__ tpos, const float4* __restrict__ src32, const int32_t* __restrict__ pleaf,
    const NodeD* __restrict__ nodes, const float4* __restrict__ quad32,
    const double* __restrict__ com, int dim, const GroupBox* __restrict__ boxes,
    const int64_t* __restrict__ coff, const CellEnt* __restrict__ clist, const int64_t* __restrict__ koff,
    const int64_t* __restrict__ ccnt, const int64_t* __restrict__ kcnt,
    const ChunkEnt* __restrict__ klist, float eps2, const int32_t* __restrict__ gt0,
    const uint32_t* __restrict__ err, uint32_t* __restrict__ next, double* __restrict__ part) {
  constexpr int CPT = 64 / CH;
  __shared__ float4 tiles[EV_WAVES][3][64];
  if (__builtin_amdgcn_readfirstlane(*err) != 0u) return;  // lists incomplete
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int lane = threadIdx.x % 64;
  const int64_t nitems = ioff[ngroups];
  // the items form EV_XCDS contiguous ranges of neighbouring groups, one per
  // XCD (workgroups are dispatched to the XCDs in turn), each with its own
  // counter; a wavefront that finds its own range done takes items from the
  // others'.  A small problem keeps one range (one counter attempt per
  // wavefront at the end instead of EV_XCDS).
  const int nx = nitems >= 2048 ? EV_XCDS : 1;
  const int x0 = (int)(blockIdx.x % nx);
  for (int r = 0; r < nx; r++) {
    const int x = (x0 + r) % nx;
    const int64_t lo_x = nitems * x / nx, hi_x = nitems * (x + 1) / nx;
    for (;;) {
      uint32_t sl = 0;
      if (lane == 0) sl = atomicAdd(next + x, 1u);
      const int64_t item = lo_x + (int64_t)__builtin_amdgcn_readfirstlane(sl);
      if (item >= hi_x) break;  // whole wavefront
      // the item's group: last gi with ioff[gi] <= item
      int64_t lo = 0, hi = ngroups - 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (ioff[mid] <= item) lo = mid;
        else hi = mid - 1;
      }
      const int64_t g = g0 + lo;
      const int64_t j = item - ioff[lo];
      const int64_t tiles_g = (ccnt[g] + 63) / 64 + (kcnt[g] + CPT - 1) / CPT;
      const int64_t ta = j * ITEM_TILES;
      const int64_t tb = ta + ITEM_TILES < tiles_g ? ta + ITEM_TILES : tiles_g;
      eval_item<TPL, CH, ZEPS, SELF, KAHAN, QUAD>(g, ta, tb, lane, tiles[w], m, gs, tpos, src32, pleaf, nodes,
                                                  quad32, com, dim,
                                                  boxes, coff, clist, koff, ccnt, kcnt, klist, eps2, gt0,
                                                  part + item * (int64_t)(64 * TPL * 3));
    }
  }
}

// Accelerations from the items' partial sums, added in item order (so the
// result does not depend on which wavefront ran which item).
__global__ __launch_bounds__(256) void k_item_reduce(int64_t g0, int64_t ngroups, const int32_t* __restrict__ ioff,
                                                     int64_t m, int gs, const int32_t* __restrict__ gt0,
                                                     const double* __restrict__ part, int tpg,
                                                     const uint32_t* __restrict__ torig, int dim, double G,
                                                     const uint32_t* __restrict__ err, double* __restrict__ out) {
  if (*err != 0u) return;
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (x >= ngroups * tpg) return;
  const int64_t lg = x / tpg, l = x % tpg;
  int64_t t0, t1;
  group_range(g0 + lg, m, gs, gt0, t0, t1);
  if (t0 + l >= t1) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int64_t it = ioff[lg]; it < ioff[lg + 1]; it++) {
    const double* p = part + (it * tpg + l) * 3;
    a0 = a0 + p[0];
    a1 = a1 + p[1];
    a2 = a2 + p[2];
  }
  const int64_t o = torig[t0 + l];
  out[o * dim] = G * a0;
  out[o * dim + 1] = G * a1;
  if (dim == 3) out[o * dim + 2] = G * a2;
}

// The one-pass lists' layout and work items, on the device: each group's
// lists sit in regions of cap entries (coff[g] = koff[g] = g * cap), or
// (preset) in the regions offs already holds; ioff[] is the prefix of the
// groups' item counts.  A list longer than its region, or more items than
// cap_items, sets bit 2 of *err (the evaluation then does nothing);
// need[0] = the longest list, need[1] = the items.
// sums (optional): cell and pair interactions (entries x the group's
// targets), cell and leaf entries.
__global__ __launch_bounds__(1024) void k_scan_lists(int64_t ng, const int64_t* __restrict__ cnt, int cpt,
                                                     int64_t* __restrict__ offs, int32_t* __restrict__ ioff,
                                                     int64_t cap, int preset, int64_t cap_items,
                                                     uint32_t* __restrict__ err,
                                                     int64_t* __restrict__ need, int64_t* __restrict__ sums,
                                                     int64_t m, int gs, const int32_t* __restrict__ gt0) {
  constexpr int IT = 8;
  __shared__ int64_t wsum[16], wmax[16], wsums[4][16];
  const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  int64_t* coff = offs;
  int64_t* koff = offs + ng + 1;
  int64_t carry_i = 0, mx = 0;
  int64_t acc[4] = {0, 0, 0, 0};
  bool ovf = false;
  for (int64_t base = 0; base < ng; base += 1024 * IT) {
    int64_t vi[IT];
    int64_t si = 0;
#pragma unroll
    for (int i = 0; i < IT; i++) {
      const int64_t g = base + (int64_t)tid * IT + i;
      const int64_t vc = g < ng ? cnt[g] : 0, vk = g < ng ? cnt[ng + g] : 0;
      vi[i] = g < ng ? group_items(vc, vk, cpt) : 0;
      si += vi[i];
      if (sums && g < ng) {
        int64_t a, b;
        group_range(g, m, gs, gt0, a, b);
        acc[0] += vc * (b - a);
        acc[1] += cnt[2 * ng + g] * (b - a);
        acc[2] += vc;
        acc[3] += vk;
      }
      mx = vc > mx ? vc : mx;
      mx = vk > mx ? vk : mx;
      // preset regions: a list longer than its own region overflows
      if (preset && g < ng && (vc > coff[g + 1] - coff[g] || vk > koff[g + 1] - koff[g])) ovf = true;
    }
    int64_t ii = si;  // inclusive wave scan of the thread sums
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t yi = __shfl_up(ii, off);
      if (lane >= off) ii += yi;
    }
    if (lane == 63) wsum[w] = ii;
    __syncthreads();
    int64_t pi = carry_i + ii - si, tn = 0;
    for (int j = 0; j < 16; j++) {
      if (j < w) pi += wsum[j];
      tn += wsum[j];
    }
    __syncthreads();  // wsum is rewritten by the next chunk
#pragma unroll
    for (int i = 0; i < IT; i++) {
      const int64_t g = base + (int64_t)tid * IT + i;
      if (g < ng) {
        if (!preset) {
          coff[g] = g * cap;
          koff[g] = g * cap;
        }
        ioff[g] = (int32_t)pi;
      }
      pi += vi[i];
    }
    carry_i += tn;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int64_t y = __shfl_xor(mx, off);
    mx = y > mx ? y : mx;
  }
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
  if (lane == 0) {
    wmax[w] = mx;
    for (int q = 0; q < 4; q++) wsums[q][w] = acc[q];
  }
  __syncthreads();
  if (tid == 0) {
    for (int j = 0; j < 16; j++) mx = wmax[j] > mx ? wmax[j] : mx;
    if (sums)
      for (int q = 0; q < 4; q++) {
        int64_t t = 0;
        for (int j = 0; j < 16; j++) t += wsums[q][j];
        sums[q] = t;
      }
    if (!preset) {
      coff[ng] = ng * cap;
      koff[ng] = ng * cap;
    }
    ioff[ng] = (int32_t)carry_i;
    need[0] = mx;
    need[1] = carry_i;
    if ((!preset && mx > cap) || carry_i > cap_items) atomicOr(err, 2u);
  }
  if (ovf) atomicOr(err, 2u);
}

// Regions for a one-pass walk from the previous call's list lengths (same
// groups): prev + prev/4 + 64 entries per group and list, as offsets in offs
// (cells, then leaves; ng + 1 each).  One workgroup.
__global__ __launch_bounds__(1024) void k_prev_regions(int64_t ng, const int64_t* __restrict__ prev,
                                                       int64_t* __restrict__ offs) {
  constexpr int IT = 8;
  __shared__ int64_t wsum[2][16];
  const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  int64_t carry[2] = {0, 0};
  for (int64_t base = 0; base < ng; base += 1024 * IT) {
    int64_t v[2][IT], sum[2] = {0, 0};
#pragma unroll
    for (int i = 0; i < IT; i++) {
      const int64_t g = base + (int64_t)tid * IT + i;
#pragma unroll
      for (int l = 0; l < 2; l++) {
        const int64_t p = g < ng ? prev[l * ng + g] : -1;
        v[l][i] = p >= 0 ? p + p / 4 + 64 : 0;
        sum[l] += v[l][i];
      }
    }
    int64_t inc[2] = {sum[0], sum[1]};
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
#pragma unroll
      for (int l = 0; l < 2; l++) {
        const int64_t y = __shfl_up(inc[l], off);
        if (lane >= off) inc[l] += y;
      }
    if (lane == 63)
      for (int l = 0; l < 2; l++) wsum[l][w] = inc[l];
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 2; l++) {
      int64_t pre = carry[l] + inc[l] - sum[l], tot = 0;
      for (int j = 0; j < 16; j++) {
        if (j < w) pre += wsum[l][j];
        tot += wsum[l][j];
      }
      for (int i = 0; i < IT; i++) {
        const int64_t g = base + (int64_t)tid * IT + i;
        if (g < ng) offs[l * (ng + 1) + g] = pre;
        pre += v[l][i];
      }
      carry[l] += tot;
    }
    __syncthreads();  // wsum is rewritten by the next chunk
  }
  if (tid == 0) {
    offs[ng] = carry[0];
    offs[2 * ng + 1] = carry[1];
  }
}

// Walk order for the next call on the same groups: a counting sort of the
// groups by list length (a proxy of their walk time), longest first, in
// ORD_BUCKETS buckets of 4 entries (longer lists share the first bucket), so
// the one-pass walk does not end on long walks started late.  The order
// within a bucket depends on atomics; the lists do not depend on the order.
constexpr int ORD_BUCKETS = 4096;

__device__ __forceinline__ int order_bucket(int64_t len) {
  // (counts left unwritten by a walk that overflowed its stack fall in a
  // bucket too: the order stays a permutation)
  const int64_t q = len > 0 ? len >> 2 : 0;
  return ORD_BUCKETS - 1 - (int)(q < ORD_BUCKETS - 1 ? q : ORD_BUCKETS - 1);
}

__global__ __launch_bounds__(256) void k_order_hist(int64_t ng, const int64_t* __restrict__ cnt,
                                                    uint32_t* __restrict__ hist) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  atomicAdd(&hist[order_bucket(cnt[g] + cnt[ng + g])], 1u);
}

// exclusive scan of the ORD_BUCKETS counts in place: one wavefront, 64
// consecutive buckets per lane
__global__ __launch_bounds__(64) void k_order_scan(uint32_t* __restrict__ hist) {
  constexpr int PER = ORD_BUCKETS / 64;
  const int lane = threadIdx.x;
  uint32_t s = 0;
  for (int i = 0; i < PER; i++) s += hist[lane * PER + i];
  uint32_t inc = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  uint32_t run = inc - s;
  for (int i = 0; i < PER; i++) {
    const uint32_t v = hist[lane * PER + i];
    hist[lane * PER + i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(256) void k_order_scatter(int64_t ng, const int64_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ cursor, int32_t* __restrict__ ord) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const uint32_t p = atomicAdd(&cursor[order_bucket(cnt[g] + cnt[ng + g])], 1u);
  ord[p] = (int32_t)g;
}

// ---------------------------------------------------------------- host
// Bytes of interaction lists materialised at once by the host-scanned path
// unless the context sets another budget: a quarter of the device memory, at
// most 64 GiB (12 GiB when the device memory is unknown).

static size_t default_list_budget(const bh_ctx* c) {
  if (c->mem_total == 0) return size_t(12) << 30;
  return std::min<size_t>(size_t(64) << 30, c->mem_total / 4);
}

// Workgroups of the persistent evaluation: all that fit at once, but no
// more than the groups need.
static unsigned eval_grid(bh_ctx* c, const void* kern, int64_t ng) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> per_cu;
  int n = 0;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = per_cu.find({kern, c->device});
    if (it != per_cu.end()) {
      n = it->second;
    } else {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, 64 * EV_WAVES, 0) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        n = 1;
      }
      per_cu[{kern, c->device}] = n;
    }
  }
  const int64_t resident = (int64_t)n * std::max(c->num_cus, 1);
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(ng, EV_WAVES), resident));
}

template <int TPL, int CH>
static void launch_eval(bh_ctx* c, int64_t g0, int64_t ng, const int32_t* ioff, bool zeps, bool self, bool kahan,
                        bool quad, int64_t m, int gs, const double4* tpos, const bh_tree* t, const GroupBox* boxes,
                        const int64_t* coff, const CellEnt* clist, const int64_t* koff, const ChunkEnt* klist,
                        const int64_t* ccnt, const int64_t* kcnt, float eps2, const int32_t* gt0,
                        const uint32_t* err, uint32_t* next, double* part, int64_t items_bound) {
#define BH_EV(Z, S, K, Q)                                                                                          \
  {                                                                                                                \
    const void* kern = reinterpret_cast<const void*>(&k_group_eval<TPL, CH, Z, S, K, Q>);                         \
    hipLaunchKernelGGL((k_group_eval<TPL, CH, Z, S, K, Q>), dim3(eval_grid(c, kern, items_bound)),                 \
                       dim3(64 * EV_WAVES), 0, c->stream, g0, ng, ioff, m, gs, tpos, t->spart32, t->pleaf,        \
                       t->nodes, t->quad32, t->com, t->dim, boxes, coff, clist, koff, ccnt, kcnt, klist, eps2, \
                       gt0, err, next, part);                                                                 \
  }
#define BH_EV_Q(Z, S, K) \
  if (quad) BH_EV(Z, S, K, true) else BH_EV(Z, S, K, false)
#define BH_EV_K(Z, S) \
  if (kahan) { BH_EV_Q(Z, S, true); } else { BH_EV_Q(Z, S, false); }
#define BH_EV_S(Z) \
  if (self) { BH_EV_K(Z, true); } else { BH_EV_K(Z, false); }
  if (zeps) {
    BH_EV_S(true);
  } else {
    BH_EV_S(false);
  }
#undef BH_EV_S
#undef BH_EV_K
#undef BH_EV_Q
#undef BH_EV
}

int walk_f32(bh_ctx* c, bh_tree* t, const double* tgt, int64_t m, double theta, double eps, double G,
             const bh_f32_opts* o, double* out, int64_t* stats, const F32Plan* plan) {
  if (m == 0) return BH_OK;
  // the group walk pushes both children of an opened node (first_child and
  // first_child + 1); an imported tree with other child counts is refused
  BH_REQ(t->binary, "the FP32 form needs a binary tree (every internal node with child_count 2)");
  const int gs = o && o->group_size ? o->group_size : 64;
  const bool kahan = o ? o->kahan != 0 : true;
  // LDS stack entries per wavefront, a multiple of 16 so what follows the
  // stacks in LDS stays 64-byte aligned
  int stack_cap = ((o && o->stack_cap ? o->stack_cap : 1024) + 15) & ~15;
  const int multipole = o && o->multipole ? o->multipole : 2;
  if (gs != 64 && gs != 128) {
    set_error("group_size must be 64 or 128, got %d", gs);
    return BH_EINVAL;
  }
  if (stack_cap < 64 || stack_cap > MAX_STACK) {
    set_error("stack_cap must be in [64, %d], got %d", MAX_STACK, stack_cap);
    return BH_EINVAL;
  }
  if (multipole != 1 && multipole != 2) {
    set_error("multipole must be 1 (monopole) or 2 (quadrupole), got %d", multipole);
    return BH_EINVAL;
  }
  std::atomic<int>& hint = t->stack_hint[tgt ? 1 : 0];
  if (!plan && hint.load() > stack_cap) stack_cap = hint.load();
  // a grown stack is kept for later walks on the tree (never shrinks)
  auto keep_hint = [&hint](int cap) {
    int cur = hint.load();
    while (cap > cur && !hint.compare_exchange_weak(cur, cap)) {
    }
  };
  if (!t->f32_ready) {
    ProfScope p(c, "f32_prepare");
    int rc = f32_prepare(c, t);
    if (rc) return rc;
  }
  const bool quad = multipole == 2 && t->quad_ready;
  const int ch = t->leaf <= 16 ? 16 : (t->leaf <= 32 ? 32 : 64);
  Scratch s(c);
  const double4* tpos;
  const uint32_t* torig;
  if (tgt) {
    uint32_t* order = s.get<uint32_t>(m);
    double4* tp = s.get<double4>(m);
    if (!order || !tp) return BH_ENOMEM;
    int rc = order_targets(c, t, tgt, m, order);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather_targets, dim3((unsigned)cdiv(m, 256)), dim3(256), 0, c->stream, tgt, order, m,
                       t->dim, tp);
    tpos = tp;
    torig = order;
  } else {
    tpos = t->spart;
    torig = reinterpret_cast<const uint32_t*>(t->perm);
  }
  // groups: runs of gs targets; a forest restarts the runs at every system
  // and walks each group from its own system's root (table kept in the tree)
  int64_t ng = cdiv(m, gs);
  const int32_t* gt0 = nullptr;
  const int32_t* groot = nullptr;
  if (!tgt && t->nsys > 1) {
    if (t->gtab_gs != gs) {
      if (c->capturing) {
        set_error("the forest group table must exist before graph capture");
        return BH_EINVAL;
      }
      std::vector<int32_t> hg;
      for (int64_t sy = 0; sy < t->nsys; sy++)
        for (int64_t a = t->sys_poff[sy]; a < t->sys_poff[sy + 1]; a += gs) hg.push_back((int32_t)a);
      const int64_t n2 = (int64_t)hg.size();
      hg.push_back((int32_t)m);
      for (int64_t sy = 0; sy < t->nsys; sy++)
        for (int64_t a = t->sys_poff[sy]; a < t->sys_poff[sy + 1]; a += gs) hg.push_back((int32_t)t->sys_noff[sy]);
      dfree(c, t->d_gtab);
      t->d_gtab = dalloc_t<int32_t>(c, hg.size());
      t->gtab_gs = 0;
      if (!t->d_gtab) return BH_ENOMEM;
      BH_HIP(hipMemcpyAsync(t->d_gtab, hg.data(), hg.size() * 4, hipMemcpyHostToDevice, c->stream));
      BH_HIP(hipStreamSynchronize(c->stream));  // hg is a temporary
      t->h_gtab.swap(hg);
      t->gtab_gs = gs;
      t->gtab_ng = n2;
    }
    ng = t->gtab_ng;
    gt0 = t->d_gtab;
    groot = t->d_gtab + ng + 1;
  }
  GroupBox* boxes = s.get<GroupBox>(ng);
  int64_t* cnt = s.get<int64_t>(3 * ng);
  int64_t* offs = s.get<int64_t>(2 * (ng + 1));
  uint32_t* err = plan ? plan->err : s.get<uint32_t>(1);
  if (!boxes || !cnt || !offs || !err) return BH_ENOMEM;
  if (!plan) {
    int rc = fill_u32(c, err, 0u, 1);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_group_box, dim3((unsigned)ng), dim3(64), 0, c->stream, tpos, m, gs, gt0, boxes);
  const double itheta2 = theta > 0.0 ? 1.0 / (theta * theta) : __builtin_inf();
  // wavefronts per workgroup: WALK_WAVES unless their stacks exceed the
  // workgroup's LDS limit
  int max_lds = 65536;
  if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess) {
    (void)hipGetLastError();
    max_lds = 65536;
  }
  int wpb = 1;
  size_t lds = 0;
  // One pass when it can: the walk writes the lists straight into regions of
  // their own and counts them at the same time -- a planned walk into
  // uniform regions of dcap entries per group, a host-scanned one into
  // regions sized per group from the context's previous call on the same
  // groups, if they fit the list budget.  Otherwise (no such call, or a list
  // outgrew its region) the groups are counted first and filled at scanned
  // offsets.
  int64_t dcap = 0;     // uniform one-pass regions of dcap entries per group
  bool sized = false;   // one-pass regions sized per group (offsets in offs)
  CellEnt* dclist = nullptr;
  ChunkEnt* dklist = nullptr;
  // the host-scanned one-pass walk visits the groups in the order the
  // previous walk over the same number of groups left (longest lists first),
  // and leaves the order for the next one (1M: walk 0.96 -> 0.86 ms; the
  // planned walk of the 4M leapfrog gained nothing from it and keeps index
  // order)
  int32_t* word = nullptr;
  int32_t* wnext = nullptr;
  if (!plan) {
    if (c->walk_order_cap < ng) {
      dfree(c, c->walk_order);
      c->walk_order = dalloc_t<int32_t>(c, (size_t)ng);
      c->walk_order_cap = c->walk_order ? ng : 0;
      c->walk_order_n = 0;
      (void)hipGetLastError();
    }
    if (c->walk_order && c->walk_order_n == ng) word = c->walk_order;
    wnext = c->walk_order;
  }
  if (plan) {
    // the planned path (no host synchronisation) always takes one pass, into
    // regions of the planned size
    dcap = std::max<int64_t>(plan->cap_region, 1);
    dclist = s.get<CellEnt>((size_t)(ng * dcap));
    dklist = s.get<ChunkEnt>((size_t)(ng * dcap));
    if (!dclist || !dklist) return BH_ENOMEM;
  } else {
    // regions from the previous call's list lengths when it had the same
    // groups (a quarter more each: memory ~1.25x the lists)
    const size_t budget = c->list_budget ? c->list_budget : default_list_budget(c);
    if (c->prev_cnt && c->prev_ng == ng) {
      const int64_t tc = c->prev_ntc + c->prev_ntc / 4 + 64 * ng, tk = c->prev_ntk + c->prev_ntk / 4 + 64 * ng;
      if ((size_t)(tc + tk) * sizeof(int32_t) <= budget) {
        dclist = s.get<CellEnt>((size_t)tc);
        dklist = s.get<ChunkEnt>((size_t)tk);
        if (dclist && dklist) {
          hipLaunchKernelGGL(k_prev_regions, dim3(1), dim3(1024), 0, c->stream, ng, c->prev_cnt, offs);
          BH_LAUNCH_CHECK();
          sized = true;
        }
        (void)hipGetLastError();
      }
    }
  }
  // the count (or one-pass) walk; an LDS stack overflow on the host-scanned
  // path reruns it with twice the stack (the lists do not depend on it)
  auto count_pass = [&]() -> int {
    const int64_t per_wave = (int64_t)stack_cap * 4 + CBUF * (int64_t)sizeof(double4);
    wpb = (int)std::max<int64_t>(1, std::min<int64_t>(WALK_WAVES, max_lds / per_wave));
    lds = (size_t)wpb * per_wave;
    if (dcap || sized) {
      ProfScope p(c, "k_group_walk_fill");
      launch_walk<true>(c, t, dim3((unsigned)cdiv(ng, wpb)), lds, gs, boxes, itheta2, stack_cap, cnt, cnt + ng,
                        cnt + 2 * ng, offs, offs + ng + 1, dclist, dklist, err, groot, 0, ng, wpb, tpos, m, gt0,
                        word, dcap);
    } else {
      ProfScope p(c, "k_group_walk_count");
      launch_walk<false>(c, t, dim3((unsigned)cdiv(ng, wpb)), lds, gs, boxes, itheta2, stack_cap, cnt,
                         cnt + ng, cnt + 2 * ng, nullptr, nullptr, nullptr, nullptr, err, groot, 0, ng, wpb, tpos,
                         m, gt0, nullptr);
    }
    BH_LAUNCH_CHECK();
    return BH_OK;
  };
  {
    int rc = count_pass();
    if (rc) return rc;
  }
  const int cpt = 64 / ch;
  const int tpg = gs;  // targets per group = 64 * TPL
  const float eps2 = (float)(eps * eps);
  const bool zeps = !(eps2 > 0.f);
  const bool self = tgt == nullptr;
  int32_t* ioff = s.get<int32_t>(ng + 1);
  uint32_t* next = s.get<uint32_t>(EV_XCDS);
  if (!ioff || !next) return BH_ENOMEM;
  // the evaluation and the reduction of groups [g0, g0 + nb): lists at the
  // offsets in offs, items[ioff], partial sums in part
  auto evaluate = [&](int64_t g0, int64_t nb, int64_t items, const CellEnt* clist, const ChunkEnt* klist,
                      double* part) -> int {
    BH_TRY(fill_u32(c, next, 0u, EV_XCDS));
    {
      ProfScope p(c, "k_group_eval_f32");
      const int64_t* coff = offs;
      const int64_t* koff = offs + ng + 1;
#define BH_EVAL(TPL, CH)                                                                                         \
  launch_eval<TPL, CH>(c, g0, nb, ioff, zeps, self, kahan, quad, m, gs, tpos, t, boxes, coff, clist, koff, klist, \
                       cnt, cnt + ng, eps2, gt0, err, next, part, items)
      if (gs == 64) {
        if (ch == 16) BH_EVAL(1, 16);
        else if (ch == 32) BH_EVAL(1, 32);
        else BH_EVAL(1, 64);
      } else {
        if (ch == 16) BH_EVAL(2, 16);
        else if (ch == 32) BH_EVAL(2, 32);
        else BH_EVAL(2, 64);
      }
#undef BH_EVAL
      BH_LAUNCH_CHECK();
    }
    {
      ProfScope p(c, "k_item_reduce");
      hipLaunchKernelGGL(k_item_reduce, dim3((unsigned)cdiv(nb * tpg, 256)), dim3(256), 0, c->stream, g0, nb, ioff,
                         m, gs, gt0, part, tpg, torig, t->dim, G, err, out);
      BH_LAUNCH_CHECK();
    }
    return BH_OK;
  };
  // the list lengths of this call, kept for the next one's regions
  auto keep_counts = [&](int64_t ntc_, int64_t ntk_) -> int {
    if (c->prev_cap < 2 * ng) {
      dfree(c, c->prev_cnt);
      c->prev_cnt = dalloc_t<int64_t>(c, (size_t)(2 * ng));
      c->prev_cap = c->prev_cnt ? 2 * ng : 0;
      c->prev_ng = -1;
      (void)hipGetLastError();
    }
    if (!c->prev_cnt) return BH_OK;
    BH_HIP(hipMemcpyAsync(c->prev_cnt, cnt, (size_t)(2 * ng) * sizeof(int64_t), hipMemcpyDeviceToDevice, c->stream));
    c->prev_ng = ng;
    c->prev_ntc = ntc_;
    c->prev_ntk = ntk_;
    return BH_OK;
  };
  if (plan || dcap > 0 || sized) {
    // The one-pass lists are in place: their layout, the work items and the
    // capacity checks come from the device, and the evaluation follows at
    // once (it does nothing if a check failed).  A planned walk stops here;
    // the host-scanned path reads the status once at the end and, if a list
    // outgrew its region or the items their buffer, redoes the lists below.
    const int64_t cap_items = std::max<int64_t>(plan ? plan->cap_items : c->items_hint, 1);
    int64_t* need = plan ? plan->need : s.get<int64_t>(6);
    double* part = s.get<double>((size_t)cap_items * tpg * 3);
    if (!need || !part) return BH_ENOMEM;
    hipLaunchKernelGGL(k_scan_lists, dim3(1), dim3(1024), 0, c->stream, ng, cnt, cpt, offs, ioff, dcap, (int)sized,
                       cap_items,
                       err, need, plan ? nullptr : need + 2, m, gs, gt0);
    BH_LAUNCH_CHECK();
    if (wnext) {
      uint32_t* hist = s.get<uint32_t>(ORD_BUCKETS);
      if (!hist) return BH_ENOMEM;
      BH_TRY(fill_u32(c, hist, 0u, ORD_BUCKETS));
      const unsigned ob = (unsigned)cdiv(ng, 256);
      hipLaunchKernelGGL(k_order_hist, dim3(ob), dim3(256), 0, c->stream, ng, cnt, hist);
      hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(64), 0, c->stream, hist);
      hipLaunchKernelGGL(k_order_scatter, dim3(ob), dim3(256), 0, c->stream, ng, cnt, hist, wnext);
      BH_LAUNCH_CHECK();
      if (!plan) c->walk_order_n = ng;
    }
    BH_TRY(evaluate(0, ng, cap_items, dclist, dklist, part));
    if (plan) return BH_OK;
    int64_t h[6];
    uint32_t herr = 0;
    BH_HIP(hipMemcpyAsync(h, need, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    BH_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
    BH_HIP(hipStreamSynchronize(c->stream));
    // the item buffer for the next call: what this one needed and a quarter
    // more
    if (!(herr & 1)) c->items_hint = std::max<int64_t>(1024, h[1] + h[1] / 4);
    if (!herr) {
      if (stats) {
        stats[0] = h[2];
        stats[1] = h[3] - (tgt ? 0 : m);  // each target's own pair contributes nothing
        stats[2] = ng;
        stats[3] = h[4] + h[5];
      }
      return keep_counts(h[4], h[5]);
    }
    BH_TRY(fill_u32(c, err, 0u, 1));
    if (herr & 1) {
      if (stack_cap >= MAX_STACK) {
        set_error("group walk stack overflow (stack_cap=%d)", stack_cap);
        return BH_EOVERFLOW;
      }
      stack_cap = std::min(MAX_STACK, 2 * stack_cap);
      keep_hint(stack_cap);
      BH_TRY(count_pass());
    }
  }
  // Host-scanned lists: materialised for batches of consecutive groups whose
  // lists fit the list budget (one batch unless the problem is very large);
  // each batch is filled, split into work items, evaluated and reduced in
  // turn.
  struct Batch {
    int64_t g0, g1, items;
  };
  std::vector<Batch> batches;
  std::vector<int64_t> ho, hc;
  CellEnt* clist;
  ChunkEnt* klist;
  int64_t cell_int = 0, pair_int = 0, ntc = 0, ntk = 0, max_items = 0;
  {
    hc.resize(3 * ng);
    for (;;) {
      uint32_t herr = 0;
      BH_HIP(hipMemcpyAsync(hc.data(), cnt, 3 * ng * 8, hipMemcpyDeviceToHost, c->stream));
      BH_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
      BH_HIP(hipStreamSynchronize(c->stream));
      if (!herr) break;
      if (stack_cap >= MAX_STACK) {
        set_error("group walk stack overflow (stack_cap=%d)", stack_cap);
        return BH_EOVERFLOW;
      }
      stack_cap = std::min(MAX_STACK, 2 * stack_cap);
      keep_hint(stack_cap);
      int rc = fill_u32(c, err, 0u, 1);
      if (!rc) rc = count_pass();
      if (rc) return rc;
    }
    const size_t budget = c->list_budget ? c->list_budget : default_list_budget(c);
    ho.assign(2 * (ng + 1), 0);
    int64_t maxc = 0, maxk = 0, bc = 0, bk = 0, bi = 0;
    size_t bytes = 0;
    batches.push_back({0, 0, 0});
    for (int64_t g = 0; g < ng; g++) {
      ho[g + 1] = ho[g] + hc[g];
      ho[ng + 1 + g + 1] = ho[ng + 1 + g] + hc[ng + g];
      const int64_t tg = gt0 ? t->h_gtab[g + 1] - t->h_gtab[g] : std::min<int64_t>(gs, m - g * gs);
      cell_int += hc[g] * tg;
      pair_int += hc[2 * ng + g] * tg;
      const size_t gb = hc[g] * sizeof(CellEnt) + hc[ng + g] * sizeof(ChunkEnt);
      if (g > batches.back().g0 && bytes + gb > budget) {
        batches.back().g1 = g;
        batches.back().items = bi;
        batches.push_back({g, 0, 0});
        maxc = std::max(maxc, bc);
        maxk = std::max(maxk, bk);
        bytes = 0;
        bc = bk = bi = 0;
      }
      bytes += gb;
      bc += hc[g];
      bk += hc[ng + g];
      bi += group_items(hc[g], hc[ng + g], cpt);
    }
    batches.back().g1 = ng;
    batches.back().items = bi;
    maxc = std::max(maxc, bc);
    maxk = std::max(maxk, bk);
    int64_t all_items = 0;
    for (const Batch& b : batches) {
      max_items = std::max(max_items, b.items);
      all_items += b.items;
    }
    c->items_hint = std::max<int64_t>(1024, all_items + all_items / 4);
    if (!tgt) pair_int -= m;  // each target's own pair contributes nothing
    for (int64_t g = 0; g < ng; g++) {
      ntc += hc[g];
      ntk += hc[ng + g];
    }
    clist = s.get<CellEnt>((size_t)std::max<int64_t>(maxc, 1));
    klist = s.get<ChunkEnt>((size_t)std::max<int64_t>(maxk, 1));
  }
  if (!clist || !klist) return BH_ENOMEM;
  double* part = s.get<double>((size_t)std::max<int64_t>(max_items, 1) * tpg * 3);
  if (!part) return BH_ENOMEM;
  std::vector<int64_t> rel;
  std::vector<int32_t> reli, hord, bucket;
  int32_t* ord = nullptr;
  {
    int64_t maxnb = 0;
    for (const Batch& b : batches) maxnb = std::max(maxnb, b.g1 - b.g0);
    ord = s.get<int32_t>((size_t)std::max<int64_t>(maxnb, 1));
    if (!ord) return BH_ENOMEM;
  }
  for (const Batch& b : batches) {
    const int64_t nb = b.g1 - b.g0;
    {
      // fill order: groups with the longest lists (the longest walks) first,
      // so the pass does not end on a few long walks started late; a
      // counting sort on the list length, in 4096 buckets
      int64_t kmax = 1;
      for (int64_t g = b.g0; g < b.g1; g++) kmax = std::max(kmax, hc[g] + hc[ng + g]);
      int sh = 0;
      while ((kmax >> sh) >= 4096) sh++;
      bucket.assign(4097, 0);
      for (int64_t g = b.g0; g < b.g1; g++) bucket[4095 - ((hc[g] + hc[ng + g]) >> sh)]++;
      for (int k = 0, acc = 0; k <= 4096; k++) {
        const int cnt_k = bucket[k];
        bucket[k] = acc;
        acc += cnt_k;
      }
      hord.resize(nb);
      for (int64_t g = b.g0; g < b.g1; g++) hord[bucket[4095 - ((hc[g] + hc[ng + g]) >> sh)]++] = (int32_t)(g - b.g0);
      BH_HIP(hipMemcpyAsync(ord, hord.data(), nb * 4, hipMemcpyHostToDevice, c->stream));
    }
    {
      // offsets of this batch's groups and items, relative to its first
      // list entries and first item
      rel.resize(2 * (nb + 1));
      reli.resize(nb + 1);
      int64_t it = 0;
      for (int64_t g = b.g0; g <= b.g1; g++) {
        rel[g - b.g0] = ho[g] - ho[b.g0];
        rel[nb + 1 + g - b.g0] = ho[ng + 1 + g] - ho[ng + 1 + b.g0];
        reli[g - b.g0] = (int32_t)it;
        if (g < b.g1) it += group_items(hc[g], hc[ng + g], cpt);
      }
      BH_HIP(hipMemcpyAsync(offs + b.g0, rel.data(), (nb + 1) * 8, hipMemcpyHostToDevice, c->stream));
      BH_HIP(hipMemcpyAsync(offs + ng + 1 + b.g0, rel.data() + nb + 1, (nb + 1) * 8, hipMemcpyHostToDevice,
                            c->stream));
      BH_HIP(hipMemcpyAsync(ioff, reli.data(), (nb + 1) * 4, hipMemcpyHostToDevice, c->stream));
      BH_HIP(hipStreamSynchronize(c->stream));  // rel, reli, hord are reused by the next batch
    }
    {
      ProfScope p(c, "k_group_walk_fill");
      launch_walk<true>(c, t, dim3((unsigned)cdiv(nb, wpb)), lds, gs, boxes, itheta2, stack_cap,
                        nullptr, nullptr, nullptr, offs, offs + ng + 1, clist, klist, err, groot, b.g0, nb, wpb, tpos,
                        m, gt0, ord);
      BH_LAUNCH_CHECK();
    }
    BH_TRY(evaluate(b.g0, nb, b.items, clist, klist, part));
  }
  {
    // the fill walks repeat the count pass's walks with the same stack, so
    // they cannot overflow; were a bit set anyway, the evaluation and the
    // reduction skipped and out holds no result
    uint32_t herr = 0;
    BH_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
    BH_HIP(hipStreamSynchronize(c->stream));
    if (herr) {
      set_error("FP32 group walk reported status %u while filling its lists", herr);
      return BH_EOVERFLOW;
    }
  }
  if (stats) {
    stats[0] = cell_int;
    stats[1] = pair_int;
    stats[2] = ng;
    stats[3] = ntc + ntk;
  }
  return keep_counts(ntc, ntk);
}

}  // namespace bh
