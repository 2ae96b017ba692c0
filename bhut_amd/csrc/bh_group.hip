// bh_group.hip -- the FP32 grouped form (the chosen, fast form).
//
// Same tree, same monopole-plus-leaf physics as bhut (bhut/traverse/bh.py:
// 485-548: leaves are always summed particle by particle, internal nodes are
// accepted as monopoles or opened), evaluated for groups of GS Morton-
// contiguous targets at once:
//
//  1. k_group_box: FP64 bounding box of each group's targets.
//  2. k_group_walk: one wavefront per group walks the tree breadth-wise
//     with an LDS stack: 64 nodes are tested per step, one per lane, with a
//     conservative group test -- accept a node only if size < theta * dmin,
//     dmin the distance from its centre of mass to the group box, so every
//     target of the group would accept it under bhut's own test
//     (bh.py:586-601).  Lanes are compacted with ballots into the monopole
//     list, the leaf list and the stack.  Run twice: count, then fill.
//  3. k_group_eval: one wavefront per group, one (or two) targets per lane,
//     over the lists in HBM.  List entries and source particles sit at
//     wave-uniform addresses and are fetched with scalar loads; every
//     interaction is FP32 on coordinates relative to the group centre, with
//     sources stored relative to their leaf centre and the leaf-to-group
//     shift formed in FP64, so precision follows the size of the leaf, not
//     the size of the coordinates.  Per-leaf (and per-32-monopole) partial
//     sums are added to the per-target total with Kahan compensation.
//
// FP64 stays in the tree build (box, quantisation, mass and centre of mass)
// and in the group-centre/leaf-shift subtractions.
#include "bh_common.h"

namespace bh {

struct GroupBox {
  double cx, cy, cz, hx, hy, hz;
};

struct __attribute__((aligned(16))) LeafEnt {
  float sx, sy, sz;  // leaf centre - group centre
  int32_t start;
  int32_t count;
  int32_t pad[3];
};
static_assert(sizeof(LeafEnt) == 32, "LeafEnt layout");

// ---------------------------------------------------------------- prepare
// Source particles relative to their leaf's centre, FP32.
template <int DIM>
__global__ __launch_bounds__(256) void k_src32(int64_t nleaves, const int32_t* __restrict__ leaf_order,
                                               const int32_t* __restrict__ start,
                                               const int32_t* __restrict__ count,
                                               const double* __restrict__ center,
                                               const double4* __restrict__ spart,
                                               float4* __restrict__ src32) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nleaves) return;
  const int v = leaf_order[t];
  const double cx = center[v * DIM], cy = center[v * DIM + 1], cz = DIM == 3 ? center[v * DIM + 2] : 0.0;
  const int64_t s = start[v], c = count[v];
  for (int64_t j = s; j < s + c; j++) {
    const double4 p = spart[j];
    src32[j] = make_float4((float)(p.x - cx), (float)(p.y - cy), (float)(p.z - cz), (float)p.w);
  }
}

int f32_prepare(bh_ctx* c, bh_tree* t) {
  if (!t->spart32) {
    t->spart32 = dalloc_t<float4>(c, (size_t)t->n);
    if (!t->spart32) return BH_ENOMEM;
  }
  const unsigned nb = (unsigned)cdiv(t->nleaves, 256);
  if (t->dim == 3)
    hipLaunchKernelGGL(k_src32<3>, dim3(nb), dim3(256), 0, c->stream, t->nleaves, t->leaf_order, t->start,
                       t->count, t->center, t->spart, t->spart32);
  else
    hipLaunchKernelGGL(k_src32<2>, dim3(nb), dim3(256), 0, c->stream, t->nleaves, t->leaf_order, t->start,
                       t->count, t->center, t->spart, t->spart32);
  BH_LAUNCH_CHECK();
  t->f32_ready = true;
  return BH_OK;
}

// ---------------------------------------------------------------- groups
__global__ __launch_bounds__(64) void k_group_box(const double4* __restrict__ tpos, int64_t m, int gs,
                                                  GroupBox* __restrict__ boxes) {
  const int64_t g = blockIdx.x;
  const int64_t t0 = g * gs;
  const int64_t t1 = t0 + gs < m ? t0 + gs : m;
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
template <int DIM, bool FILL>
__global__ __launch_bounds__(64) void k_group_walk(const NodeD* __restrict__ nodes,
                                                   const double* __restrict__ center, int64_t ngroups,
                                                   const GroupBox* __restrict__ boxes, double theta2,
                                                   int stack_cap, int64_t* __restrict__ mcount,
                                                   int64_t* __restrict__ lcount,
                                                   int64_t* __restrict__ pcount,
                                                   const int64_t* __restrict__ moff,
                                                   const int64_t* __restrict__ loff,
                                                   float4* __restrict__ mlist,
                                                   LeafEnt* __restrict__ llist,
                                                   uint32_t* __restrict__ err) {
  extern __shared__ int32_t stack[];
  const int64_t g = blockIdx.x;
  const int lane = threadIdx.x;
  const GroupBox B = boxes[g];
  if (lane == 0) stack[0] = 0;
  __syncthreads();
  int sp = 1;
  int64_t nm = 0, nl = 0, np = 0;
  const int64_t mb = FILL ? moff[g] : 0, me = FILL ? moff[g + 1] : 0;
  const int64_t lb = FILL ? loff[g] : 0, le = FILL ? loff[g + 1] : 0;
  const uint64_t lt = lanemask_lt();
  while (sp > 0) {
    const int k = sp < 64 ? sp : 64;
    int v = -1;
    if (lane < k) v = stack[sp - k + lane];
    __syncthreads();
    sp -= k;
    int kind = 0;  // 1 monopole, 2 leaf, 3 open
    NodeD nd;
    if (v >= 0) {
      nd = nodes[v];
      if (nd.m != 0.0) {
        if (nd.fc < 0) {
          kind = 2;
        } else {
          double ex = fabs(nd.cx - B.cx) - B.hx;
          double ey = fabs(nd.cy - B.cy) - B.hy;
          double ez = DIM == 3 ? fabs(nd.cz - B.cz) - B.hz : 0.0;
          ex = ex > 0.0 ? ex : 0.0;
          ey = ey > 0.0 ? ey : 0.0;
          ez = ez > 0.0 ? ez : 0.0;
          const double d2 = ex * ex + ey * ey + ez * ez;
          const bool open = !(d2 > 0.0) || !(nd.size * nd.size < theta2 * d2);
          kind = open ? 3 : 1;
        }
      }
    }
    const uint64_t bm = __ballot(kind == 1), bl = __ballot(kind == 2), bo = __ballot(kind == 3);
    if (kind == 3) {
      const int o = sp + 2 * __popcll(bo & lt);
      if (o + 1 < stack_cap) {
        stack[o] = nd.fc;
        stack[o + 1] = nd.fc + 1;
      }
    }
    if (FILL) {
      if (kind == 1) {
        const int64_t o = mb + nm + __popcll(bm & lt);
        if (o < me)
          mlist[o] = make_float4((float)(nd.cx - B.cx), (float)(nd.cy - B.cy), (float)(nd.cz - B.cz), (float)nd.m);
      }
      if (kind == 2) {
        const int64_t o = lb + nl + __popcll(bl & lt);
        if (o < le) {
          LeafEnt e;
          e.sx = (float)(center[(int64_t)v * DIM] - B.cx);
          e.sy = (float)(center[(int64_t)v * DIM + 1] - B.cy);
          e.sz = DIM == 3 ? (float)(center[(int64_t)v * DIM + 2] - B.cz) : 0.f;
          e.start = nd.start;
          e.count = nd.count;
          e.pad[0] = e.pad[1] = e.pad[2] = 0;
          llist[o] = e;
        }
      }
    } else {
      // particles summed directly: wave total of leaf counts
      int pc = kind == 2 ? nd.count : 0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) pc += __shfl_xor(pc, off);
      np += pc;
    }
    nm += __popcll(bm);
    nl += __popcll(bl);
    const int push = 2 * __popcll(bo);
    if (sp + push > stack_cap) {
      if (lane == 0) atomicOr(err, 1u);
      return;
    }
    sp += push;
    __syncthreads();
  }
  if (!FILL && lane == 0) {
    mcount[g] = nm;
    lcount[g] = nl;
    pcount[g] = np;
  }
}

// ---------------------------------------------------------------- eval
template <bool ZEPS>
__device__ __forceinline__ void interact(float dx, float dy, float dz, float m, float eps2, float& px, float& py,
                                         float& pz) {
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
  float ri = __builtin_amdgcn_rsqf(r2);
  if (ZEPS) ri = r2 > 0.f ? ri : 0.f;
  const float mg = m * (ri * ri * ri);
  px = fmaf(mg, dx, px);
  py = fmaf(mg, dy, py);
  pz = fmaf(mg, dz, pz);
}

__device__ __forceinline__ void kahan_add(float v, float& s, float& c) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <int TPL, bool ZEPS, bool SELF, bool KAHAN>
__global__ __launch_bounds__(64) void k_group_eval(int64_t m, int gs, const double4* __restrict__ tpos,
                                                   const uint32_t* __restrict__ torig,
                                                   const GroupBox* __restrict__ boxes,
                                                   const int64_t* __restrict__ moff,
                                                   const float4* __restrict__ mlist,
                                                   const int64_t* __restrict__ loff,
                                                   const LeafEnt* __restrict__ llist,
                                                   const float4* __restrict__ src32, float eps2, double G,
                                                   int dim, double* __restrict__ out) {
  const int64_t g = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t t0 = g * gs;
  const int64_t t1 = t0 + gs < m ? t0 + gs : m;
  const GroupBox B = boxes[g];
  float tx[TPL], ty[TPL], tz[TPL];
  int64_t ti[TPL];
  float sx[TPL], sy[TPL], sz[TPL], cx[TPL], cy[TPL], cz[TPL];
#pragma unroll
  for (int q = 0; q < TPL; q++) {
    ti[q] = t0 + lane + 64 * q;
    const double4 p = tpos[ti[q] < t1 ? ti[q] : t0];
    tx[q] = (float)(p.x - B.cx);
    ty[q] = (float)(p.y - B.cy);
    tz[q] = (float)(p.z - B.cz);
    sx[q] = sy[q] = sz[q] = 0.f;
    cx[q] = cy[q] = cz[q] = 0.f;
  }
  // monopoles, flushed every 32
  const int64_t m0 = moff[g], m1 = moff[g + 1];
  for (int64_t e = m0; e < m1; e += 32) {
    const int64_t ee = e + 32 < m1 ? e + 32 : m1;
    float px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int q = 0; q < TPL; q++) px[q] = py[q] = pz[q] = 0.f;
#pragma unroll 4
    for (int64_t k = e; k < ee; k++) {
      const float4 c = mlist[k];
#pragma unroll
      for (int q = 0; q < TPL; q++) interact<ZEPS>(c.x - tx[q], c.y - ty[q], c.z - tz[q], c.w, eps2, px[q], py[q], pz[q]);
    }
#pragma unroll
    for (int q = 0; q < TPL; q++) {
      if (KAHAN) {
        kahan_add(px[q], sx[q], cx[q]);
        kahan_add(py[q], sy[q], cy[q]);
        kahan_add(pz[q], sz[q], cz[q]);
      } else {
        sx[q] += px[q];
        sy[q] += py[q];
        sz[q] += pz[q];
      }
    }
  }
  // leaves
  const int64_t l0 = loff[g], l1 = loff[g + 1];
  for (int64_t e = l0; e < l1; e++) {
    const LeafEnt L = llist[e];
    float ux[TPL], uy[TPL], uz[TPL], px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int q = 0; q < TPL; q++) {
      ux[q] = tx[q] - L.sx;
      uy[q] = ty[q] - L.sy;
      uz[q] = tz[q] - L.sz;
      px[q] = py[q] = pz[q] = 0.f;
    }
    const int32_t j0 = L.start, j1 = L.start + L.count;
    const bool overlap = SELF && j0 < t1 && t0 < j1;
    if (!overlap) {
#pragma unroll 4
      for (int32_t j = j0; j < j1; j++) {
        const float4 s = src32[j];
#pragma unroll
        for (int q = 0; q < TPL; q++) interact<ZEPS>(s.x - ux[q], s.y - uy[q], s.z - uz[q], s.w, eps2, px[q], py[q], pz[q]);
      }
    } else {
      for (int32_t j = j0; j < j1; j++) {
        const float4 s = src32[j];
#pragma unroll
        for (int q = 0; q < TPL; q++) {
          const float w = j == ti[q] ? 0.f : s.w;  // skip the target itself
          const float dx = s.x - ux[q], dy = s.y - uy[q], dz = s.z - uz[q];
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
          float ri = __builtin_amdgcn_rsqf(r2);
          ri = (j == ti[q] || (ZEPS && !(r2 > 0.f))) ? 0.f : ri;
          const float mg = w * (ri * ri * ri);
          px[q] = fmaf(mg, dx, px[q]);
          py[q] = fmaf(mg, dy, py[q]);
          pz[q] = fmaf(mg, dz, pz[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < TPL; q++) {
      if (KAHAN) {
        kahan_add(px[q], sx[q], cx[q]);
        kahan_add(py[q], sy[q], cy[q]);
        kahan_add(pz[q], sz[q], cz[q]);
      } else {
        sx[q] += px[q];
        sy[q] += py[q];
        sz[q] += pz[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < TPL; q++) {
    if (ti[q] < t1) {
      const int64_t o = torig[ti[q]];
      out[o * dim] = G * ((double)sx[q] - (double)cx[q]);
      out[o * dim + 1] = G * ((double)sy[q] - (double)cy[q]);
      if (dim == 3) out[o * dim + 2] = G * ((double)sz[q] - (double)cz[q]);
    }
  }
}

// ---------------------------------------------------------------- host
int walk_f32(bh_ctx* c, bh_tree* t, const double* tgt, int64_t m, double theta, double eps, double G,
             const bh_f32_opts* o, double* out, int64_t* stats) {
  if (m == 0) return BH_OK;
  const int gs = o && o->group_size ? o->group_size : 64;
  const bool kahan = o ? o->kahan != 0 : true;
  const int stack_cap = o && o->stack_cap ? o->stack_cap : 2048;
  if (gs != 64 && gs != 128) {
    set_error("group_size must be 64 or 128, got %d", gs);
    return BH_EINVAL;
  }
  if (stack_cap < 64 || stack_cap > 8192) {
    set_error("stack_cap must be in [64, 8192], got %d", stack_cap);
    return BH_EINVAL;
  }
  if (!t->f32_ready) {
    int rc = f32_prepare(c, t);
    if (rc) return rc;
  }
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
  const int64_t ng = cdiv(m, gs);
  GroupBox* boxes = s.get<GroupBox>(ng);
  int64_t* cnt = s.get<int64_t>(3 * ng);
  int64_t* offs = s.get<int64_t>(2 * (ng + 1));
  uint32_t* err = s.get<uint32_t>(1);
  if (!boxes || !cnt || !offs || !err) return BH_ENOMEM;
  int rc = fill_u32(c, err, 0u, 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_group_box, dim3((unsigned)ng), dim3(64), 0, c->stream, tpos, m, gs, boxes);
  const double theta2 = theta * theta;
  const size_t lds = (size_t)stack_cap * sizeof(int32_t);
  {
    ProfScope p(c, "k_group_walk_count");
    if (t->dim == 3)
      hipLaunchKernelGGL((k_group_walk<3, false>), dim3((unsigned)ng), dim3(64), lds, c->stream, t->nodes,
                         t->center, ng, boxes, theta2, stack_cap, cnt, cnt + ng, cnt + 2 * ng, nullptr, nullptr,
                         nullptr, nullptr, err);
    else
      hipLaunchKernelGGL((k_group_walk<2, false>), dim3((unsigned)ng), dim3(64), lds, c->stream, t->nodes,
                         t->center, ng, boxes, theta2, stack_cap, cnt, cnt + ng, cnt + 2 * ng, nullptr, nullptr,
                         nullptr, nullptr, err);
    BH_LAUNCH_CHECK();
  }
  std::vector<int64_t> hc(3 * ng);
  uint32_t herr = 0;
  BH_HIP(hipMemcpyAsync(hc.data(), cnt, 3 * ng * 8, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
  BH_HIP(hipStreamSynchronize(c->stream));
  if (herr) {
    set_error("group walk stack overflow (stack_cap=%d); raise bh_f32_opts.stack_cap", stack_cap);
    return BH_EOVERFLOW;
  }
  std::vector<int64_t> ho(2 * (ng + 1));
  int64_t mono_int = 0, pair_int = 0;
  ho[0] = 0;
  ho[ng + 1] = 0;
  for (int64_t g = 0; g < ng; g++) {
    ho[g + 1] = ho[g] + hc[g];
    ho[ng + 1 + g + 1] = ho[ng + 1 + g] + hc[ng + g];
    const int64_t tg = std::min<int64_t>(gs, m - g * gs);
    mono_int += hc[g] * tg;
    pair_int += hc[2 * ng + g] * tg;
  }
  if (!tgt) pair_int -= m;  // each target skips itself once in its own leaf
  const int64_t ntm = ho[ng], ntl = ho[2 * ng + 1];
  BH_HIP(hipMemcpyAsync(offs, ho.data(), 2 * (ng + 1) * 8, hipMemcpyHostToDevice, c->stream));
  float4* mlist = s.get<float4>((size_t)std::max<int64_t>(ntm, 1));
  LeafEnt* llist = s.get<LeafEnt>((size_t)std::max<int64_t>(ntl, 1));
  if (!mlist || !llist) return BH_ENOMEM;
  {
    ProfScope p(c, "k_group_walk_fill");
    if (t->dim == 3)
      hipLaunchKernelGGL((k_group_walk<3, true>), dim3((unsigned)ng), dim3(64), lds, c->stream, t->nodes, t->center,
                         ng, boxes, theta2, stack_cap, nullptr, nullptr, nullptr, offs, offs + ng + 1, mlist, llist,
                         err);
    else
      hipLaunchKernelGGL((k_group_walk<2, true>), dim3((unsigned)ng), dim3(64), lds, c->stream, t->nodes, t->center,
                         ng, boxes, theta2, stack_cap, nullptr, nullptr, nullptr, offs, offs + ng + 1, mlist, llist,
                         err);
    BH_LAUNCH_CHECK();
  }
  const float eps2 = (float)(eps * eps);
  const bool zeps = !(eps2 > 0.f);
  const bool self = tgt == nullptr;
  {
    ProfScope p(c, "k_group_eval_f32");
#define BH_EV(TPL, Z, S, K)                                                                                  \
  hipLaunchKernelGGL((k_group_eval<TPL, Z, S, K>), dim3((unsigned)ng), dim3(64), 0, c->stream, m, gs, tpos, torig, \
                     boxes, offs, mlist, offs + ng + 1, llist, t->spart32, eps2, G, t->dim, out)
#define BH_EV_K(TPL, Z, S) \
  if (kahan) BH_EV(TPL, Z, S, true); else BH_EV(TPL, Z, S, false)
#define BH_EV_S(TPL, Z) \
  if (self) { BH_EV_K(TPL, Z, true); } else { BH_EV_K(TPL, Z, false); }
#define BH_EV_Z(TPL) \
  if (zeps) { BH_EV_S(TPL, true); } else { BH_EV_S(TPL, false); }
    if (gs == 64) {
      BH_EV_Z(1);
    } else {
      BH_EV_Z(2);
    }
#undef BH_EV_Z
#undef BH_EV_S
#undef BH_EV_K
#undef BH_EV
    BH_LAUNCH_CHECK();
  }
  if (stats) {
    stats[0] = mono_int;
    stats[1] = pair_int;
    stats[2] = ng;
    stats[3] = ntm + ntl;
  }
  return BH_OK;
}

}  // namespace bh
