// bh_direct.hip -- FP64 direct summation (accelerations and potential).
//
// The theta = 0 limit of the walk (bhut/traverse/bh.py:600-601 opens every
// node), used as the accuracy reference for the FP32 form and for the
// energy of the leapfrog driver.  Sources are staged through LDS in tiles of
// 256 (coalesced double4 reads); the source range is split into chunks over
// blockIdx.y so small target sets still fill the 256 CUs, and the chunk
// partials are added in chunk order so the result is deterministic.
#include "bh_common.h"

namespace bh {

constexpr int DS_T = 256;

template <int DIM, bool SELF>
__global__ __launch_bounds__(DS_T) void k_direct(const double4* __restrict__ src, int64_t n,
                                                 const double* __restrict__ tgt, int64_t m,
                                                 int64_t chunk, double eps2, double G,
                                                 double* __restrict__ part_acc,
                                                 double* __restrict__ part_pot) {
  __shared__ double4 tile[DS_T];
  const int64_t i = blockIdx.x * (int64_t)DS_T + threadIdx.x;
  const bool valid = i < m;
  double tx = 0, ty = 0, tz = 0;
  if (valid) {
    tx = tgt[i * DIM];
    ty = tgt[i * DIM + 1];
    if (DIM == 3) tz = tgt[i * DIM + 2];
  }
  const int64_t j0 = blockIdx.y * chunk;
  const int64_t j1 = j0 + chunk < n ? j0 + chunk : n;
  double ax = 0, ay = 0, az = 0, phi = 0;
  for (int64_t b = j0; b < j1; b += DS_T) {
    __syncthreads();
    if (b + threadIdx.x < j1) tile[threadIdx.x] = src[b + threadIdx.x];
    __syncthreads();
    const int cnt = (int)(j1 - b < DS_T ? j1 - b : DS_T);
    for (int k = 0; k < cnt; k++) {
      const double4 q = tile[k];
      if (SELF && b + k == i) continue;
      const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
      double rsq = dx * dx;
      rsq = rsq + dy * dy;
      if (DIM == 3) rsq = rsq + dz * dz;
      rsq = rsq + eps2;
      if (rsq == 0.0) continue;
      const double rinv = 1.0 / sqrt(rsq);
      const double gm = G * q.w;
      const double mag = gm * (rinv * rinv * rinv);
      ax = ax + mag * dx;
      ay = ay + mag * dy;
      if (DIM == 3) az = az + mag * dz;
      phi = phi - gm * rinv;
    }
  }
  if (valid) {
    const int64_t o = (int64_t)blockIdx.y * m + i;
    if (part_acc) {
      part_acc[o * DIM] = ax;
      part_acc[o * DIM + 1] = ay;
      if (DIM == 3) part_acc[o * DIM + 2] = az;
    }
    if (part_pot) part_pot[o] = phi;
  }
}

template <int DIM>
__global__ void k_direct_reduce(const double* __restrict__ part_acc,
                                const double* __restrict__ part_pot, int64_t m, int nchunks,
                                double* __restrict__ acc, double* __restrict__ pot) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (acc) {
    for (int d = 0; d < DIM; d++) {
      double s = 0.0;
      for (int k = 0; k < nchunks; k++) s = s + part_acc[((int64_t)k * m + i) * DIM + d];
      acc[i * DIM + d] = s;
    }
  }
  if (pot) {
    double s = 0.0;
    for (int k = 0; k < nchunks; k++) s = s + part_pot[(int64_t)k * m + i];
    pot[i] = s;
  }
}

__global__ void k_pack_src(const double* __restrict__ pos, const double* __restrict__ mass,
                           int64_t n, int dim, double4* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 r;
  r.x = pos[i * dim];
  r.y = pos[i * dim + 1];
  r.z = dim == 3 ? pos[i * dim + 2] : 0.0;
  r.w = mass[i];
  out[i] = r;
}

int direct(bh_ctx* c, const double* src, const double* mass, int64_t n, const double* tgt,
           int64_t m, int dim, int self_mode, double eps, double G, double* acc, double* pot) {
  if (m == 0) return BH_OK;
  Scratch s(c);
  double4* s4 = s.get<double4>(n);
  if (!s4) return BH_ENOMEM;
  if (n > 0)
    hipLaunchKernelGGL(k_pack_src, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, c->stream, src, mass,
                       n, dim, s4);
  const int64_t tblocks = cdiv(m, DS_T);
  int64_t nchunks = std::max<int64_t>(1, std::min<int64_t>(cdiv(2048, tblocks), cdiv(n, 4 * DS_T)));
  nchunks = std::min<int64_t>(nchunks, 1024);
  const int64_t chunk = std::max<int64_t>(cdiv(std::max<int64_t>(n, 1), nchunks), 1);
  nchunks = cdiv(std::max<int64_t>(n, 1), chunk);
  double* pa = acc ? s.get<double>((size_t)nchunks * m * dim) : nullptr;
  double* pp = pot ? s.get<double>((size_t)nchunks * m) : nullptr;
  if ((acc && !pa) || (pot && !pp)) return BH_ENOMEM;
  const double eps2 = eps * eps;
  const dim3 grid((unsigned)tblocks, (unsigned)nchunks);
#define BH_D(D, S)                                                                               \
  hipLaunchKernelGGL((k_direct<D, S>), grid, dim3(DS_T), 0, c->stream, s4, n, tgt, m, chunk, eps2, \
                     G, pa, pp)
  if (dim == 3) {
    if (self_mode) BH_D(3, true); else BH_D(3, false);
  } else {
    if (self_mode) BH_D(2, true); else BH_D(2, false);
  }
#undef BH_D
  if (dim == 3)
    hipLaunchKernelGGL(k_direct_reduce<3>, dim3((unsigned)cdiv(m, 256)), dim3(256), 0, c->stream, pa,
                       pp, m, (int)nchunks, acc, pot);
  else
    hipLaunchKernelGGL(k_direct_reduce<2>, dim3((unsigned)cdiv(m, 256)), dim3(256), 0, c->stream, pa,
                       pp, m, (int)nchunks, acc, pot);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

}  // namespace bh
