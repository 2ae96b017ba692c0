// bh_sort.hip -- bounding box, Morton encoding, stable LSD radix sort, gather.
//
// Reference steps (maroba/bhut): aabb (bhut/space/bbox.py:14-46),
// normalize_to_unit (bbox.py:49-96), morton_codes (bhut/utils/morton.py:48-102),
// np.argsort(kind='stable') (morton.py:135), positions[perm] (build.py:91-92).
//
// All of these are HBM-bound streaming passes.  The sort is an LSD radix
// sort over 8-bit digits with per-wave LDS histograms, a per-digit row scan
// across tiles, and a scatter whose in-tile ranks come from wavefront ballots
// (eight ballots match the lanes holding the same digit), which keeps equal
// keys in input order -- the stability np.argsort(kind='stable') guarantees.
#include "bh_common.h"

#include <algorithm>
#include <thread>

namespace bh {

// ---------------------------------------------------------------- bbox
template <int DIM>
__global__ __launch_bounds__(256) void k_bbox_partial(const double* __restrict__ pos, int64_t n,
                                                      double* __restrict__ part) {
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    lo[d] = __builtin_inf();
    hi[d] = -__builtin_inf();
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      double v = pos[i * DIM + d];
      lo[d] = np_min(lo[d], v);
      hi[d] = np_max(hi[d], v);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      lo[d] = np_min(lo[d], __shfl_xor(lo[d], off));
      hi[d] = np_max(hi[d], __shfl_xor(hi[d], off));
    }
  }
  __shared__ double s[4][2 * DIM];
  const int w = threadIdx.x / WAVE;
  if (__lane_id() == 0) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      s[w][d] = lo[d];
      s[w][DIM + d] = hi[d];
    }
  }
  __syncthreads();
  if (threadIdx.x < DIM) {
    const int d = threadIdx.x;
    double a = s[0][d], b = s[0][DIM + d];
    for (int k = 1; k < (int)(blockDim.x / WAVE); k++) {
      a = np_min(a, s[k][d]);
      b = np_max(b, s[k][DIM + d]);
    }
    part[blockIdx.x * 2 * DIM + d] = a;
    part[blockIdx.x * 2 * DIM + DIM + d] = b;
  }
}

template <int DIM>
__global__ __launch_bounds__(64) void k_bbox_final(const double* __restrict__ part, int nb,
                                                   double* __restrict__ box) {
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    lo[d] = __builtin_inf();
    hi[d] = -__builtin_inf();
  }
  for (int b = threadIdx.x; b < nb; b += WAVE) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      lo[d] = np_min(lo[d], part[b * 2 * DIM + d]);
      hi[d] = np_max(hi[d], part[b * 2 * DIM + DIM + d]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      lo[d] = np_min(lo[d], __shfl_xor(lo[d], off));
      hi[d] = np_max(hi[d], __shfl_xor(hi[d], off));
    }
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      box[d] = lo[d];
      box[DIM + d] = hi[d];
    }
  }
}

int bbox(bh_ctx* c, const double* pos, int64_t n, int dim, double* d_box) {
  const int nb = (int)std::min<int64_t>(std::max<int64_t>(cdiv(n, 256 * 8), 1), 2048);
  Scratch s(c);
  double* part = s.get<double>((size_t)nb * 2 * dim);
  if (!part) return BH_ENOMEM;
  if (dim == 3) {
    hipLaunchKernelGGL(k_bbox_partial<3>, dim3(nb), dim3(256), 0, c->stream, pos, n, part);
    hipLaunchKernelGGL(k_bbox_final<3>, dim3(1), dim3(64), 0, c->stream, part, nb, d_box);
  } else {
    hipLaunchKernelGGL(k_bbox_partial<2>, dim3(nb), dim3(256), 0, c->stream, pos, n, part);
    hipLaunchKernelGGL(k_bbox_final<2>, dim3(1), dim3(64), 0, c->stream, part, nb, d_box);
  }
  BH_LAUNCH_CHECK();
  return BH_OK;
}

// ---------------------------------------------------------------- Morton
// Bit spreading for 21-bit coordinates: the standard shift-and-mask
// interleave (each step halves the run length of consecutive bits).
__host__ __device__ __forceinline__ uint64_t spread3(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__host__ __device__ __forceinline__ uint64_t spread2(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

// normalize_to_unit + quantisation (bbox.py:85-94, morton.py:86-89).  A NaN
// coordinate quantises to 2^21-1, which is what the x86-64 float->uint64
// conversion NumPy uses (2^63) becomes after the clamp.  clamp=1 is for
// external targets, which may lie outside the box: they are clamped into it
// (their codes only order the targets for coherent waves).
// Also compiled for the host (host_target_order): the same IEEE operations,
// with floating-point contraction off in both.
template <int DIM>
__host__ __device__ __forceinline__ uint64_t morton_code(const double* __restrict__ p,
                                                         const double* __restrict__ box, int clamp) {
  const double cap = 1.0 - 1e-12;
  uint64_t code = 0;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const double lo = box[d], hi = box[DIM + d];
    const double extent = hi - lo;
    const double ext = (extent <= 0) ? 1.0 : extent;
    const double off = (extent <= 0) ? 0.5 : 0.0;
    double v = (p[d] - lo) / ext + off;
    v = (v < cap || v != v) ? v : cap;
    if (clamp) v = (v > 0.0) ? v : 0.0;  // also maps NaN to 0
    uint64_t q;
    if (v != v) {
      q = (1ull << 21) - 1;
    } else {
      const double s = v * 2097152.0;
      q = (s <= 0.0) ? 0ull : (uint64_t)s;
      q = q > ((1ull << 21) - 1) ? ((1ull << 21) - 1) : q;
    }
    code |= (DIM == 3 ? spread3(q) : spread2(q)) << d;
  }
  return code;
}

// Morton code of particle i relative to its own system's box.
template <int DIM>
__global__ __launch_bounds__(256) void k_morton_seg(const double* __restrict__ pos, int64_t n,
                                                    const int32_t* __restrict__ poff, int64_t nsys,
                                                    const double* __restrict__ boxes, uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = find_sys(poff, nsys, i);
  keys[i] = morton_code<DIM>(pos + i * DIM, boxes + s * 2 * DIM, 0);
  vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_system_keys(const uint32_t* __restrict__ vals, int64_t n,
                                                     const int32_t* __restrict__ poff, int64_t nsys,
                                                     uint64_t* __restrict__ keys) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint64_t)find_sys(poff, nsys, vals[i]);
}

// One workgroup per system: NaN-propagating box of its particle run.
template <int DIM>
__global__ __launch_bounds__(256) void k_bbox_seg(const double* __restrict__ pos, const int32_t* __restrict__ poff,
                                                  double* __restrict__ boxes) {
  const int64_t s = blockIdx.x;
  const int64_t i0 = poff[s], i1 = poff[s + 1];
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    lo[d] = __builtin_inf();
    hi[d] = -__builtin_inf();
  }
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const double v = pos[i * DIM + d];
      lo[d] = np_min(lo[d], v);
      hi[d] = np_max(hi[d], v);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      lo[d] = np_min(lo[d], __shfl_xor(lo[d], off));
      hi[d] = np_max(hi[d], __shfl_xor(hi[d], off));
    }
  }
  __shared__ double sh[4][2 * DIM];
  const int w = threadIdx.x / WAVE;
  if (__lane_id() == 0) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      sh[w][d] = lo[d];
      sh[w][DIM + d] = hi[d];
    }
  }
  __syncthreads();
  if (threadIdx.x < DIM) {
    const int d = threadIdx.x;
    double a = sh[0][d], b = sh[0][DIM + d];
    for (int k = 1; k < 4; k++) {
      a = np_min(a, sh[k][d]);
      b = np_max(b, sh[k][DIM + d]);
    }
    boxes[s * 2 * DIM + d] = a;
    boxes[s * 2 * DIM + DIM + d] = b;
  }
}

int bbox_seg(bh_ctx* c, const double* pos, const int32_t* d_poff, int64_t nsys, int dim, double* d_boxes) {
  if (dim == 3)
    hipLaunchKernelGGL(k_bbox_seg<3>, dim3((unsigned)nsys), dim3(256), 0, c->stream, pos, d_poff, d_boxes);
  else
    hipLaunchKernelGGL(k_bbox_seg<2>, dim3((unsigned)nsys), dim3(256), 0, c->stream, pos, d_poff, d_boxes);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

int morton_seg(bh_ctx* c, const double* pos, int64_t n, int dim, const int32_t* d_poff, int64_t nsys,
               const double* d_boxes, uint64_t* keys, uint32_t* vals) {
  if (n == 0) return BH_OK;
  const unsigned nb = (unsigned)cdiv(n, 256);
  if (dim == 3)
    hipLaunchKernelGGL(k_morton_seg<3>, dim3(nb), dim3(256), 0, c->stream, pos, n, d_poff, nsys, d_boxes, keys, vals);
  else
    hipLaunchKernelGGL(k_morton_seg<2>, dim3(nb), dim3(256), 0, c->stream, pos, n, d_poff, nsys, d_boxes, keys, vals);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

int system_keys(bh_ctx* c, const uint32_t* vals, int64_t n, const int32_t* d_poff, int64_t nsys, uint64_t* keys) {
  if (n == 0) return BH_OK;
  hipLaunchKernelGGL(k_system_keys, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, c->stream, vals, n, d_poff, nsys,
                     keys);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_morton(const double* __restrict__ pos, int64_t n,
                                                const double* __restrict__ box, int clamp,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = morton_code<DIM>(pos + i * DIM, box, clamp);
  vals[i] = (uint32_t)i;
}

int morton_iota(bh_ctx* c, const double* pos, int64_t n, int dim, const double* d_box, int clamp,
                uint64_t* keys, uint32_t* vals) {
  if (n == 0) return BH_OK;
  const int nb = (int)cdiv(n, 256);
  if (dim == 3)
    hipLaunchKernelGGL(k_morton<3>, dim3(nb), dim3(256), 0, c->stream, pos, n, d_box, clamp, keys, vals);
  else
    hipLaunchKernelGGL(k_morton<2>, dim3(nb), dim3(256), 0, c->stream, pos, n, d_box, clamp, keys, vals);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

// Stable LSD radix sort of the targets' codes on the host threads: 8-bit
// digits; each thread counts and scatters its own contiguous block, and the
// blocks' offsets within a digit follow block order, so equal codes keep
// input order (as the device sort keeps them).
void host_target_order(const double* tgt, int64_t m, int dim, const double* box, uint32_t* order) {
  const int nthr = (int)std::max<int64_t>(1, std::min<int64_t>(
      std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())), m / 65536 + 1));
  std::vector<uint64_t> k0(m), k1(m);
  std::vector<uint32_t> v1(m);
  auto blk = [&](int t) { return std::make_pair(m * t / nthr, m * (t + 1) / nthr); };
  auto run = [&](auto&& fn) {
    std::vector<std::thread> th;
    for (int t = 1; t < nthr; t++) th.emplace_back(fn, t);
    fn(0);
    for (auto& x : th) x.join();
  };
  uint64_t kor = 0;
  std::vector<uint64_t> kor_t(nthr, 0);
  run([&](int t) {
    const auto [a, b] = blk(t);
    uint64_t o = 0;
    for (int64_t i = a; i < b; i++) {
      k0[i] = dim == 3 ? morton_code<3>(tgt + i * 3, box, 1) : morton_code<2>(tgt + i * 2, box, 1);
      order[i] = (uint32_t)i;
      o |= k0[i];
    }
    kor_t[t] = o;
  });
  for (uint64_t o : kor_t) kor |= o;
  uint64_t *ka = k0.data(), *kb = k1.data();
  uint32_t *va = order, *vb = v1.data();
  std::vector<int64_t> hist((size_t)nthr * 256);
  for (int shift = 0; shift < 64 && (kor >> shift) != 0; shift += 8) {
    run([&](int t) {
      const auto [a, b] = blk(t);
      int64_t* h = hist.data() + (size_t)t * 256;
      std::fill(h, h + 256, 0);
      for (int64_t i = a; i < b; i++) h[(ka[i] >> shift) & 255]++;
    });
    int64_t acc = 0;  // exclusive offsets: digit-major, then block order
    for (int d = 0; d < 256; d++)
      for (int t = 0; t < nthr; t++) {
        const int64_t c = hist[(size_t)t * 256 + d];
        hist[(size_t)t * 256 + d] = acc;
        acc += c;
      }
    run([&](int t) {
      const auto [a, b] = blk(t);
      int64_t* h = hist.data() + (size_t)t * 256;
      for (int64_t i = a; i < b; i++) {
        const int64_t p = h[(ka[i] >> shift) & 255]++;
        kb[p] = ka[i];
        vb[p] = va[i];
      }
    });
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (va != order) std::copy(va, va + m, order);
}

// ---------------------------------------------------------------- radix sort
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 4096 keys per tile
static_assert(RS_TILE == RADIX_TILE, "radix tile size");
constexpr int RS_WAVES = RS_THREADS / WAVE;

// Per-tile digit counts, stored digit-major: counts[d * ntiles + tile], so
// an exclusive scan along each digit row gives every tile's base offset.
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint64_t* __restrict__ keys,
                                                           int64_t n, int shift, int ntiles,
                                                           uint32_t* __restrict__ counts) {
  __shared__ uint32_t h[RS_WAVES][256];
  const int w = threadIdx.x / WAVE;
  for (int k = threadIdx.x; k < RS_WAVES * 256; k += RS_THREADS) (&h[0][0])[k] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int it = 0; it < RS_ITEMS; it++) {
    const int64_t i = base + it * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[w][(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 256) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) s += h[k][threadIdx.x];
    counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = s;
  }
}

// Exclusive scan of one 1024-thread block's values (wave shuffles + LDS).
__device__ __forceinline__ uint32_t block_exscan_1024(uint32_t v, uint32_t* lds16, uint32_t* total) {
  const int lane = __lane_id(), w = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == WAVE - 1) lds16[w] = x;
  __syncthreads();
  if (threadIdx.x < 16) {
    uint32_t s = lds16[threadIdx.x];
    uint32_t t = s;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      uint32_t y = __shfl_up(t, off);
      if ((int)threadIdx.x >= off) t += y;
    }
    lds16[16 + threadIdx.x] = t - s;  // exclusive wave offsets
    if (threadIdx.x == 15) lds16[32] = t;
  }
  __syncthreads();
  const uint32_t r = x - v + lds16[16 + w];
  *total = lds16[32];
  __syncthreads();
  return r;
}

// One block per digit: exclusive scan of counts[d, :] in place; row total
// to totals[d].
__global__ __launch_bounds__(1024) void k_radix_rowscan(uint32_t* __restrict__ counts, int ntiles,
                                                        uint32_t* __restrict__ totals) {
  __shared__ uint32_t lds[40];
  uint32_t* row = counts + (int64_t)blockIdx.x * ntiles;
  uint32_t carry = 0;
  for (int base = 0; base < ntiles; base += 1024 * 4) {
    uint32_t v[4];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = base + threadIdx.x * 4 + k;
      v[k] = i < ntiles ? row[i] : 0u;
      s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_exscan_1024(s, lds, &tot) + carry;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = base + threadIdx.x * 4 + k;
      if (i < ntiles) row[i] = ex;
      ex += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n, int shift,
    int ntiles, const uint32_t* __restrict__ counts, const uint32_t* __restrict__ totals,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t wcount[RS_WAVES][256];
  __shared__ uint32_t dbase[256];
  __shared__ uint32_t wtot[RS_WAVES];
  const int w = threadIdx.x / WAVE, lane = __lane_id();
  for (int k = threadIdx.x; k < RS_WAVES * 256; k += RS_THREADS) (&wcount[0][0])[k] = 0;
  // digit base offsets: exclusive scan of the 256 digit totals
  {
    const uint32_t v = totals[threadIdx.x];
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
      uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == WAVE - 1) wtot[w] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int k = 0; k < w; k++) add += wtot[k];
    dbase[threadIdx.x] = x - v + add;
  }
  __syncthreads();

  const int64_t base = (int64_t)blockIdx.x * RS_TILE + (int64_t)w * (WAVE * RS_ITEMS);
  uint64_t key[RS_ITEMS];
  uint32_t val[RS_ITEMS];
  uint32_t rank[RS_ITEMS];
  uint32_t dig[RS_ITEMS];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; it++) {
    const int64_t i = base + it * WAVE + lane;
    const bool valid = i < n;
    key[it] = valid ? kin[i] : 0ull;
    val[it] = valid ? vin[i] : 0u;
    const uint32_t d = (uint32_t)(key[it] >> shift) & 255u;
    dig[it] = d;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t r = 0;
    if (valid) {
      const uint32_t before = wcount[w][d];
      r = before + (uint32_t)__popcll(peers & lt);
      // the lowest lane of each digit group advances the wave's counter
      if ((peers & lt) == 0) wcount[w][d] = before + (uint32_t)__popcll(peers);
    }
    rank[it] = r;
  }
  __syncthreads();
  // per-wave bases: global digit base + this tile's offset + earlier waves
  {
    const int d = threadIdx.x;
    uint32_t run = dbase[d] + counts[(int64_t)d * ntiles + blockIdx.x];
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) {
      const uint32_t cnt = wcount[k][d];
      wcount[k][d] = run;
      run += cnt;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; it++) {
    const int64_t i = base + it * WAVE + lane;
    if (i < n) {
      const uint32_t p = wcount[w][dig[it]] + rank[it];
      kout[p] = key[it];
      vout[p] = val[it];
    }
  }
}

__global__ void k_iota(uint32_t* v, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

int iota(bh_ctx* c, uint32_t* v, int64_t n) {
  if (n == 0) return BH_OK;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, c->stream, v, n);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

// Sorts (keys, vals) by the low `bits` bits of the keys, stably.  Uses the
// alt buffers as ping-pong partners; *keys_out/*vals_out receive whichever
// buffer holds the result.
int radix_sort_ws(bh_ctx* c, uint64_t* keys, uint32_t* vals, int64_t n, int bits, uint64_t** keys_out,
                  uint32_t** vals_out, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t* counts,
                  uint32_t* totals) {
  *keys_out = keys;
  *vals_out = vals;
  if (n <= 1) return BH_OK;
  if (n > 0xffffffffll) return BH_EINVAL;
  const int ntiles = (int)radix_tiles(n);
  uint64_t *ka = keys, *kb = keys_alt;
  uint32_t *va = vals, *vb = vals_alt;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(RS_THREADS), 0, c->stream, ka, n, shift,
                       ntiles, counts);
    hipLaunchKernelGGL(k_radix_rowscan, dim3(256), dim3(1024), 0, c->stream, counts, ntiles, totals);
    hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(RS_THREADS), 0, c->stream, ka, va, n,
                       shift, ntiles, counts, totals, kb, vb);
    BH_LAUNCH_CHECK();
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  *keys_out = ka;
  *vals_out = va;
  return BH_OK;
}

int radix_sort(bh_ctx* c, uint64_t* keys, uint32_t* vals, int64_t n, int bits, uint64_t** keys_out,
               uint32_t** vals_out, uint64_t* keys_alt, uint32_t* vals_alt) {
  *keys_out = keys;
  *vals_out = vals;
  if (n <= 1) return BH_OK;
  if (n > 0xffffffffll) return BH_EINVAL;
  Scratch s(c);
  uint32_t* counts = s.get<uint32_t>((size_t)256 * radix_tiles(n));
  uint32_t* totals = s.get<uint32_t>(256);
  if (!counts || !totals) return BH_ENOMEM;
  return radix_sort_ws(c, keys, vals, n, bits, keys_out, vals_out, keys_alt, vals_alt, counts, totals);
}

// ---------------------------------------------------------------- gather
template <int DIM>
__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ pos,
                                                const double* __restrict__ mass,
                                                const uint32_t* __restrict__ perm, int64_t n,
                                                double4* spart) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = perm[i];
  double4 r;
  r.x = pos[p * DIM + 0];
  r.y = pos[p * DIM + 1];
  r.z = DIM == 3 ? pos[p * DIM + 2] : 0.0;
  r.w = mass ? mass[p] : spart[i].w;  // no masses: keep the current ones
  spart[i] = r;
}

int gather(bh_ctx* c, const double* pos, const double* mass, const uint32_t* perm, int64_t n,
           int dim, double4* spart) {
  if (n == 0) return BH_OK;
  const int nb = (int)cdiv(n, 256);
  if (dim == 3)
    hipLaunchKernelGGL(k_gather<3>, dim3(nb), dim3(256), 0, c->stream, pos, mass, perm, n, spart);
  else
    hipLaunchKernelGGL(k_gather<2>, dim3(nb), dim3(256), 0, c->stream, pos, mass, perm, n, spart);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

}  // namespace bh
