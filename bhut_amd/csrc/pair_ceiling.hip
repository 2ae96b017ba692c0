// pair_ceiling.hip -- the rate at which one MI355X evaluates softened FP32
// particle pairs in the leaf-pair loop of k_group_eval alone: sources
// broadcast from an LDS tile (x, y, z read as 12 B), 8 pairs unrolled, the
// same launch shape (4 wavefronts per workgroup, 8 per SIMD), no lists,
// windows or partial sums.  It bounds what the evaluation kernel can reach
// with this instruction mix: 3 sub, 3 fma (r^2 + eps^2), rsq, 2 mul (r^-3),
// 3 fma per pair, plus 1 mul when each source carries its own mass.
//
// Output: one JSON line {"uniform": pairs/s, "mass": pairs/s,
// "no_rsq": pairs/s, "fma_lane_per_s": ..., ...}; "no_rsq" replaces the
// reciprocal square root by a multiply, to show its share of the loop;
// "fma_lane_per_s" is the issue rate of independent FP32 FMAs alone (eight
// chains per lane, same launch shape): the VALU ceiling on this device
// without packed instructions, in lane-FMAs per second; "rsq_lane_per_s" the
// same for v_rsq_f32.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "%s: %s (line %d)\n", #x, hipGetErrorString(e_), __LINE__); \
      return 1;                                                                \
    }                                                                          \
  } while (0)

constexpr int TILE_SRC = 4096;  // distinct sources the tiles are filled from

// MODE 0: common mass (no per-pair mass product), 1: per-source mass,
// 2: as 0 with rsq replaced by a multiply.
template <int MODE>
__global__ __launch_bounds__(256, 8) void k_pair_loop(const float4* __restrict__ src, int reps, float eps2,
                                                      float* __restrict__ out) {
  __shared__ float4 tile[4][64];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  tile[w][lane] = src[(gid / 64 * 64 + lane) % TILE_SRC];
  __syncthreads();
  const float tx = 0.001f * lane, ty = 0.002f * lane, tz = -0.001f * lane;
  float px = 0.f, py = 0.f, pz = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
  for (int r = 0; r < reps; r++) {
#pragma unroll 8
    for (int k = 0; k < 64; k++) {
      float4 v;
      if (MODE == 1) {
        v = tile[w][k];
      } else {
        const float3 v3 = *reinterpret_cast<const float3*>(&tile[w][k]);
        v = make_float4(v3.x, v3.y, v3.z, 1.f);
      }
      const float dx = v.x - tx, dy = v.y - ty, dz = v.z - tz;
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
      const float ri = MODE == 2 ? r2 * 0.5f : __builtin_amdgcn_rsqf(r2);
      const float mg = MODE == 1 ? v.w * (ri * ri * ri) : ri * ri * ri;
      px = fmaf(mg, dx, px);
      py = fmaf(mg, dy, py);
      pz = fmaf(mg, dz, pz);
    }
    sx += px;
    sy += py;
    sz += pz;
    px = py = pz = 0.f;
  }
  out[gid] = sx + sy + sz;  // keeps the loop live
}

// Eight independent FMA chains per lane: VALU issue rate alone.
__global__ __launch_bounds__(256, 8) void k_fma_chains(int reps, float* __restrict__ out) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  float a[8];
#pragma unroll
  for (int j = 0; j < 8; j++) a[j] = 0.001f * (threadIdx.x + j);
  for (int r = 0; r < reps; r++) {
#pragma unroll 16
    for (int k = 0; k < 64; k++) {
#pragma unroll
      for (int j = 0; j < 8; j++) a[j] = fmaf(a[j], 0.999f, 0.001f);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) s += a[j];
  out[gid] = s;  // keeps the chains live
}

// Eight independent v_rsq_f32 chains per lane: the transcendental's issue
// rate alone.
__global__ __launch_bounds__(256, 8) void k_rsq_chains(int reps, float* __restrict__ out) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  float a[8];
#pragma unroll
  for (int j = 0; j < 8; j++) a[j] = 1.f + 0.001f * (threadIdx.x + j);
  for (int r = 0; r < reps; r++) {
#pragma unroll 16
    for (int k = 0; k < 64; k++) {
#pragma unroll
      for (int j = 0; j < 8; j++) a[j] = __builtin_amdgcn_rsqf(a[j]);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) s += a[j];
  out[gid] = s;  // keeps the chains live
}

int main() {
  int ncu = 0;
  CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
  std::vector<float4> h(TILE_SRC);
  for (int i = 0; i < TILE_SRC; i++)
    h[i] = make_float4(0.01f * (i % 97), 0.013f * (i % 89), -0.007f * (i % 83), 1.f + 0.001f * (i % 7));
  const int blocks = ncu * 8;  // 8 workgroups of 4 wavefronts per CU
  float4* d = nullptr;
  float* o = nullptr;
  CK(hipMalloc(&d, TILE_SRC * sizeof(float4)));
  CK(hipMalloc(&o, (size_t)blocks * 256 * sizeof(float)));
  CK(hipMemcpy(d, h.data(), TILE_SRC * sizeof(float4), hipMemcpyHostToDevice));
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  const int reps = 2000;
  const double pairs = (double)blocks * 256 * reps * 64;
  double rate[3] = {0, 0, 0};
  for (int mode = 0; mode < 3; mode++) {
    for (int it = 0; it < 3; it++) {  // the last of three launches is kept
      CK(hipEventRecord(a));
      if (mode == 0) hipLaunchKernelGGL(k_pair_loop<0>, dim3(blocks), dim3(256), 0, 0, d, reps, 1e-4f, o);
      else if (mode == 1) hipLaunchKernelGGL(k_pair_loop<1>, dim3(blocks), dim3(256), 0, 0, d, reps, 1e-4f, o);
      else hipLaunchKernelGGL(k_pair_loop<2>, dim3(blocks), dim3(256), 0, 0, d, reps, 1e-4f, o);
      CK(hipGetLastError());
      CK(hipEventRecord(b));
      CK(hipEventSynchronize(b));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, a, b));
      rate[mode] = pairs / (ms * 1e-3);
    }
  }
  double fma_rate = 0.0;
  const int freps = 1000;
  for (int it = 0; it < 3; it++) {
    CK(hipEventRecord(a));
    hipLaunchKernelGGL(k_fma_chains, dim3(blocks), dim3(256), 0, 0, freps, o);
    CK(hipGetLastError());
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, a, b));
    fma_rate = (double)blocks * 256 * freps * 64 * 8 / (ms * 1e-3);
  }
  double rsq_rate = 0.0;
  for (int it = 0; it < 3; it++) {
    CK(hipEventRecord(a));
    hipLaunchKernelGGL(k_rsq_chains, dim3(blocks), dim3(256), 0, 0, freps, o);
    CK(hipGetLastError());
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, a, b));
    rsq_rate = (double)blocks * 256 * freps * 64 * 8 / (ms * 1e-3);
  }
  printf("{\"uniform\": %.6e, \"mass\": %.6e, \"no_rsq\": %.6e, \"fma_lane_per_s\": %.6e, "
         "\"rsq_lane_per_s\": %.6e, \"cus\": %d, \"pairs_per_launch\": %.6e}\n", rate[0], rate[1], rate[2],
         fma_rate, rsq_rate, ncu, pairs);
  CK(hipFree(d));
  CK(hipFree(o));
  return 0;
}
