// fp64_check.hip -- checks the middle-range FP64 square root and division of
// bh_fp64.h against sqrt() and '/' on this device, bit for bit: random
// operands spread over the whole middle range (uniform exponent, uniform
// 52-bit mantissa), exact squares, and every exponent of the range with the
// mantissas 0, 1, 2^51 and 2^52 - 1.  Prints one JSON line with the number
// of operands and mismatches per function; exits 1 on any mismatch.
//
//   fp64_check [random operands per function, default 2^28]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bh_fp64.h"

#define CHECK(x)                                                                      \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return 2;                                                                       \
    }                                                                                 \
  } while (0)

enum { F_SQRT = 0, F_RECIP = 1, F_DIV = 2, F_RSQRT = 3, NF = 4 };

__host__ __device__ inline uint64_t mix(uint64_t z) {  // splitmix64 finaliser
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// a double with unbiased exponent uniform in [lo, hi] and a uniform mantissa
__host__ __device__ inline double draw(uint64_t h, int lo, int hi) {
  const uint64_t e = (uint64_t)(lo + (int)((h >> 52) % (uint64_t)(hi - lo + 1)) + 1023);
  const uint64_t bits = (e << 52) | (mix(h) & 0xfffffffffffffull);
  double x;
  std::memcpy(&x, &bits, 8);
  return x;
}

__device__ inline uint64_t bits_of(double x) { return (uint64_t)__double_as_longlong(x); }

// One comparison per function for operand pair (a, b); counts per function
// in cnt[f], the first mismatch's operands in bad[2f], bad[2f+1].
__device__ inline void compare(double a, double b, unsigned long long* cnt, double* bad) {
  const bool ms = bits_of(bh::sqrt_mid(a)) != bits_of(sqrt(a));
  const bool ain = a >= 0x1p-350 && a <= 0x1p350;  // a in the denominator range
  const bool mr = bits_of(bh::recip_mid(b)) != bits_of(1.0 / b) ||
                  (ain && bits_of(bh::recip_mid(a)) != bits_of(1.0 / a));
  const bool md = ain && bits_of(bh::div_mid(b, a)) != bits_of(b / a);
  const bool mq = bits_of(bh::recip_mid(bh::sqrt_mid(a))) != bits_of(1.0 / sqrt(a));
  const bool m[NF] = {ms, mr, md, mq};
  for (int f = 0; f < NF; f++)
    if (m[f]) {
      atomicAdd(cnt + f, 1ull);
      bad[2 * f] = a;
      bad[2 * f + 1] = b;
    }
}

// random operands: a (sqrt, div denominator) from the square root's range
// restricted to the division's where needed, b from [2^-300, 2^300]
__global__ void k_random(uint64_t seed, int64_t n, unsigned long long* cnt, double* bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h = mix(seed ^ mix((uint64_t)i));
    const uint64_t h2 = mix(h ^ 0x5bd1e995ull);
    const bool narrow = (h & 1) != 0;  // half the operands in the division's denominator range
    const double a = draw(h, narrow ? -350 : -700, narrow ? 350 : 700);
    const double b = draw(h2, -300, 300);
    compare(a, b, cnt, bad);
  }
}

// exact squares k^2 2^(2e) (k < 2^26) and their neighbours one ulp away
__global__ void k_squares(uint64_t seed, int64_t n, unsigned long long* cnt, double* bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h = mix(seed ^ mix((uint64_t)i * 3 + 1));
    const double k = (double)((h & ((1ull << 26) - 1)) | 1ull);
    const int e = (int)((h >> 32) % 600) - 300;
    double x = ldexp(k * k, 2 * e);
    const int nb = (int)((h >> 60) % 3) - 1;
    x = __longlong_as_double(__double_as_longlong(x) + nb);
    const double b = draw(mix(h), -300, 300);
    compare(x, b, cnt, bad);
  }
}

__global__ void k_list(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                       unsigned long long* cnt, double* bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  compare(a[i], b[i], cnt, bad);
}

int main(int argc, char** argv) {
  const int64_t n = argc > 1 ? std::atoll(argv[1]) : (1ll << 28);
  // the edge list: every exponent of the square root's range with four mantissas,
  // paired with every exponent of the numerator range
  std::vector<double> la, lb;
  const uint64_t mants[4] = {0ull, 1ull, 1ull << 51, (1ull << 52) - 1};
  for (int e = -700; e <= 700; e++)
    for (int k = 0; k < 4; k++)
      for (int k2 = 0; k2 < 4; k2++) {
        const uint64_t ba = ((uint64_t)(e + 1023) << 52) | mants[k];
        const int eb = -300 + (int)(mix((uint64_t)(e * 16 + k * 4 + k2)) % 601);
        const uint64_t bb = ((uint64_t)(eb + 1023) << 52) | mants[k2];
        double x, y;
        std::memcpy(&x, &ba, 8);
        std::memcpy(&y, &bb, 8);
        la.push_back(x);
        lb.push_back(y);
      }
  const int64_t nl = (int64_t)la.size();
  while (la.size() % 64) {
    la.push_back(1.0);
    lb.push_back(1.0);
  }
  double *d_a, *d_b, *d_bad;
  unsigned long long* d_cnt;
  CHECK(hipMalloc(&d_a, la.size() * 8));
  CHECK(hipMalloc(&d_b, lb.size() * 8));
  CHECK(hipMalloc(&d_cnt, NF * 8));
  CHECK(hipMalloc(&d_bad, 2 * NF * 8));
  CHECK(hipMemcpy(d_a, la.data(), la.size() * 8, hipMemcpyHostToDevice));
  CHECK(hipMemcpy(d_b, lb.data(), lb.size() * 8, hipMemcpyHostToDevice));
  CHECK(hipMemset(d_cnt, 0, NF * 8));
  CHECK(hipMemset(d_bad, 0, 2 * NF * 8));
  hipLaunchKernelGGL(k_list, dim3((unsigned)(la.size() / 64)), dim3(64), 0, 0, d_a, d_b, (int64_t)la.size(), d_cnt,
                     d_bad);
  CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_random, dim3(8192), dim3(256), 0, 0, 0x243f6a8885a308d3ull, n, d_cnt, d_bad);
  CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_squares, dim3(8192), dim3(256), 0, 0, 0x13198a2e03707344ull, n / 4, d_cnt, d_bad);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  unsigned long long cnt[NF];
  double bad[2 * NF];
  CHECK(hipMemcpy(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(bad, d_bad, sizeof bad, hipMemcpyDeviceToHost));
  const char* names[NF] = {"sqrt", "recip", "div", "recip_sqrt"};
  std::printf("{\"operands\": %lld", (long long)(n + n / 4 + nl));
  unsigned long long tot = 0;
  for (int f = 0; f < NF; f++) {
    std::printf(", \"%s\": {\"mismatch\": %llu, \"first\": [%.17g, %.17g]}", names[f], cnt[f], bad[2 * f],
                bad[2 * f + 1]);
    tot += cnt[f];
  }
  std::printf("}\n");
  return tot ? 1 : 0;
}
