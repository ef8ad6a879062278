// Streaming-bandwidth kernels: the achievable HBM rate that the memory-bound kernels of
// the solve are compared with (besides the 8 TB/s datasheet figure). Each thread moves
// 4 x 16 bytes per step with non-temporal loads; the grid is a multiple of the CU count
// and strides over the arrays.
//   read:  partials[b] = sum of x over block b's elements (8 bytes per element moved)
//   copy:  y = x                                           (16 bytes per element)
//   triad: y = a + s * b                                   (24 bytes per element)
#include "common.hpp"

namespace dds {
namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;

typedef double double2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2_t ld_nt2(const double* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const double2_t*>(p));
}

__global__ __launch_bounds__(kThreads) void stream_read_kernel(int64_t n2, const double* __restrict__ x,
                                                               double* __restrict__ partials) {
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kThreads * kUnroll;
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; base < n2; base += stride) {
    double2_t v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kThreads;
      v[u] = i < n2 ? ld_nt2(x + 2 * i) : double2_t{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc += v[u][0] + v[u][1];
  }
  __shared__ double scratch[kThreads / 64];
  acc = block_sum(acc, scratch);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kThreads) void stream_copy_kernel(int64_t n2, const double* __restrict__ x,
                                                               double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * kThreads * kUnroll;
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; base < n2; base += stride) {
    double2_t v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kThreads;
      if (i < n2) v[u] = ld_nt2(x + 2 * i);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kThreads;
      if (i < n2) __builtin_nontemporal_store(v[u], reinterpret_cast<double2_t*>(y + 2 * i));
    }
  }
}

__global__ __launch_bounds__(kThreads) void stream_triad_kernel(int64_t n2, const double* __restrict__ a,
                                                                const double* __restrict__ b, double s,
                                                                double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * kThreads * kUnroll;
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; base < n2; base += stride) {
    double2_t va[kUnroll], vb[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kThreads;
      if (i < n2) {
        va[u] = ld_nt2(a + 2 * i);
        vb[u] = ld_nt2(b + 2 * i);
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kThreads;
      if (i < n2) __builtin_nontemporal_store(va[u] + s * vb[u], reinterpret_cast<double2_t*>(y + 2 * i));
    }
  }
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_stream(int op, int64_t n, const double* a, const double* b, double s, double* y,
                          int64_t grid, void* stream) {
  DDS_REQUIRE(n >= 0 && n % 2 == 0, "dds_stream: n must be even and >= 0 (got %lld)", (long long)n);
  DDS_REQUIRE(grid >= 1 && grid <= kMaxPartials, "dds_stream: grid must be in [1, %d]", kMaxPartials);
  DDS_REQUIRE(a && y, "dds_stream: null pointer");
  const int64_t n2 = n / 2;
  hipStream_t st = as_stream(stream);
  if (op == 0) {
    hipLaunchKernelGGL(stream_read_kernel, dim3((unsigned)grid), dim3(kThreads), 0, st, n2, a, y);
  } else if (op == 1) {
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)grid), dim3(kThreads), 0, st, n2, a, y);
  } else if (op == 2) {
    DDS_REQUIRE(b != nullptr, "dds_stream: triad needs b");
    hipLaunchKernelGGL(stream_triad_kernel, dim3((unsigned)grid), dim3(kThreads), 0, st, n2, a, b, s, y);
  } else {
    set_error("dds_stream: op must be 0 (read), 1 (copy) or 2 (triad), got %d", op);
    return DDS_EINVAL;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}
