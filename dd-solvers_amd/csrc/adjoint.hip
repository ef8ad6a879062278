// Per-element coefficient gradient of the Q1 Poisson solve (adjoint method):
//   grad[e] = -h * exp(theta[e]) * lam_e^T Khat u_e
// for the N^3 hexahedra of the unit cube, h = 1/N, Khat the unit-coefficient Q1 element
// stiffness (entry (a, b) depends only on how many coordinates of the vertices a and b
// differ: 1/3, 0, -1/12, -1/12), lam and u the adjoint and forward solutions on the
// (N-1)^3 interior nodes (Dirichlet nodes contribute 0).
#include "common.hpp"

namespace dds {
namespace {

__global__ __launch_bounds__(256) void q1_element_grad_kernel(int N, const double* __restrict__ theta,
                                                              const double* __restrict__ lam,
                                                              const double* __restrict__ u,
                                                              const int32_t* __restrict__ map,
                                                              double* __restrict__ grad) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n_el = (int64_t)N * N * N;
  if (e >= n_el) return;
  const int m = N - 1;
  const int ex = (int)(e % N);
  const int ey = (int)((e / N) % N);
  const int ez = (int)(e / ((int64_t)N * N));
  double l[8], w[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ix = ex + (a & 1), iy = ey + ((a >> 1) & 1), iz = ez + ((a >> 2) & 1);
    const bool inside = ix >= 1 && ix <= m && iy >= 1 && iy <= m && iz >= 1 && iz <= m;
    double lv = 0.0, uv = 0.0;
    if (inside) {
      int64_t idx = ((int64_t)(iz - 1) * m + (iy - 1)) * m + (ix - 1);
      if (map != nullptr) idx = map[idx];
      lv = lam[idx];
      uv = u[idx];
    }
    l[a] = lv;
    w[a] = uv;
  }
  constexpr double kd[4] = {1.0 / 3.0, 0.0, -1.0 / 12.0, -1.0 / 12.0};
  double q = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b) t += kd[__builtin_popcount(a ^ b)] * w[b];
    q += l[a] * t;
  }
  grad[e] = -exp(theta[e]) * q / (double)N;
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_q1_element_grad(int64_t N, const double* theta, const double* lam,
                                   const double* u, const int32_t* map, double* grad,
                                   void* stream) {
  DDS_REQUIRE(N >= 2 && N <= 2048, "dds_q1_element_grad: bad grid size %lld", (long long)N);
  DDS_REQUIRE(theta && lam && u && grad, "dds_q1_element_grad: null pointer");
  const int64_t n_el = N * N * N;
  hipLaunchKernelGGL(q1_element_grad_kernel, dim3((unsigned)ceil_div(n_el, 256)), dim3(256), 0,
                     as_stream(stream), (int)N, theta, lam, u, map, grad);
  DDS_LAUNCHED();
  return DDS_OK;
}
