// Dense coarse-space solver: blocked right-looking Cholesky (setup), blocked forward and
// backward substitution in one workgroup (per application), explicit inverse and a dense
// symmetric matvec for large coarse spaces. Row-major storage, lower triangle.
#include "common.hpp"

namespace dds {
namespace {

constexpr int NB = 64;  // tile size

// ---------------------------------------------------------------------------------------
// Factorization
// ---------------------------------------------------------------------------------------

// Unblocked Cholesky of the diagonal tile starting at k0 (size nb <= NB), in LDS.
template <typename T>
__global__ __launch_bounds__(256) void potrf_tile_kernel(int64_t n, T* a, int64_t lda, int64_t k0,
                                                         int* info) {
  __shared__ T t[NB][NB + 1];
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    t[i][j] = (j <= i) ? a[(k0 + i) * lda + k0 + j] : T(0);
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (threadIdx.x == 0) {
      const T d = t[j][j];
      if (!(d > T(0))) {
        if (*info == 0) *info = (int)(k0 + j + 1);
        t[j][j] = T(1);  // keep going; the caller inspects *info
      } else {
        t[j][j] = sqrt(d);
      }
    }
    __syncthreads();
    const T d = t[j][j];
    for (int i = j + 1 + threadIdx.x; i < nb; i += blockDim.x) t[i][j] /= d;
    __syncthreads();
    const int m = nb - j - 1;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int i = j + 1 + e / m, l = j + 1 + e % m;
      if (l <= i) t[i][l] -= t[i][j] * t[l][j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    if (j <= i) a[(k0 + i) * lda + k0 + j] = t[i][j];
  }
}

// Panel below the diagonal tile: X = A_ik L_kk^{-T}, one workgroup of 64 lanes per tile row.
template <typename T>
__global__ __launch_bounds__(64) void trsm_panel_kernel(int64_t n, T* a, int64_t lda, int64_t k0) {
  __shared__ T L[NB][NB + 1];
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  for (int e = threadIdx.x; e < nb * nb; e += 64) {
    const int i = e / nb, j = e % nb;
    L[i][j] = (j <= i) ? a[(k0 + i) * lda + k0 + j] : T(0);
  }
  __syncthreads();
  const int64_t row = k0 + nb + (int64_t)blockIdx.x * NB + threadIdx.x;
  if (row >= n) return;
  T x[NB];
  T* ar = a + row * lda + k0;
#pragma unroll
  for (int j = 0; j < NB; ++j) x[j] = j < nb ? ar[j] : T(0);
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < nb) {
      T v = x[j];
#pragma unroll
      for (int l = 0; l < NB; ++l)
        if (l < j) v -= x[l] * L[j][l];
      x[j] = v / L[j][j];
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j)
    if (j < nb) ar[j] = x[j];
}

// Trailing update A_IJ -= L_Ik L_Jk^T for kb < J <= I; 2D grid of tile pairs, 4x4 per lane.
template <typename T>
__global__ __launch_bounds__(256) void syrk_update_kernel(int64_t n, T* a, int64_t lda,
                                                          int64_t k0) {
  const int64_t t1 = k0 + NB;  // first trailing row
  const int64_t I0 = t1 + (int64_t)blockIdx.y * NB;
  const int64_t J0 = t1 + (int64_t)blockIdx.x * NB;
  if (J0 > I0) return;
  constexpr int KC = 32;  // k-chunk: two 64 x 33 tiles fit the 64 KiB static LDS limit
  __shared__ T Li[NB][KC + 1];
  __shared__ T Lj[NB][KC + 1];
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
  T acc[4][4] = {};
  for (int kc = 0; kc < NB; kc += KC) {
    __syncthreads();
    for (int e = threadIdx.x; e < NB * KC; e += 256) {
      const int i = e / KC, l = e % KC;
      Li[i][l] = (I0 + i < n) ? a[(I0 + i) * lda + k0 + kc + l] : T(0);
      Lj[i][l] = (J0 + i < n) ? a[(J0 + i) * lda + k0 + kc + l] : T(0);
    }
    __syncthreads();
    for (int l = 0; l < KC; ++l) {
      T xi[4], xj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xi[u] = Li[ti + 16 * u][l];
        xj[u] = Lj[tj + 16 * u][l];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += xi[u] * xj[v];
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t i = I0 + ti + 16 * u, j = J0 + tj + 16 * v;
      if (i < n && j < n && j <= i) a[i * lda + j] -= acc[u][v];
    }
}

// ---------------------------------------------------------------------------------------
// Solve: forward y = L^{-1} x, backward x = L^{-T} y. One workgroup of 1024 lanes per
// right-hand side (blockIdx.x); x and y with column stride n. `identity` solves e_b.
// The solution lives in LDS (at most kMaxSolveN entries).
// ---------------------------------------------------------------------------------------
constexpr int kSolveThreads = 1024;
constexpr int kSolveWaves = kSolveThreads / 64;
constexpr int kMaxSolveN = 6144;  // keeps the LDS stage under 64 KiB

template <typename T>
__global__ __launch_bounds__(kSolveThreads) void chol_solve_kernel(int64_t n, const T* __restrict__ L,
                                                                   int64_t ldl, const T* __restrict__ x,
                                                                   T* __restrict__ y, int identity,
                                                                   const dds_pcg_state* __restrict__ gate) {
  if (gate != nullptr && gate->done) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sol = reinterpret_cast<T*>(smem_raw);  // n
  T* part = sol + n;                         // NB
  T* red = part + NB;                        // kSolveWaves * NB
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t rhs = blockIdx.x;
  const T* xb = x + rhs * n;
  T* yb = y + rhs * n;
  const int nt = (int)ceil_div(n, NB);

  // forward substitution
  for (int kb = 0; kb < nt; ++kb) {
    const int64_t k0 = (int64_t)kb * NB;
    const int nb = (int)std::min<int64_t>(NB, n - k0);
    // part[i] = x[k0+i] - sum_{j<k0} L[k0+i][j] sol[j]; 16 waves x 4 rows each
    for (int i = w; i < nb; i += kSolveWaves) {
      const T* Lr = L + (k0 + i) * ldl;
      T acc = 0;
      for (int64_t j = lane; j < k0; j += 64) acc += Lr[j] * sol[j];
      acc = wave_sum(acc);
      if (lane == 0) {
        const T b = identity ? T(k0 + i == rhs ? 1 : 0) : xb[k0 + i];
        part[i] = b - acc;
      }
    }
    __syncthreads();
    if (w == 0) {
      T v = lane < nb ? part[lane] : T(0);
      const T* Ld = L + (k0 + lane) * ldl + k0;
      for (int j = 0; j < nb; ++j) {
        const T yj = __shfl(v, j, 64) / L[(k0 + j) * ldl + k0 + j];
        if (lane == j) v = yj;
        if (lane > j && lane < nb) v -= Ld[j] * yj;
      }
      if (lane < nb) sol[k0 + lane] = v;
    }
    __syncthreads();
  }

  // backward substitution
  for (int kb = nt - 1; kb >= 0; --kb) {
    const int64_t k0 = (int64_t)kb * NB;
    const int nb = (int)std::min<int64_t>(NB, n - k0);
    const int64_t k1 = k0 + nb;
    // red[w][i] = sum over rows j >= k1, j = k1 + w (mod 16), of L[j][k0+i] * sol[j]
    T acc = 0;
    if (lane < nb)
      for (int64_t j = k1 + w; j < n; j += kSolveWaves) acc += L[j * ldl + k0 + lane] * sol[j];
    red[w * NB + lane] = acc;
    __syncthreads();
    if (w == 0) {
      T s = 0;
      for (int ww = 0; ww < kSolveWaves; ++ww) s += red[ww * NB + lane];
      T v = lane < nb ? sol[k0 + lane] - s : T(0);
      for (int j = nb - 1; j >= 0; --j) {
        const T xj = __shfl(v, j, 64) / L[(k0 + j) * ldl + k0 + j];
        if (lane == j) v = xj;
        if (lane < j) v -= L[(k0 + j) * ldl + k0 + lane] * xj;
      }
      if (lane < nb) sol[k0 + lane] = v;
    }
    __syncthreads();
  }
  for (int64_t i = threadIdx.x; i < n; i += kSolveThreads) yb[i] = sol[i];
}

// y = S x, S dense symmetric row-major: one wavefront per row. A row that starts off a
// 16-byte boundary (an odd leading dimension puts every other row there) first takes
// the h < V elements up to the boundary one per lane; the rest of the row is read in whole
// 16-byte chunks, eight per lane at a time (chunks c, c + 64, ..., c + 448 in flight
// together, those past the row end skipped), with x read in 16-byte chunks when it shares
// that alignment and element by element (cached) when it does not; the last elements
// follow one per lane. Lanes' sums are added by a butterfly. A row of a few thousand
// entries is one round trip to memory per lane.
template <typename T>
__global__ __launch_bounds__(256) void dense_symv_kernel(int64_t n, const T* __restrict__ S,
                                                         int64_t lds, const T* __restrict__ x,
                                                         T* __restrict__ y,
                                                         const dds_pcg_state* __restrict__ gate) {
  if (gate != nullptr && gate->done) return;
  constexpr int V = Vec16<T>::N;
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const T* Sr = S + row * lds;
  T a0 = 0, a1 = 0;
  const uintptr_t mis = reinterpret_cast<uintptr_t>(Sr) & 15;
  if (mis % sizeof(T) == 0) {
    // elements before the first 16-byte boundary of the row
    const int64_t h = mis == 0 ? 0 : std::min<int64_t>(n, (int64_t)((16 - mis) / sizeof(T)));
    if (lane < h) a0 += Sr[lane] * x[lane];
    const T* Sa = Sr + h;
    const T* xa = x + h;
    const bool xal = (reinterpret_cast<uintptr_t>(xa) & 15) == 0;
    const int64_t nchunk = (n - h) / V;
    for (int64_t c0 = lane; c0 < nchunk; c0 += 8 * 64) {
      Vec16<T> sv[8], xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t c = c0 + (int64_t)u * 64;
        if (c < nchunk) {
          sv[u] = load16_nt(Sa + c * V);
          if (xal) {
            xv[u] = load16(xa + c * V);
          } else {
#pragma unroll
            for (int e = 0; e < V; ++e) xv[u].v[e] = xa[c * V + e];
          }
        } else {
#pragma unroll
          for (int e = 0; e < V; ++e) sv[u].v[e] = xv[u].v[e] = T{};
        }
      }
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
#pragma unroll
        for (int e = 0; e < V; ++e) {
          a0 += sv[u].v[e] * xv[u].v[e];
          a1 += sv[u + 1].v[e] * xv[u + 1].v[e];
        }
      }
    }
    const int64_t t = h + nchunk * V + lane;
    if (t < n) a0 += Sr[t] * x[t];
  } else {
    for (int64_t j = lane; j < n; j += 64) a0 += Sr[j] * x[j];
  }
  const T acc = wave_sum(a0 + a1);
  if (lane == 0) y[row] = acc;
}

template <typename T>
int factor(int64_t n, void* a_, int64_t lda, int* info, hipStream_t s) {
  T* a = static_cast<T*>(a_);
  DDS_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  const int64_t nt = ceil_div(n, NB);
  for (int64_t kb = 0; kb < nt; ++kb) {
    const int64_t k0 = kb * NB;
    hipLaunchKernelGGL(potrf_tile_kernel<T>, dim3(1), dim3(256), 0, s, n, a, lda, k0, info);
    DDS_LAUNCHED();
    const int64_t m = nt - kb - 1;
    if (m == 0) break;
    hipLaunchKernelGGL(trsm_panel_kernel<T>, dim3((unsigned)m), dim3(64), 0, s, n, a, lda, k0);
    DDS_LAUNCHED();
    hipLaunchKernelGGL(syrk_update_kernel<T>, dim3((unsigned)m, (unsigned)m), dim3(256), 0, s, n,
                       a, lda, k0);
    DDS_LAUNCHED();
  }
  return DDS_OK;
}

template <typename T>
size_t solve_lds(int64_t n) {
  return (size_t)(n + NB + kSolveWaves * NB) * sizeof(T);
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_chol_factor(int dtype, int64_t n, void* a, int64_t lda, int* info,
                               void* stream) {
  DDS_REQUIRE(n >= 0 && lda >= n && info, "dds_chol_factor: bad arguments");
  if (n == 0) return DDS_OK;
  DDS_REQUIRE(a, "dds_chol_factor: null matrix");
  DDS_REQUIRE(ceil_div(n, NB) <= 65535, "dds_chol_factor: n too large");
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64) return factor<double>(n, a, lda, info, s);
  if (dtype == DDS_F32) return factor<float>(n, a, lda, info, s);
  set_error("dds_chol_factor: unsupported dtype %d", dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_chol_solve(int dtype, int64_t n, const void* L, int64_t ldl, const void* x,
                              void* y, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n >= 0 && ldl >= n, "dds_chol_solve: bad sizes");
  if (n == 0) return DDS_OK;
  DDS_REQUIRE(L && x && y, "dds_chol_solve: null pointer");
  DDS_REQUIRE(n <= kMaxSolveN, "dds_chol_solve: n=%lld exceeds %d (use the inverse form)",
              (long long)n, kMaxSolveN);
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(chol_solve_kernel<double>, dim3(1), dim3(kSolveThreads),
                       solve_lds<double>(n), s, n, static_cast<const double*>(L), ldl,
                       static_cast<const double*>(x), static_cast<double*>(y), 0, gate);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(chol_solve_kernel<float>, dim3(1), dim3(kSolveThreads),
                       solve_lds<float>(n), s, n, static_cast<const float*>(L), ldl,
                       static_cast<const float*>(x), static_cast<float*>(y), 0, gate);
  } else {
    set_error("dds_chol_solve: unsupported dtype %d", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_chol_inverse(int dtype, int64_t n, const void* L, int64_t ldl, void* ainv,
                                int64_t ldi, void* stream) {
  DDS_REQUIRE(n >= 0 && ldl >= n && ldi == n, "dds_chol_inverse: bad sizes (ldi must be n)");
  if (n == 0) return DDS_OK;
  DDS_REQUIRE(L && ainv, "dds_chol_inverse: null pointer");
  DDS_REQUIRE(n <= kMaxSolveN, "dds_chol_inverse: n=%lld exceeds %d", (long long)n, kMaxSolveN);
  hipStream_t s = as_stream(stream);
  // Column b of the inverse is the solve with e_b; symmetric, so it is also row b.
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(chol_solve_kernel<double>, dim3((unsigned)n), dim3(kSolveThreads),
                       solve_lds<double>(n), s, n, static_cast<const double*>(L), ldl,
                       static_cast<const double*>(ainv), static_cast<double*>(ainv), 1,
                       (const dds_pcg_state*)nullptr);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(chol_solve_kernel<float>, dim3((unsigned)n), dim3(kSolveThreads),
                       solve_lds<float>(n), s, n, static_cast<const float*>(L), ldl,
                       static_cast<const float*>(ainv), static_cast<float*>(ainv), 1,
                       (const dds_pcg_state*)nullptr);
  } else {
    set_error("dds_chol_inverse: unsupported dtype %d", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_dense_symv(int dtype, int64_t n, const void* S, int64_t lds, const void* x,
                              void* y, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n >= 0 && lds >= n, "dds_dense_symv: bad sizes");
  if (n == 0) return DDS_OK;
  DDS_REQUIRE(S && x && y, "dds_dense_symv: null pointer");
  hipStream_t s = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n, 4));
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(dense_symv_kernel<double>, grid, dim3(256), 0, s, n,
                       static_cast<const double*>(S), lds, static_cast<const double*>(x),
                       static_cast<double*>(y), gate);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(dense_symv_kernel<float>, grid, dim3(256), 0, s, n,
                       static_cast<const float*>(S), lds, static_cast<const float*>(x),
                       static_cast<float*>(y), gate);
  } else {
    set_error("dds_dense_symv: unsupported dtype %d", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}
