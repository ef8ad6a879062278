// Batched inverse of symmetric positive definite matrices (the local subdomain matrices
// at preconditioner setup) by blocked Gauss-Jordan elimination without pivoting, in place.
//
// Matrix b is row-major k_b x k_b at A + b * stride with leading dimension ld. For each
// pivot block p (NB rows/columns at p0) with P = A_pp^-1:
//   A_pj <- P A_pj            (j outside p: the row panel)
//   A_ij <- A_ij - A_ip A_pj' (i, j outside p, A_pj' the new row panel)
//   A_ip <- -A_ip P           (i outside p)
//   A_pp <- P
// after the last pivot block A holds A^-1. On an SPD matrix every pivot is a Schur
// complement diagonal and positive; a pivot <= 0 (the matrix is not SPD) sets info[b] = 1
// and the caller falls back to a pivoting inverse for that batch.
//
// One workgroup of 256 threads per matrix; the trailing update runs over 64 x 64 tiles,
// each thread owning a 4 x 4 micro-tile (rows 4 * (t / 16) + r, columns t % 16 + 16 c),
// with the tile's pivot columns A_Ip and the row panel staged in LDS.
#include "common.hpp"

namespace dds {
namespace {

constexpr int kNB = 32;
constexpr int kTile = 64;
constexpr int kThreads = 256;

template <typename T>
struct GjLds {
  T P[kNB][kNB + 1];
  T Ti[kTile][kNB + 1];
  T Rt[kNB][kTile + 1];
};

// Eliminates the first kp pivots (kp <= k) of the k x k matrix at a (row-major, leading
// dimension ld) in place, by the steps above applied to every row and column of the
// matrix: with kp = k the result is A^-1; with kp < k and A = [A_JJ A_JS; A_SJ A_SS]
// (J the first kp), it is [A_JJ^-1, A_JJ^-1 A_JS; -A_SJ A_JJ^-1, A_SS - A_SJ A_JJ^-1 A_JS].
// Sets bad = 1 on a pivot <= 0. Called by all kThreads threads of the workgroup; the
// workgroup's global writes are visible to all its threads on return.
template <typename T>
__device__ void gj_sweep(T* __restrict__ a, int k, int kp, int64_t ld, GjLds<T>& L, int& bad) {
  auto& P = L.P;
  auto& Ti = L.Ti;
  auto& Rt = L.Rt;
  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  for (int p0 = 0; p0 < kp; p0 += kNB) {
    const int nb = min(kNB, kp - p0);
    // 1. pivot block into LDS (identity outside nb x nb) and its inverse by scalar sweeps
    for (int e = tid; e < kNB * kNB; e += kThreads) {
      const int r = e / kNB, c = e % kNB;
      P[r][c] = (r < nb && c < nb) ? a[(int64_t)(p0 + r) * ld + p0 + c] : (r == c ? T(1) : T(0));
    }
    __syncthreads();
    for (int q = 0; q < nb; ++q) {
      const T piv = P[q][q];
      if (tid == 0 && !(piv > T(0))) bad = 1;
      const T ipiv = T(1) / piv;
      T nv[(kNB * kNB) / kThreads];
#pragma unroll
      for (int u = 0; u < (kNB * kNB) / kThreads; ++u) {
        const int e = tid + u * kThreads;
        const int i = e / kNB, j = e % kNB;
        T v;
        if (i == q && j == q) v = ipiv;
        else if (i == q) v = P[q][j] * ipiv;
        else if (j == q) v = -P[i][q] * ipiv;
        else v = P[i][j] - P[i][q] * (P[q][j] * ipiv);
        nv[u] = v;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < (kNB * kNB) / kThreads; ++u) {
        const int e = tid + u * kThreads;
        P[e / kNB][e % kNB] = nv[u];
      }
      __syncthreads();
    }
    // 2. row panel: A_pj <- P A_pj for the columns outside the pivot block
    for (int j0 = 0; j0 < k; j0 += kTile) {
      for (int e = tid; e < kNB * kTile; e += kThreads) {
        const int q = e / kTile, c = e % kTile;
        const int j = j0 + c;
        Rt[q][c] = (q < nb && j < k) ? a[(int64_t)(p0 + q) * ld + j] : T(0);
      }
      __syncthreads();
      for (int e = tid; e < kNB * kTile; e += kThreads) {
        const int r = e / kTile, c = e % kTile;
        const int j = j0 + c;
        if (r < nb && j < k && (j < p0 || j >= p0 + nb)) {
          T s = T(0);
          for (int q = 0; q < nb; ++q) s += P[r][q] * Rt[q][c];
          a[(int64_t)(p0 + r) * ld + j] = s;
        }
      }
      __syncthreads();
    }
    // 3. rows outside the pivot block, tile by tile
    for (int i0 = 0; i0 < k; i0 += kTile) {
      for (int e = tid; e < kTile * kNB; e += kThreads) {
        const int r = e / kNB, q = e % kNB;
        const int i = i0 + r;
        Ti[r][q] = (q < nb && i < k && (i < p0 || i >= p0 + nb)) ? a[(int64_t)i * ld + p0 + q] : T(0);
      }
      __syncthreads();
      for (int j0 = 0; j0 < k; j0 += kTile) {
        // Rt: the new row panel, with P in the pivot columns
        for (int e = tid; e < kNB * kTile; e += kThreads) {
          const int q = e / kTile, c = e % kTile;
          const int j = j0 + c;
          T v = T(0);
          if (q < nb && j < k) {
            if (j >= p0 && j < p0 + nb) v = P[q][j - p0];
            else v = a[(int64_t)(p0 + q) * ld + j];
          }
          Rt[q][c] = v;
        }
        __syncthreads();
        T acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] = T(0);
        for (int q = 0; q < nb; ++q) {
          T ti[4], rt[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) ti[r] = Ti[4 * ty + r][q];
#pragma unroll
          for (int c = 0; c < 4; ++c) rt[c] = Rt[q][tx + 16 * c];
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] += ti[r] * rt[c];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = i0 + 4 * ty + r;
          if (i >= k || (i >= p0 && i < p0 + nb)) continue;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = j0 + tx + 16 * c;
            if (j >= k) continue;
            T* dst = a + (int64_t)i * ld + j;
            const bool pivot_col = j >= p0 && j < p0 + nb;
            *dst = (pivot_col ? T(0) : *dst) - acc[r][c];
          }
        }
        __syncthreads();
      }
    }
    // 4. the pivot block itself
    for (int e = tid; e < kNB * kNB; e += kThreads) {
      const int r = e / kNB, c = e % kNB;
      if (r < nb && c < nb) a[(int64_t)(p0 + r) * ld + p0 + c] = P[r][c];
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void gj_inverse_kernel(T* __restrict__ A, int64_t stride, int64_t ld,
                                                              const int32_t* __restrict__ ks,
                                                              int32_t* __restrict__ info) {
  __shared__ GjLds<T> lds;
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const int k = ks[blockIdx.x];
  gj_sweep<T>(A + (int64_t)blockIdx.x * stride, k, k, ld, lds, bad);
  if (threadIdx.x == 0 && info != nullptr) info[blockIdx.x] = bad;
}

// One front of the multifrontal block LDL^T factorisation per workgroup (fronts lo ..
// lo + gridDim.x - 1 of one height): assembly of F (M x M, row-major at work +
// work_off[b]) from the matrix entries in the front's columns and the children's update
// matrices (child by child, in slot order: every entry's sum has the same order on every
// run), then (mode 0) the partial Gauss-Jordan sweep of the nJ pivots and the packing of
// F_JJ^-1, W = -(swept F_SJ) and U = swept F_SS.
__global__ __launch_bounds__(kThreads) void nd_factor_kernel(
    int mode, int64_t lo, const int64_t* __restrict__ fj0, const int64_t* __restrict__ fnj,
    const int64_t* __restrict__ fns, const int64_t* __restrict__ foff,
    const int64_t* __restrict__ wtoff, const int64_t* __restrict__ moff,
    const int64_t* __restrict__ uoff2, const int64_t* __restrict__ child_ptr,
    const int64_t* __restrict__ child_idx, const int32_t* __restrict__ maps,
    const int64_t* __restrict__ s_idx, const int64_t* __restrict__ ap_ptr,
    const int32_t* __restrict__ ap_idx, const double* __restrict__ ap_val,
    const int64_t* __restrict__ work_off, double* __restrict__ work,
    double* __restrict__ ustore, double* __restrict__ fac, int32_t* __restrict__ info) {
  __shared__ GjLds<double> lds;
  __shared__ int bad;
  const int tid = threadIdx.x;
  const int64_t t = lo + blockIdx.x;
  const int64_t a = fj0[t], nJ = fnj[t], nS = fns[t], M = nJ + nS;
  double* F = work + work_off[blockIdx.x];
  if (tid == 0) bad = 0;
  for (int64_t e = tid; e < M * M; e += kThreads) F[e] = 0.0;
  __syncthreads();
  // entries of the permuted matrix in the columns J: row a + r gives F[pos(col)][r] (and
  // the mirror F[r][pos(col)] for a coupling column); columns below a were eliminated below
  const int64_t* S = s_idx + moff[t];
  for (int64_t r = tid; r < nJ; r += kThreads) {
    for (int64_t q = ap_ptr[a + r]; q < ap_ptr[a + r + 1]; ++q) {
      const int64_t col = ap_idx[q];
      if (col < a) continue;
      const double v = ap_val[q];
      if (col < a + nJ) {
        F[(col - a) * M + r] = v;
      } else {
        int64_t lo_s = 0, hi_s = nS;  // position of col in the ascending coupling set
        while (lo_s < hi_s) {
          const int64_t mid = (lo_s + hi_s) >> 1;
          if (S[mid] < col) lo_s = mid + 1;
          else hi_s = mid;
        }
        const int64_t pc = nJ + lo_s;
        F[pc * M + r] = v;
        F[r * M + pc] = v;
      }
    }
  }
  __syncthreads();
  for (int64_t q = child_ptr[t]; q < child_ptr[t + 1]; ++q) {
    const int64_t c = child_idx[q];
    const int64_t nc = fns[c];
    const double* U = ustore + uoff2[c];
    const int32_t* p = maps + moff[c];
    for (int64_t e = tid; e < nc * nc; e += kThreads) {
      const int64_t i = e / nc, j = e - i * nc;
      F[(int64_t)p[i] * M + p[j]] += U[e];
    }
    __syncthreads();
  }
  if (mode != 0) return;
  gj_sweep<double>(F, (int)M, (int)nJ, M, lds, bad);
  const int64_t m = M;
  double* st = fac + foff[t];
  for (int64_t e = tid; e < nJ * nJ; e += kThreads) {
    const int64_t c = e / nJ, r = e - c * nJ;
    st[c * m + r] = F[r * M + c];
  }
  double* wt = fac + wtoff[t];
  for (int64_t e = tid; e < nS * nJ; e += kThreads) {
    const int64_t s = e / nJ, c = e - s * nJ;
    const double w = -F[(nJ + s) * M + c];
    st[c * m + nJ + s] = w;
    wt[s * nJ + c] = w;
  }
  double* u = ustore + uoff2[t];
  for (int64_t e = tid; e < nS * nS; e += kThreads) {
    const int64_t i = e / nS, j = e - i * nS;
    u[e] = F[(nJ + i) * M + nJ + j];
  }
  if (tid == 0) info[blockIdx.x] = bad;
}

// out[i] = vals[seg_ptr[i]] + vals[seg_ptr[i] + 1] + ... + vals[seg_ptr[i + 1] - 1], added
// in index order (one thread per segment: the same sums on every run). Plain: in T, the
// sequential sum. Compensated (Neumaier): the rounding error of each addition is carried in
// a second term, so a long run whose terms cancel (a coarse-matrix entry sums a whole cell
// block of the fine matrix, nearly all of it cancelling) keeps about full precision.
template <typename T, bool kCompensated>
__global__ __launch_bounds__(256) void segment_sum_kernel(int64_t n_seg,
                                                          const int64_t* __restrict__ seg_ptr,
                                                          const T* __restrict__ vals,
                                                          T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_seg) return;
  const int64_t e = seg_ptr[i + 1];
  T acc = T{};
  T comp = T{};
  for (int64_t j = seg_ptr[i]; j < e; ++j) {
    const T v = vals[j];
    if (kCompensated) {
      const T t = acc + v;
      comp += fabs(acc) >= fabs(v) ? (acc - t) + v : (v - t) + acc;
      acc = t;
    } else {
      acc += v;
    }
  }
  out[i] = kCompensated ? acc + comp : acc;
}

// Batched inverse of general (nonsingular) matrices by Gauss-Jordan elimination with
// partial pivoting, in place: one workgroup per matrix (row-major k x k at a, leading
// dimension ld). Step q takes the row p >= q of largest |a_pq| (the smallest such p on
// ties), swaps rows p and q, scales row q by 1/a_qq and eliminates column q from every
// other row: a_ij <- (j == q ? 0 : a_ij) - a_iq a_qj, which leaves the inverse's columns in
// the eliminated positions. The row swaps are undone at the end as column swaps in reverse
// order. A zero pivot column sets info[b] = q + 1 (the matrix is singular) and stops.
// LDS (dynamic): the pivot rows chosen (k ints), column q and row q of the step (k values
// each), and the argmax reduction.
template <typename T>
__global__ __launch_bounds__(kThreads) void pivoted_inverse_kernel(T* __restrict__ A,
                                                                   int64_t stride, int64_t ld,
                                                                   const int32_t* __restrict__ ks,
                                                                   int32_t* __restrict__ info) {
  extern __shared__ __align__(16) unsigned char piv_smem[];
  const int k = ks[blockIdx.x];
  const int kmax = (int)ld;
  T* colq = reinterpret_cast<T*>(piv_smem);
  T* rowq = colq + kmax;
  T* rv = rowq + kmax;
  int* ri = reinterpret_cast<int*>(rv + kThreads);
  int* pr = ri + kThreads;
  T* a = A + (int64_t)blockIdx.x * stride;
  const int tid = threadIdx.x;
  for (int q = 0; q < k; ++q) {
    T best = T(-1);
    int bi = q;
    for (int i = q + tid; i < k; i += kThreads) {
      const T v = fabs(a[(int64_t)i * ld + q]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    rv[tid] = best;
    ri[tid] = bi;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
      if (tid < h && (rv[tid + h] > rv[tid] || (rv[tid + h] == rv[tid] && ri[tid + h] < ri[tid]))) {
        rv[tid] = rv[tid + h];
        ri[tid] = ri[tid + h];
      }
      __syncthreads();
    }
    const int p = ri[0];
    const bool zero = !(rv[0] > T(0));
    __syncthreads();  // everyone has read the result before rv / ri are reused
    if (zero) {
      if (tid == 0 && info != nullptr) info[blockIdx.x] = q + 1;
      return;  // uniform: every thread read the same rv[0]
    }
    if (tid == 0) pr[q] = p;
    if (p != q) {
      for (int j = tid; j < k; j += kThreads) {
        const T t = a[(int64_t)p * ld + j];
        a[(int64_t)p * ld + j] = a[(int64_t)q * ld + j];
        a[(int64_t)q * ld + j] = t;
      }
    }
    __syncthreads();
    const T ipiv = T(1) / a[(int64_t)q * ld + q];
    for (int i = tid; i < k; i += kThreads) colq[i] = a[(int64_t)i * ld + q];
    __syncthreads();  // the pivot and column q are read before row q is overwritten
    for (int j = tid; j < k; j += kThreads) {
      const T v = j == q ? ipiv : a[(int64_t)q * ld + j] * ipiv;
      rowq[j] = v;
      a[(int64_t)q * ld + j] = v;
    }
    __syncthreads();
    for (int i = 0; i < k; ++i) {
      if (i == q) continue;
      const T f = colq[i];
      T* ar = a + (int64_t)i * ld;
      for (int j = tid; j < k; j += kThreads) ar[j] = (j == q ? T(0) : ar[j]) - f * rowq[j];
    }
    __syncthreads();
  }
  for (int q = k - 1; q >= 0; --q) {
    const int p = pr[q];
    if (p != q) {
      for (int i = tid; i < k; i += kThreads) {
        const T t = a[(int64_t)i * ld + p];
        a[(int64_t)i * ld + p] = a[(int64_t)i * ld + q];
        a[(int64_t)i * ld + q] = t;
      }
    }
    __syncthreads();
  }
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_segment_sum(int dtype, int compensated, int64_t n_seg, const int64_t* seg_ptr,
                               const void* vals, void* out, void* stream) {
  DDS_REQUIRE(n_seg >= 0 && n_seg <= (int64_t)INT32_MAX * 256, "dds_segment_sum: bad count");
  if (n_seg == 0) return DDS_OK;
  DDS_REQUIRE(seg_ptr && vals && out, "dds_segment_sum: null pointer");
  hipStream_t s = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n_seg, 256));
  const double* vd = static_cast<const double*>(vals);
  const float* vf = static_cast<const float*>(vals);
  if (dtype == DDS_F64 && compensated) {
    hipLaunchKernelGGL((segment_sum_kernel<double, true>), grid, dim3(256), 0, s, n_seg, seg_ptr,
                       vd, static_cast<double*>(out));
  } else if (dtype == DDS_F64) {
    hipLaunchKernelGGL((segment_sum_kernel<double, false>), grid, dim3(256), 0, s, n_seg,
                       seg_ptr, vd, static_cast<double*>(out));
  } else if (dtype == DDS_F32 && compensated) {
    hipLaunchKernelGGL((segment_sum_kernel<float, true>), grid, dim3(256), 0, s, n_seg, seg_ptr,
                       vf, static_cast<float*>(out));
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL((segment_sum_kernel<float, false>), grid, dim3(256), 0, s, n_seg, seg_ptr,
                       vf, static_cast<float*>(out));
  } else {
    set_error("dds_segment_sum: dtype must be F64 or F32 (got %d)", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_spd_inverse_batched(int dtype, int64_t n_mats, const int32_t* k, int64_t ld,
                                       int64_t stride, void* A, int32_t* info, void* stream) {
  DDS_REQUIRE(n_mats >= 0 && n_mats <= INT32_MAX, "dds_spd_inverse_batched: bad batch size");
  DDS_REQUIRE(k && A, "dds_spd_inverse_batched: null pointer");
  DDS_REQUIRE(ld >= 1 && stride >= ld, "dds_spd_inverse_batched: bad ld/stride");
  if (n_mats == 0) return DDS_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(gj_inverse_kernel<double>, dim3((unsigned)n_mats), dim3(kThreads), 0, s,
                       static_cast<double*>(A), stride, ld, k, info);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(gj_inverse_kernel<float>, dim3((unsigned)n_mats), dim3(kThreads), 0, s,
                       static_cast<float*>(A), stride, ld, k, info);
  } else {
    set_error("dds_spd_inverse_batched: dtype must be F64 or F32 (got %d)", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_nd_factor(int mode, int64_t lo, int64_t hi, const int64_t* fj0,
                             const int64_t* fnj, const int64_t* fns, const int64_t* foff,
                             const int64_t* wtoff, const int64_t* moff, const int64_t* uoff2,
                             const int64_t* child_ptr, const int64_t* child_idx,
                             const int32_t* maps, const int64_t* s_idx, const int64_t* ap_ptr,
                             const int32_t* ap_idx, const double* ap_val,
                             const int64_t* work_off, double* work, double* ustore, double* fac,
                             int32_t* info, void* stream) {
  DDS_REQUIRE((mode == 0 || mode == 1) && lo >= 0 && hi >= lo && hi - lo <= INT32_MAX,
              "dds_nd_factor: bad mode or front range");
  if (hi == lo) return DDS_OK;
  DDS_REQUIRE(fj0 && fnj && fns && moff && uoff2 && child_ptr && ap_ptr && work_off && work,
              "dds_nd_factor: null table");
  DDS_REQUIRE(mode == 1 || (foff && wtoff && ustore && fac && info), "dds_nd_factor: null output");
  hipLaunchKernelGGL(nd_factor_kernel, dim3((unsigned)(hi - lo)), dim3(kThreads), 0,
                     as_stream(stream), mode, lo, fj0, fnj, fns, foff, wtoff, moff, uoff2,
                     child_ptr, child_idx, maps, s_idx, ap_ptr, ap_idx, ap_val, work_off, work,
                     ustore, fac, info);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_inverse_batched(int dtype, int64_t n_mats, const int32_t* k, int64_t ld,
                                   int64_t stride, void* A, int32_t* info, void* stream) {
  DDS_REQUIRE(n_mats >= 0 && n_mats <= INT32_MAX, "dds_inverse_batched: bad batch size");
  DDS_REQUIRE(ld >= 1 && ld <= 4096 && stride >= ld * ld,
              "dds_inverse_batched: bad ld/stride (ld <= 4096, stride >= ld * ld)");
  if (n_mats == 0) return DDS_OK;
  DDS_REQUIRE(k && A, "dds_inverse_batched: null pointer");
  hipStream_t s = as_stream(stream);
  const size_t esz = dtype == DDS_F64 ? 8 : 4;
  const size_t lds = (2 * (size_t)ld + kThreads) * esz + (kThreads + (size_t)ld) * 4;
#define DDS_PIVINV(T)                                                                       \
  do {                                                                                      \
    if (lds > 65536)                                                                        \
      DDS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&pivoted_inverse_kernel<T>), \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));   \
    hipLaunchKernelGGL(pivoted_inverse_kernel<T>, dim3((unsigned)n_mats), dim3(kThreads), lds, \
                       s, static_cast<T*>(A), stride, ld, k, info);                         \
  } while (0)
  if (dtype == DDS_F64) {
    DDS_PIVINV(double);
  } else if (dtype == DDS_F32) {
    DDS_PIVINV(float);
  } else {
    set_error("dds_inverse_batched: dtype must be F64 or F32 (got %d)", dtype);
    return DDS_EUNSUPPORTED;
  }
#undef DDS_PIVINV
  DDS_LAUNCHED();
  return DDS_OK;
}

