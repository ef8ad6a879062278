// Sparse coarse solve: the two sweeps of a nested-dissection block LDL^T factor
// (A = L D L^T with L_SJ = W = F_SJ F_JJ^-1 and D_J = F_JJ per front), one launch per
// height of the dissection tree.
//
// Front t is described by fronts[8t .. 8t+7] = (j0, nJ, nS, foff, fpos, fu, fsidx, -):
//   unknowns j0 .. j0+nJ-1 of the nested-dissection numbering (perm[i] = original index);
//   factor rows [Finv (nJ x nJ); W (nS x nJ)] row-major at fac + foff, so row r of the
//   stack is at fac + foff + r*nJ;
//   its nJ+nS front positions at gather-list positions fpos .. fpos+nJ+nS-1: position p
//   receives u[gsrc[k]] for k in [gptr[fpos+p], gptr[fpos+p+1]) (the children's update
//   vectors, in child order);
//   its update vector at u + fu (nS), its coupling set (nested-dissection indices of the
//   ancestors' unknowns it touches) at sidx + fsidx (nS).
// Every sum runs in a fixed order, so a solve is bitwise reproducible.
#include "common.hpp"

namespace dds {
namespace {

constexpr int kNdThreads = 256;
constexpr int kNdWaves = kNdThreads / 64;
constexpr int64_t kNdLdsCap = 7680;  // elements of f_J (forward) or x_S (backward): 60 KiB

struct Front {
  int64_t j0, nJ, nS, foff, fpos, fu, fsidx;
};

__device__ __forceinline__ Front load_front(const int64_t* __restrict__ fronts, int t) {
  const int64_t* f = fronts + 8 * (int64_t)t;
  return Front{f[0], f[1], f[2], f[3], f[4], f[5], f[6]};
}

// Work item (t, r0, nr): rows r0 .. r0+nr-1 of front t's stacked factor.
//   f_J = b[perm[j0 + i]] + gathered updates            (every workgroup of the front)
//   r <  nJ:  g[j0 + r]      = Finv[r] . f_J
//   r >= nJ:  u[fu + r - nJ] = (gathered updates at r) - W[r - nJ] . f_J
// Wave w takes rows r0 + w, r0 + w + 4, ...; lanes split a row's columns and a butterfly
// adds them.
template <typename T>
__global__ __launch_bounds__(kNdThreads) void nd_forward_kernel(
    const int32_t* __restrict__ items, const int64_t* __restrict__ fronts,
    const int64_t* __restrict__ gptr, const int32_t* __restrict__ gsrc,
    const int32_t* __restrict__ perm, const T* __restrict__ fac, const T* __restrict__ b,
    T* __restrict__ g, T* __restrict__ u, const dds_pcg_state* __restrict__ gate) {
  extern __shared__ double f[];
  if (gate != nullptr && gate->done) return;
  const int t = items[3 * blockIdx.x];
  const int r0 = items[3 * blockIdx.x + 1];
  const int nr = items[3 * blockIdx.x + 2];
  const Front fr = load_front(fronts, t);
  const int nJ = (int)fr.nJ;
  for (int i = threadIdx.x; i < nJ; i += kNdThreads) {
    double v = static_cast<double>(b[perm[fr.j0 + i]]);
    const int64_t k1 = gptr[fr.fpos + i + 1];
    for (int64_t k = gptr[fr.fpos + i]; k < k1; ++k) v += static_cast<double>(u[gsrc[k]]);
    f[i] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  for (int r = r0 + w; r < r0 + nr; r += kNdWaves) {
    const T* row = fac + fr.foff + (int64_t)r * nJ;
    double a0 = 0.0, a1 = 0.0;
    int c = lane;
    for (; c + 64 < nJ; c += 128) {
      a0 += static_cast<double>(load_nt(row + c)) * f[c];
      a1 += static_cast<double>(load_nt(row + c + 64)) * f[c + 64];
    }
    if (c < nJ) a0 += static_cast<double>(load_nt(row + c)) * f[c];
    const double acc = wave_sum(a0 + a1);
    if (lane == 0) {
      if (r < nJ) {
        g[fr.j0 + r] = static_cast<T>(acc);
      } else {
        double v = 0.0;
        const int64_t k1 = gptr[fr.fpos + r + 1];
        for (int64_t k = gptr[fr.fpos + r]; k < k1; ++k) v += static_cast<double>(u[gsrc[k]]);
        u[fr.fu + (r - nJ)] = static_cast<T>(v - acc);
      }
    }
  }
}

// Work item (t, c0, nc), nc <= 64: x_J[c] = g[j0 + c] - sum_s W[s][c] x_S[s] for the nc
// columns c0 .. c0+nc-1 (lane = column), x_S gathered from the ancestors' solution. Wave
// w sums the rows s = w, w+4, ...; the four wave sums are added in wave order.
template <typename T>
__global__ __launch_bounds__(kNdThreads) void nd_backward_kernel(
    const int32_t* __restrict__ items, const int64_t* __restrict__ fronts,
    const int32_t* __restrict__ sidx, const int32_t* __restrict__ perm,
    const T* __restrict__ fac, const T* __restrict__ g, T* __restrict__ xnd,
    T* __restrict__ out, const dds_pcg_state* __restrict__ gate) {
  extern __shared__ double xs[];
  __shared__ double red[kNdWaves][64];
  if (gate != nullptr && gate->done) return;
  const int t = items[3 * blockIdx.x];
  const int c0 = items[3 * blockIdx.x + 1];
  const int nc = items[3 * blockIdx.x + 2];
  const Front fr = load_front(fronts, t);
  const int nJ = (int)fr.nJ;
  const int nS = (int)fr.nS;
  for (int i = threadIdx.x; i < nS; i += kNdThreads)
    xs[i] = static_cast<double>(xnd[sidx[fr.fsidx + i]]);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  double acc = 0.0;
  if (lane < nc) {
    const T* W = fac + fr.foff + (int64_t)nJ * nJ + c0 + lane;
    double a0 = 0.0, a1 = 0.0;
    int s = w;
    for (; s + kNdWaves < nS; s += 2 * kNdWaves) {
      a0 += static_cast<double>(load_nt(W + (int64_t)s * nJ)) * xs[s];
      a1 += static_cast<double>(load_nt(W + (int64_t)(s + kNdWaves) * nJ)) * xs[s + kNdWaves];
    }
    if (s < nS) a0 += static_cast<double>(load_nt(W + (int64_t)s * nJ)) * xs[s];
    acc = a0 + a1;
  }
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && lane < nc) {
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < kNdWaves; ++i) sum += red[i][lane];
    const int64_t j = fr.j0 + c0 + lane;
    const double x = static_cast<double>(g[j]) - sum;
    xnd[j] = static_cast<T>(x);
    out[perm[j]] = static_cast<T>(x);
  }
}

template <typename T>
int launch_nd_forward(int64_t n_items, const int32_t* items, const int64_t* fronts,
                      const int64_t* gptr, const int32_t* gsrc, const int32_t* perm,
                      const void* fac, const void* b, void* g, void* u, int64_t lds,
                      const dds_pcg_state* gate, hipStream_t s) {
  hipLaunchKernelGGL(nd_forward_kernel<T>, dim3((unsigned)n_items), dim3(kNdThreads),
                     (size_t)std::max<int64_t>(lds, 1) * sizeof(double), s, items, fronts, gptr,
                     gsrc, perm, static_cast<const T*>(fac), static_cast<const T*>(b),
                     static_cast<T*>(g), static_cast<T*>(u), gate);
  DDS_LAUNCHED();
  return DDS_OK;
}

template <typename T>
int launch_nd_backward(int64_t n_items, const int32_t* items, const int64_t* fronts,
                       const int32_t* sidx, const int32_t* perm, const void* fac,
                       const void* g, void* xnd, void* out, int64_t lds,
                       const dds_pcg_state* gate, hipStream_t s) {
  hipLaunchKernelGGL(nd_backward_kernel<T>, dim3((unsigned)n_items), dim3(kNdThreads),
                     (size_t)std::max<int64_t>(lds, 1) * sizeof(double), s, items, fronts, sidx,
                     perm, static_cast<const T*>(fac), static_cast<const T*>(g),
                     static_cast<T*>(xnd), static_cast<T*>(out), gate);
  DDS_LAUNCHED();
  return DDS_OK;
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_nd_forward(int dtype, int64_t n_items, const int32_t* items,
                              const int64_t* fronts, const int64_t* gptr, const int32_t* gsrc,
                              const int32_t* perm, const void* fac, const void* b, void* g,
                              void* u, int64_t lds_elems, const dds_pcg_state* gate,
                              void* stream) {
  DDS_REQUIRE(n_items >= 0 && n_items <= INT32_MAX, "dds_nd_forward: bad item count");
  DDS_REQUIRE(lds_elems >= 0 && lds_elems <= kNdLdsCap, "dds_nd_forward: front of %lld > %lld",
              (long long)lds_elems, (long long)kNdLdsCap);
  if (n_items == 0) return DDS_OK;
  DDS_REQUIRE(items && fronts && gptr && perm && fac && b && g && u,
              "dds_nd_forward: null pointer");
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64)
    return launch_nd_forward<double>(n_items, items, fronts, gptr, gsrc, perm, fac, b, g, u,
                                     lds_elems, gate, s);
  if (dtype == DDS_F32)
    return launch_nd_forward<float>(n_items, items, fronts, gptr, gsrc, perm, fac, b, g, u,
                                    lds_elems, gate, s);
  set_error("dds_nd_forward: unsupported dtype %d", dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_nd_backward(int dtype, int64_t n_items, const int32_t* items,
                               const int64_t* fronts, const int32_t* sidx, const int32_t* perm,
                               const void* fac, const void* g, void* x_nd, void* out,
                               int64_t lds_elems, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_items >= 0 && n_items <= INT32_MAX, "dds_nd_backward: bad item count");
  DDS_REQUIRE(lds_elems >= 0 && lds_elems <= kNdLdsCap, "dds_nd_backward: front of %lld > %lld",
              (long long)lds_elems, (long long)kNdLdsCap);
  if (n_items == 0) return DDS_OK;
  DDS_REQUIRE(items && fronts && perm && fac && g && x_nd && out,
              "dds_nd_backward: null pointer");
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64)
    return launch_nd_backward<double>(n_items, items, fronts, sidx, perm, fac, g, x_nd, out,
                                      lds_elems, gate, s);
  if (dtype == DDS_F32)
    return launch_nd_backward<float>(n_items, items, fronts, sidx, perm, fac, g, x_nd, out,
                                     lds_elems, gate, s);
  set_error("dds_nd_backward: unsupported dtype %d", dtype);
  return DDS_EUNSUPPORTED;
}
