// Sparse coarse solve: the sweeps of a nested-dissection block LDL^T factor
// (A = L D L^T with L_SJ = W = F_SJ F_JJ^-1 and D_J = F_JJ per front), one launch per
// height of the dissection tree below a dense top front.
//
// A front owns the unknowns j0 .. j0+nJ-1 of the nested-dissection numbering
// (perm[i] = original index) and couples to nS unknowns of its ancestors; its stacked
// factor [F_JJ^-1 (nJ x nJ); W (nS x nJ)] is column-major at fac + foff, element (r, c) at
// foff + c*m + r with m = nJ + nS. Data moves between levels only by pushes into fixed
// buffers, so every kernel reads its inputs with one level of indexing:
//   forward:  a child leaves its update vector u in its own slot of its parent's front,
//             slots[wbase + maps[moff + s]] = u_s; a front adds its children's slots in
//             child order (entries that no child maps stay zero);
//   backward: a front writes x at each child's coupling set into that child's buffer,
//             xs[push_dst[k]] = x at front position push_src[k].
// Every sum runs in a fixed order, so a solve is bitwise reproducible.
#include "common.hpp"

namespace dds {
namespace {

constexpr int kNdThreads = 256;
constexpr int kNdWaves = kNdThreads / 64;
static_assert(kNdWaves == 4, "the top assembly adds four wave sums");
constexpr int64_t kNdLdsCap = 7680;  // dynamic LDS elements per workgroup: 60 KiB of FP64

// Forward item (12 x int64): j0, nJ, nS, foff, soff, n_children, wbase, moff, r0, nr, R.
// The item covers rows r0 .. r0+nr-1 (nr <= R) of the stacked factor; R (64, 128 or 256)
// rows take the 256 threads as R lanes x 256/R column parts: thread t has row r0 + t % R
// and part t / R, which sums a contiguous quarter, half or all of the columns (four
// accumulators, c ascending); the parts are added in part order through LDS. Then
//   f_J = b[perm[j0 + .]] + the children's slots (child order), staged in LDS;
//   r <  nJ:  g[j0 + r] = sum_c Finv[r][c] f_J[c]
//   r >= nJ:  the parent's slot gets (children's slots at r) - sum_c W[r-nJ][c] f_J[c].
// Splitting the columns shortens each thread's chain of dependent loads for the wide
// fronts near the roots, where few workgroups are in flight.
template <typename T, typename F>
__global__ __launch_bounds__(kNdThreads) void nd_forward_kernel(
    const int64_t* __restrict__ items, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ maps, const F* __restrict__ fac, const T* __restrict__ b,
    T* __restrict__ g, T* __restrict__ slots, const dds_pcg_state* __restrict__ gate) {
  extern __shared__ double f[];
  __shared__ double red[kNdThreads];
  if (gate != nullptr && gate->done) return;
  const int64_t* it = items + 12 * (int64_t)blockIdx.x;
  const int64_t j0 = it[0], foff = it[3], soff = it[4], wbase = it[6], moff = it[7];
  const int nJ = (int)it[1], nS = (int)it[2], nch = (int)it[5], r0 = (int)it[8], nr = (int)it[9];
  const int R = (int)it[10];  // 64, 128 or 256 (checked on the host)
  const int K = kNdThreads / R;
  const int m = nJ + nS;
  for (int p = threadIdx.x; p < nJ; p += kNdThreads) {
    double v = static_cast<double>(b[perm[j0 + p]]);
    for (int q = 0; q < nch; ++q) v += static_cast<double>(slots[soff + (int64_t)q * m + p]);
    f[p] = v;
  }
  const int lr = (int)threadIdx.x & (R - 1);
  const int part = (int)threadIdx.x / R;
  const int r = r0 + lr;
  const bool active = lr < nr;
  double fs = 0.0;  // the children's update at an S row
  int32_t dst = 0;
  if (part == 0 && active && r >= nJ) {
    for (int q = 0; q < nch; ++q) fs += static_cast<double>(slots[soff + (int64_t)q * m + r]);
    dst = maps[moff + (r - nJ)];
  }
  __syncthreads();
  double acc = 0.0;
  if (active) {
    const int q = (nJ + K - 1) / K;
    const int cb = part * q;
    const int ce = min(nJ, cb + q);
    const F* col = fac + foff + r;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int c = cb;
    for (; c + 7 < ce; c += 8) {  // eight loads in flight, then the same four sums
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = static_cast<double>(load_nt(col + (int64_t)(c + u) * m));
      a0 += v[0] * f[c];
      a1 += v[1] * f[c + 1];
      a2 += v[2] * f[c + 2];
      a3 += v[3] * f[c + 3];
      a0 += v[4] * f[c + 4];
      a1 += v[5] * f[c + 5];
      a2 += v[6] * f[c + 6];
      a3 += v[7] * f[c + 7];
    }
    for (; c + 3 < ce; c += 4) {
      a0 += static_cast<double>(load_nt(col + (int64_t)c * m)) * f[c];
      a1 += static_cast<double>(load_nt(col + (int64_t)(c + 1) * m)) * f[c + 1];
      a2 += static_cast<double>(load_nt(col + (int64_t)(c + 2) * m)) * f[c + 2];
      a3 += static_cast<double>(load_nt(col + (int64_t)(c + 3) * m)) * f[c + 3];
    }
    for (; c < ce; ++c) a0 += static_cast<double>(load_nt(col + (int64_t)c * m)) * f[c];
    acc = (a0 + a1) + (a2 + a3);
  }
  if (K > 1) {
    red[threadIdx.x] = acc;
    __syncthreads();
    if (part == 0) {
      acc = red[lr];
      for (int k = 1; k < K; ++k) acc += red[k * R + lr];
    }
  }
  if (part != 0 || !active) return;
  if (r < nJ)
    g[j0 + r] = static_cast<T>(acc);
  else
    slots[wbase + dst] = static_cast<T>(fs - acc);
}

// Backward item (12 x int64): j0, nJ, nS, wtoff, xoff, c0, nc, k0, k1, C, with W
// row-major at fac + wtoff. The item covers columns c0 .. c0+nc-1 (nc <= C); C (16, 32 or
// 64) columns take each wavefront as C lanes x 64/C row groups, so the 256 threads form
// P = 256/C parts (part = wave * 64/C + lane / C), and part p sums the rows s of the p-th
// contiguous 1/P of [0, nS) (four accumulators, s ascending); the parts are added in part
// order through LDS:
//   x_J[c] = g[j0 + c] - sum_s W[s][c] x_S[s]   (x_S = xs[xoff .. xoff+nS)),
// written to out[perm[j0 + c]]; then the push entries [k0, k1) copy x_J (sources in
// [c0, c0+nc)) or x_S (sources >= nJ) into the children's buffers. Narrow fronts use
// narrow C, so fewer lanes idle.
template <typename T, typename F>
__global__ __launch_bounds__(kNdThreads) void nd_backward_kernel(
    const int64_t* __restrict__ items, const int32_t* __restrict__ perm,
    const int64_t* __restrict__ push_dst, const int32_t* __restrict__ push_src,
    const F* __restrict__ fac, const T* __restrict__ g, T* __restrict__ xs,
    T* __restrict__ out, const dds_pcg_state* __restrict__ gate) {
  extern __shared__ double xsl[];
  __shared__ double red[kNdThreads];
  __shared__ double xj[64];
  if (gate != nullptr && gate->done) return;
  const int64_t* it = items + 12 * (int64_t)blockIdx.x;
  const int64_t j0 = it[0], wtoff = it[3], xoff = it[4], k0 = it[7], k1 = it[8];
  const int nJ = (int)it[1], nS = (int)it[2], c0 = (int)it[5], nc = (int)it[6];
  const int C = (int)it[9];  // 16, 32 or 64 (checked on the host)
  const int P = kNdThreads / C;
  for (int s = threadIdx.x; s < nS; s += kNdThreads) xsl[s] = static_cast<double>(xs[xoff + s]);
  __syncthreads();
  const int lc = (int)threadIdx.x & (C - 1);
  const int part = (int)threadIdx.x / C;
  const int q = (nS + P - 1) / P;
  const int s0 = min(nS, part * q);
  const int s1 = min(nS, s0 + q);
  double acc = 0.0;
  if (lc < nc) {
    const F* W = fac + wtoff + c0 + lc;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int s = s0;
    for (; s + 7 < s1; s += 8) {  // eight loads in flight, then the same four sums
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = static_cast<double>(load_nt(W + (int64_t)(s + u) * nJ));
      a0 += v[0] * xsl[s];
      a1 += v[1] * xsl[s + 1];
      a2 += v[2] * xsl[s + 2];
      a3 += v[3] * xsl[s + 3];
      a0 += v[4] * xsl[s + 4];
      a1 += v[5] * xsl[s + 5];
      a2 += v[6] * xsl[s + 6];
      a3 += v[7] * xsl[s + 7];
    }
    for (; s + 3 < s1; s += 4) {
      a0 += static_cast<double>(load_nt(W + (int64_t)s * nJ)) * xsl[s];
      a1 += static_cast<double>(load_nt(W + (int64_t)(s + 1) * nJ)) * xsl[s + 1];
      a2 += static_cast<double>(load_nt(W + (int64_t)(s + 2) * nJ)) * xsl[s + 2];
      a3 += static_cast<double>(load_nt(W + (int64_t)(s + 3) * nJ)) * xsl[s + 3];
    }
    for (; s < s1; ++s) a0 += static_cast<double>(load_nt(W + (int64_t)s * nJ)) * xsl[s];
    acc = (a0 + a1) + (a2 + a3);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if ((int)threadIdx.x < nc) {
    const int c = (int)threadIdx.x;
    double sum = red[c];
    for (int p = 1; p < P; ++p) sum += red[p * C + c];
    const int64_t j = j0 + c0 + c;
    const double x = static_cast<double>(g[j]) - sum;
    xj[c] = x;
    out[perm[j]] = static_cast<T>(x);
  }
  __syncthreads();
  for (int64_t k = k0 + threadIdx.x; k < k1; k += kNdThreads) {
    const int src = push_src[k];
    const double v = src < nJ ? xj[src - c0] : xsl[src - nJ];
    xs[push_dst[k]] = static_cast<T>(v);
  }
}

// Top front (unknowns j0 .. j0+nT-1): f = b[perm[j0 + p]] + its children's slots. A
// workgroup takes 64 positions; wave g sums the children q = g, g+4, ... (two loads in
// flight per step), and the four wave sums are added in wave order.
template <typename T>
__global__ __launch_bounds__(kNdThreads) void nd_top_assemble_kernel(
    int64_t j0, int64_t nT, int64_t soff, int nch, const int32_t* __restrict__ perm,
    const T* __restrict__ b, const T* __restrict__ slots, T* __restrict__ ftop,
    const dds_pcg_state* __restrict__ gate) {
  __shared__ double part[kNdWaves][64];
  if (gate != nullptr && gate->done) return;
  const int pl = threadIdx.x & 63;
  const int g = threadIdx.x >> 6;
  const int64_t p = (int64_t)blockIdx.x * 64 + pl;
  double a0 = 0.0, a1 = 0.0;
  if (p < nT) {
    int q = g;
    for (; q + kNdWaves < nch; q += 2 * kNdWaves) {
      a0 += static_cast<double>(slots[soff + (int64_t)q * nT + p]);
      a1 += static_cast<double>(slots[soff + (int64_t)(q + kNdWaves) * nT + p]);
    }
    if (q < nch) a0 += static_cast<double>(slots[soff + (int64_t)q * nT + p]);
  }
  part[g][pl] = a0 + a1;
  __syncthreads();
  if (g == 0 && p < nT) {
    const double c = (part[0][pl] + part[1][pl]) + (part[2][pl] + part[3][pl]);
    ftop[p] = static_cast<T>(static_cast<double>(b[perm[j0 + p]]) + c);
  }
}

// Top front solution x (from the dense inverse): out[perm[j0 + p]] = x[p], and the
// pushes of x into the top's children's buffers.
template <typename T>
__global__ __launch_bounds__(kNdThreads) void nd_top_push_kernel(
    int64_t j0, int64_t nT, int64_t n_push, const int32_t* __restrict__ perm,
    const T* __restrict__ xtop, const int64_t* __restrict__ push_dst,
    const int32_t* __restrict__ push_src, T* __restrict__ xs, T* __restrict__ out,
    const dds_pcg_state* __restrict__ gate) {
  if (gate != nullptr && gate->done) return;
  const int64_t i = (int64_t)blockIdx.x * kNdThreads + threadIdx.x;
  if (i < nT) out[perm[j0 + i]] = xtop[i];
  if (i < n_push) xs[push_dst[i]] = xtop[push_src[i]];
}

template <typename T, typename F>
int nd_forward(int64_t n_items, const int64_t* items, const int32_t* perm, const int32_t* maps,
               const void* fac, const void* b, void* g, void* slots, int64_t lds,
               const dds_pcg_state* gate, hipStream_t s) {
  hipLaunchKernelGGL((nd_forward_kernel<T, F>), dim3((unsigned)n_items), dim3(kNdThreads),
                     (size_t)std::max<int64_t>(lds, 1) * sizeof(double), s, items, perm, maps,
                     static_cast<const F*>(fac), static_cast<const T*>(b), static_cast<T*>(g),
                     static_cast<T*>(slots), gate);
  DDS_LAUNCHED();
  return DDS_OK;
}

template <typename T, typename F>
int nd_backward(int64_t n_items, const int64_t* items, const int32_t* perm,
                const int64_t* push_dst, const int32_t* push_src, const void* fac,
                const void* g, void* xs, void* out, int64_t lds, const dds_pcg_state* gate,
                hipStream_t s) {
  hipLaunchKernelGGL((nd_backward_kernel<T, F>), dim3((unsigned)n_items), dim3(kNdThreads),
                     (size_t)std::max<int64_t>(lds, 1) * sizeof(double), s, items, perm,
                     push_dst, push_src, static_cast<const F*>(fac), static_cast<const T*>(g),
                     static_cast<T*>(xs), static_cast<T*>(out), gate);
  DDS_LAUNCHED();
  return DDS_OK;
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_nd_forward(int dtype, int fac_dtype, int64_t n_items,
                              const int64_t* items, const int32_t* perm, const int32_t* maps,
                              const void* fac, const void* b, void* g, void* slots,
                              int64_t lds_elems, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_items >= 0 && n_items <= INT32_MAX, "dds_nd_forward: bad item count");
  DDS_REQUIRE(lds_elems >= 0 && lds_elems <= kNdLdsCap, "dds_nd_forward: front of %lld > %lld",
              (long long)lds_elems, (long long)kNdLdsCap);
  if (n_items == 0) return DDS_OK;
  DDS_REQUIRE(items && perm && fac && b && g && slots, "dds_nd_forward: null pointer");
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64 && fac_dtype == DDS_F64)
    return nd_forward<double, double>(n_items, items, perm, maps, fac, b, g, slots, lds_elems,
                                      gate, s);
  if (dtype == DDS_F64 && fac_dtype == DDS_F32)
    return nd_forward<double, float>(n_items, items, perm, maps, fac, b, g, slots, lds_elems,
                                     gate, s);
  if (dtype == DDS_F32 && fac_dtype == DDS_F32)
    return nd_forward<float, float>(n_items, items, perm, maps, fac, b, g, slots, lds_elems,
                                    gate, s);
  set_error("dds_nd_forward: unsupported dtypes %d / %d", dtype, fac_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_nd_backward(int dtype, int fac_dtype, int64_t n_items,
                               const int64_t* items, const int32_t* perm,
                               const int64_t* push_dst, const int32_t* push_src,
                               const void* fac, const void* g, void* xs, void* out,
                               int64_t lds_elems, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_items >= 0 && n_items <= INT32_MAX, "dds_nd_backward: bad item count");
  DDS_REQUIRE(lds_elems >= 0 && lds_elems <= kNdLdsCap, "dds_nd_backward: front of %lld > %lld",
              (long long)lds_elems, (long long)kNdLdsCap);
  if (n_items == 0) return DDS_OK;
  DDS_REQUIRE(items && perm && fac && g && xs && out, "dds_nd_backward: null pointer");
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64 && fac_dtype == DDS_F64)
    return nd_backward<double, double>(n_items, items, perm, push_dst, push_src, fac, g, xs,
                                       out, lds_elems, gate, s);
  if (dtype == DDS_F64 && fac_dtype == DDS_F32)
    return nd_backward<double, float>(n_items, items, perm, push_dst, push_src, fac, g, xs,
                                      out, lds_elems, gate, s);
  if (dtype == DDS_F32 && fac_dtype == DDS_F32)
    return nd_backward<float, float>(n_items, items, perm, push_dst, push_src, fac, g, xs, out,
                                     lds_elems, gate, s);
  set_error("dds_nd_backward: unsupported dtypes %d / %d", dtype, fac_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_nd_top_assemble(int dtype, int64_t j0, int64_t n_top, int64_t soff,
                                   int64_t n_children, const int32_t* perm, const void* b,
                                   const void* slots, void* ftop, const dds_pcg_state* gate,
                                   void* stream) {
  DDS_REQUIRE(j0 >= 0 && n_top >= 0 && soff >= 0 && n_children >= 0 && n_children <= INT32_MAX,
              "dds_nd_top_assemble: bad sizes");
  if (n_top == 0) return DDS_OK;
  DDS_REQUIRE(perm && b && ftop && (slots || n_children == 0), "dds_nd_top_assemble: null pointer");
  hipStream_t s = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n_top, 64));
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(nd_top_assemble_kernel<double>, grid, dim3(kNdThreads), 0, s, j0, n_top,
                       soff, (int)n_children, perm, static_cast<const double*>(b),
                       static_cast<const double*>(slots), static_cast<double*>(ftop), gate);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(nd_top_assemble_kernel<float>, grid, dim3(kNdThreads), 0, s, j0, n_top,
                       soff, (int)n_children, perm, static_cast<const float*>(b),
                       static_cast<const float*>(slots), static_cast<float*>(ftop), gate);
  } else {
    set_error("dds_nd_top_assemble: unsupported dtype %d", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_nd_top_push(int dtype, int64_t j0, int64_t n_top, int64_t n_push,
                               const int32_t* perm, const void* xtop, const int64_t* push_dst,
                               const int32_t* push_src, void* xs, void* out,
                               const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(j0 >= 0 && n_top >= 0 && n_push >= 0, "dds_nd_top_push: bad sizes");
  const int64_t n = std::max(n_top, n_push);
  if (n == 0) return DDS_OK;
  DDS_REQUIRE(perm && xtop && out && (n_push == 0 || (push_dst && push_src && xs)),
              "dds_nd_top_push: null pointer");
  hipStream_t s = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n, kNdThreads));
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(nd_top_push_kernel<double>, grid, dim3(kNdThreads), 0, s, j0, n_top,
                       n_push, perm, static_cast<const double*>(xtop), push_dst, push_src,
                       static_cast<double*>(xs), static_cast<double*>(out), gate);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(nd_top_push_kernel<float>, grid, dim3(kNdThreads), 0, s, j0, n_top,
                       n_push, perm, static_cast<const float*>(xtop), push_dst, push_src,
                       static_cast<float*>(xs), static_cast<float*>(out), gate);
  } else {
    set_error("dds_nd_top_push: unsupported dtype %d", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}
