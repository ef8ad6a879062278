// PCG vector kernels with device-side scalars and norm test, the overlapped prolongation
// sum, the coarse restriction and dtype casts.
#include "common.hpp"

namespace dds {
namespace {

constexpr int kBlock = 256;

int64_t vec_grid(int64_t n) {
  return std::min<int64_t>(ceil_div(std::max<int64_t>(n, 1), kBlock), kVecGrid);
}

__device__ __forceinline__ bool gated(const dds_pcg_state* g) { return g != nullptr && g->done; }

__device__ __forceinline__ void write_partial(double v, double* partials, double* scratch) {
  const double t = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void residual_kernel(int64_t n, const double* __restrict__ b,
                                                          const double* __restrict__ q,
                                                          double* __restrict__ r,
                                                          double* __restrict__ partials,
                                                          const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kBlock / 64];
  if (gated(gate)) return;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double v = b[i] - q[i];
    r[i] = v;
    acc += v * v;
  }
  write_partial(acc, partials, scratch);
}

__global__ __launch_bounds__(kBlock) void update_p_kernel(int64_t n, const double* __restrict__ z,
                                                          double* __restrict__ p,
                                                          const dds_pcg_state* __restrict__ st) {
  if (st->done) return;
  const bool first = st->it == 0;
  const double beta = st->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    p[i] = first ? z[i] : p[i] * beta + z[i];
  }
}

__global__ __launch_bounds__(kBlock) void update_xr_kernel(int64_t n, const double* __restrict__ p,
                                                           const double* __restrict__ q,
                                                           double* __restrict__ x,
                                                           double* __restrict__ r,
                                                           double* __restrict__ partials,
                                                           const dds_pcg_state* __restrict__ st) {
  __shared__ double scratch[kBlock / 64];
  if (st->done) return;
  const double alpha = st->alpha;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double v = r[i] - alpha * q[i];
    r[i] = v;
    acc += v * v;
  }
  write_partial(acc, partials, scratch);
}

__global__ __launch_bounds__(kBlock) void dot_kernel(int64_t n, const double* __restrict__ a,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ partials,
                                                     const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kBlock / 64];
  if (gated(gate)) return;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    acc += a[i] * b[i];
  write_partial(acc, partials, scratch);
}

// Fixed-order sum of the partials: lane t keeps four sums over t + 256 (4i + q), i = 0, 1,
// ..., so four independent loads are in flight per step, adds them in order, then the
// block tree.
__device__ double sum_partials_block(const double* partials, int count, double* scratch) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * kBlock < count; i += 4 * kBlock) {
    a0 += partials[i];
    a1 += partials[i + kBlock];
    a2 += partials[i + 2 * kBlock];
    a3 += partials[i + 3 * kBlock];
  }
  for (; i < count; i += kBlock) a0 += partials[i];
  return block_sum((a0 + a1) + (a2 + a3), scratch);
}

__device__ __forceinline__ void check_norm(dds_pcg_state* st, double rr, double* norms) {
  st->rr = rr;
  const double res = sqrt(rr);
  norms[st->it] = res;
  if (res < st->atol) st->done = 1;
}

__global__ __launch_bounds__(kBlock) void finalize_kernel(int op, const double* __restrict__ partials,
                                                          int count, dds_pcg_state* st,
                                                          double* norms, double* alphas,
                                                          double* betas) {
  __shared__ double scratch[kBlock / 64];
  if (st->done) return;
  const double sum = sum_partials_block(partials, count, scratch);
  if (threadIdx.x != 0) return;
  switch (op) {
    case DDS_FIN_CHECK:
      check_norm(st, sum, norms);
      break;
    case DDS_FIN_RHO:
      st->rho = sum;
      if (st->it > 0) st->beta = sum / st->rho_prev;
      break;
    case DDS_FIN_ALPHA:
      st->pq = sum;
      st->alpha = st->rho / sum;
      alphas[st->it] = st->alpha;
      betas[st->it] = st->it > 0 ? st->beta : 0.0;
      break;
    case DDS_FIN_NEXT:
      st->rho_prev = st->rho;
      st->it += 1;
      if (st->it >= st->maxiter) {
        st->done = 2;
      } else {
        check_norm(st, sum, norms);
      }
      break;
  }
}

__global__ __launch_bounds__(kBlock) void sum_kernel(const double* __restrict__ partials, int count,
                                                     double* __restrict__ out) {
  __shared__ double scratch[kBlock / 64];
  const double sum = sum_partials_block(partials, count, scratch);
  if (threadIdx.x == 0) *out = sum;
}

// TY: the type of the local results Y (T, or float under FP64 vectors: the matrix-core apply
// accumulates in FP32 and writes half the bytes); the sums run in T.
template <typename T, typename TY>
__global__ __launch_bounds__(kBlock) void overlap_sum_kernel(
    int64_t n, const int64_t* __restrict__ dof_ptr, const int32_t* __restrict__ dof_pos,
    const TY* Y, const int32_t* __restrict__ cell, const T* __restrict__ yc, T sign, T* z,
    const double* __restrict__ r, double* __restrict__ partials,
    const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kBlock / 64];
  if (gated(gate)) return;
  double acc = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * kBlock) {
    T v;
    if (dof_ptr != nullptr) {
      v = 0;
      for (int64_t p = dof_ptr[j]; p < dof_ptr[j + 1]; ++p) v += static_cast<T>(Y[dof_pos[p]]);
    } else {
      v = static_cast<T>(Y[j]);
    }
    if (yc != nullptr) v += sign * yc[cell[j]];
    z[j] = v;
    if (r != nullptr) acc += r[j] * static_cast<double>(v);
  }
  if (partials != nullptr) write_partial(acc, partials, scratch);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void restrict_kernel(const int64_t* __restrict__ coarse_ptr,
                                                          const T* __restrict__ x,
                                                          T* __restrict__ xc,
                                                          const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kBlock / 64];
  if (gated(gate)) return;
  const int64_t s = coarse_ptr[blockIdx.x];
  const int64_t e = coarse_ptr[blockIdx.x + 1];
  double acc = 0.0;
  for (int64_t j = s + threadIdx.x; j < e; j += kBlock) acc += static_cast<double>(x[j]);
  acc = block_sum(acc, scratch);
  if (threadIdx.x == 0) xc[blockIdx.x] = static_cast<T>(acc);
}

// Small cells (the reference's DG coarse spaces: 3-48 unknowns per cell): G lanes per cell,
// grid-stride over cells; lane t of a group adds t, t + G, ... then a butterfly.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void restrict_group_kernel(int64_t n_c,
                                                                const int64_t* __restrict__ coarse_ptr,
                                                                const T* __restrict__ x,
                                                                T* __restrict__ xc,
                                                                const dds_pcg_state* __restrict__ gate) {
  if (gated(gate)) return;
  const int lg = threadIdx.x & (G - 1);
  const int64_t n_groups = (int64_t)gridDim.x * (kBlock / G);
  for (int64_t c = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G; c < n_c; c += n_groups) {
    const int64_t s = coarse_ptr[c];
    const int64_t e = coarse_ptr[c + 1];
    double acc = 0.0;
    for (int64_t j = s + lg; j < e; j += G) acc += static_cast<double>(x[j]);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    if (lg == 0) xc[c] = static_cast<T>(acc);
  }
}

template <typename T>
int launch_restrict(int64_t n_c, const int64_t* coarse_ptr, const void* x, void* xc, int lanes,
                    const dds_pcg_state* gate, hipStream_t s) {
  const T* xv = static_cast<const T*>(x);
  T* xcv = static_cast<T*>(xc);
  if (lanes == 0) {
    hipLaunchKernelGGL(restrict_kernel<T>, dim3((unsigned)n_c), dim3(kBlock), 0, s, coarse_ptr,
                       xv, xcv, gate);
  } else {
    const int64_t grid = std::min<int64_t>(ceil_div(n_c * lanes, kBlock), 1 << 20);
#define DDS_RESTRICT_G(G)                                                                       \
  hipLaunchKernelGGL((restrict_group_kernel<T, G>), dim3((unsigned)grid), dim3(kBlock), 0, s, \
                     n_c, coarse_ptr, xv, xcv, gate)
    switch (lanes) {
      case 8: DDS_RESTRICT_G(8); break;
      case 16: DDS_RESTRICT_G(16); break;
      case 32: DDS_RESTRICT_G(32); break;
      default: DDS_RESTRICT_G(64); break;
    }
#undef DDS_RESTRICT_G
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void cast_kernel(int64_t n, const S* __restrict__ src,
                                                      D* __restrict__ dst,
                                                      const dds_pcg_state* __restrict__ gate) {
  if (gated(gate)) return;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    dst[i] = static_cast<D>(src[i]);
}

}  // namespace
}  // namespace dds

using namespace dds;

#define DDS_PARTIALS_OUT(np, grid) \
  do {                             \
    if (np) *(np) = (int)(grid);   \
  } while (0)

extern "C" int dds_pcg_residual(int64_t n, const double* b, const double* q, double* r,
                                double* partials, int* n_partials, const dds_pcg_state* gate,
                                void* stream) {
  DDS_REQUIRE(n >= 0 && partials, "dds_pcg_residual: bad arguments");
  const int64_t g = vec_grid(n);
  hipLaunchKernelGGL(residual_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n,
                     b, q, r, partials, gate);
  DDS_LAUNCHED();
  DDS_PARTIALS_OUT(n_partials, g);
  return DDS_OK;
}

extern "C" int dds_pcg_update_p(int64_t n, const double* z, double* p, const dds_pcg_state* st,
                                void* stream) {
  DDS_REQUIRE(n >= 0 && st, "dds_pcg_update_p: bad arguments");
  const int64_t g = vec_grid(n);
  hipLaunchKernelGGL(update_p_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n,
                     z, p, st);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_pcg_update_xr(int64_t n, const double* p, const double* q, double* x,
                                 double* r, double* partials, int* n_partials,
                                 const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n >= 0 && st && partials, "dds_pcg_update_xr: bad arguments");
  const int64_t g = vec_grid(n);
  hipLaunchKernelGGL(update_xr_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n,
                     p, q, x, r, partials, st);
  DDS_LAUNCHED();
  DDS_PARTIALS_OUT(n_partials, g);
  return DDS_OK;
}

extern "C" int dds_dot_partials(int64_t n, const double* a, const double* b, double* partials,
                                int* n_partials, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n >= 0 && partials, "dds_dot_partials: bad arguments");
  const int64_t g = vec_grid(n);
  hipLaunchKernelGGL(dot_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n, a, b,
                     partials, gate);
  DDS_LAUNCHED();
  DDS_PARTIALS_OUT(n_partials, g);
  return DDS_OK;
}

extern "C" int dds_pcg_finalize(int op, const double* partials, int count, dds_pcg_state* st,
                                double* norms, double* alphas, double* betas, void* stream) {
  DDS_REQUIRE(op >= DDS_FIN_CHECK && op <= DDS_FIN_NEXT, "dds_pcg_finalize: bad op %d", op);
  DDS_REQUIRE(count >= 0 && count <= kMaxPartials, "dds_pcg_finalize: bad count %d", count);
  DDS_REQUIRE(st && norms && alphas && betas && (partials || count == 0),
              "dds_pcg_finalize: null pointer");
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kBlock), 0, as_stream(stream), op, partials,
                     count, st, norms, alphas, betas);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_sum_partials(const double* partials, int count, double* out, void* stream) {
  DDS_REQUIRE(count >= 0 && count <= kMaxPartials && out, "dds_sum_partials: bad arguments");
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kBlock), 0, as_stream(stream), partials, count,
                     out);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_overlap_sum(int dtype, int64_t n, const int64_t* dof_ptr,
                               const int32_t* dof_pos, const void* Y, int y_dtype,
                               const int32_t* cell,
                               const void* yc, double coarse_sign, void* z, const double* r,
                               double* dot_partials, int* n_partials,
                               const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n >= 0 && Y && z, "dds_overlap_sum: bad arguments");
  DDS_REQUIRE(!dof_ptr || dof_pos, "dds_overlap_sum: dof_ptr without dof_pos");
  DDS_REQUIRE(!yc || cell, "dds_overlap_sum: yc without cell");
  DDS_REQUIRE(!dot_partials || r, "dds_overlap_sum: partials need r");
  const int64_t g = vec_grid(n);
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64 && y_dtype == DDS_F64) {
    hipLaunchKernelGGL((overlap_sum_kernel<double, double>), dim3((unsigned)g), dim3(kBlock), 0, s,
                       n, dof_ptr, dof_pos, static_cast<const double*>(Y), cell,
                       static_cast<const double*>(yc), coarse_sign, static_cast<double*>(z), r,
                       dot_partials, gate);
  } else if (dtype == DDS_F64 && y_dtype == DDS_F32) {
    hipLaunchKernelGGL((overlap_sum_kernel<double, float>), dim3((unsigned)g), dim3(kBlock), 0, s,
                       n, dof_ptr, dof_pos, static_cast<const float*>(Y), cell,
                       static_cast<const double*>(yc), coarse_sign, static_cast<double*>(z), r,
                       dot_partials, gate);
  } else if (dtype == DDS_F32 && y_dtype == DDS_F32) {
    hipLaunchKernelGGL((overlap_sum_kernel<float, float>), dim3((unsigned)g), dim3(kBlock), 0, s,
                       n, dof_ptr, dof_pos, static_cast<const float*>(Y), cell,
                       static_cast<const float*>(yc), (float)coarse_sign,
                       static_cast<float*>(z), r, dot_partials, gate);
  } else {
    set_error("dds_overlap_sum: unsupported dtypes (vectors %d, local results %d)", dtype, y_dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  if (n_partials) *n_partials = dot_partials ? (int)g : 0;
  return DDS_OK;
}

extern "C" int dds_restrict(int dtype, int64_t n_c, const int64_t* coarse_ptr, const void* x,
                            void* xc, int lanes, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_c >= 0 && n_c <= INT32_MAX, "dds_restrict: bad size");
  DDS_REQUIRE(lanes == 0 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64,
              "dds_restrict: lanes must be 0, 8, 16, 32 or 64 (got %d)", lanes);
  if (n_c == 0) return DDS_OK;
  DDS_REQUIRE(coarse_ptr && x && xc, "dds_restrict: null pointer");
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64) return launch_restrict<double>(n_c, coarse_ptr, x, xc, lanes, gate, s);
  if (dtype == DDS_F32) return launch_restrict<float>(n_c, coarse_ptr, x, xc, lanes, gate, s);
  set_error("dds_restrict: unsupported dtype %d", dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_cast(int src_dtype, int dst_dtype, int64_t n, const void* src, void* dst,
                        const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n >= 0 && src && dst, "dds_cast: bad arguments");
  const int64_t g = vec_grid(n);
  hipStream_t s = as_stream(stream);
  if (src_dtype == DDS_F64 && dst_dtype == DDS_F32) {
    hipLaunchKernelGGL((cast_kernel<double, float>), dim3((unsigned)g), dim3(kBlock), 0, s, n,
                       static_cast<const double*>(src), static_cast<float*>(dst), gate);
  } else if (src_dtype == DDS_F32 && dst_dtype == DDS_F64) {
    hipLaunchKernelGGL((cast_kernel<float, double>), dim3((unsigned)g), dim3(kBlock), 0, s, n,
                       static_cast<const float*>(src), static_cast<double*>(dst), gate);
  } else if (src_dtype == dst_dtype && (src_dtype == DDS_F64 || src_dtype == DDS_F32)) {
    const size_t bytes = (size_t)n * (src_dtype == DDS_F64 ? 8 : 4);
    if (src != dst) DDS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    return DDS_OK;
  } else {
    set_error("dds_cast: unsupported dtypes %d -> %d", src_dtype, dst_dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}
