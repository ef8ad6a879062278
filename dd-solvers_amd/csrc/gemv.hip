// Batched dense GEMV kernels: the drop-in strided-batched op and the ragged local-inverse
// apply used by the Schwarz preconditioners.
#include "common.hpp"

namespace dds {
namespace {

// ---------------------------------------------------------------------------------------
// Drop-in op: one output row per lane, columns in ascending order, accumulator of the
// vector type (the semantics of gemv_kernels.cu:31-93). Wave64 blocks; the LDS variant
// stages each batch's vector once per workgroup.
// ---------------------------------------------------------------------------------------
// The Schwarz preconditioners run the same kernels on their local inverses when the
// subdomains are the reference's (equal, contiguous, non-overlapping): `scale` (optional)
// multiplies output row block b by scale[b], and `gate` (optional) skips the launch once
// the PCG state says done. The drop-in op passes neither.
template <typename TM, typename TV>
__global__ __launch_bounds__(256) void gemv_rows_kernel(const TM* __restrict__ mat,
                                                        const TV* __restrict__ vec,
                                                        TV* __restrict__ out, int64_t n,
                                                        int64_t k, const float* __restrict__ scale,
                                                        const dds_pcg_state* __restrict__ gate) {
  if (gate != nullptr && gate->done) return;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (tid >= n * k) return;
  const int64_t b = tid / k;
  const int64_t r = tid - b * k;
  const TM* M = mat + b * k * k + r;
  const TV* v = vec + b * k;
  TV acc = 0;
  for (int64_t c = 0; c < k; ++c) acc += static_cast<TV>(widen(M[c * k])) * v[c];
  if (scale != nullptr) acc *= static_cast<TV>(scale[b]);
  out[b * k + r] = acc;
}

template <typename TM, typename TV>
__global__ __launch_bounds__(1024) void gemv_rows_lds_kernel(const TM* __restrict__ mat,
                                                             const TV* __restrict__ vec,
                                                             TV* __restrict__ out, int64_t n,
                                                             int k, int per_block,
                                                             const float* __restrict__ scale,
                                                             const dds_pcg_state* __restrict__ gate) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  if (gate != nullptr && gate->done) return;
  TV* vs = reinterpret_cast<TV*>(smem_raw);
  const int local = threadIdx.x / k;
  const int r = threadIdx.x - local * k;
  const int64_t b = (int64_t)blockIdx.x * per_block + local;
  const bool active = local < per_block && b < n;
  if (active) vs[local * k + r] = vec[b * k + r];
  __syncthreads();
  if (!active) return;
  const TM* M = mat + b * (int64_t)k * k + r;
  const TV* v = vs + local * k;
  TV acc = 0;
  for (int c = 0; c < k; ++c) acc += static_cast<TV>(widen(M[(int64_t)c * k])) * v[c];
  if (scale != nullptr) acc *= static_cast<TV>(scale[b]);
  out[b * k + r] = acc;
}

template <typename TM, typename TV>
int launch_gemv(const void* mat, const void* vec, void* out, int64_t n, int64_t k, int use_lds,
                hipStream_t s, const float* scale = nullptr,
                const dds_pcg_state* gate = nullptr) {
  if (n == 0 || k == 0) return DDS_OK;
  const TM* m = static_cast<const TM*>(mat);
  const TV* v = static_cast<const TV*>(vec);
  TV* o = static_cast<TV*>(out);
  if (use_lds && k <= 1024) {
    int block, per_block;
    if (k <= 256) {
      per_block = 256 / (int)k;
      block = 256;
    } else {
      per_block = 1;
      block = (int)ceil_div(k, 64) * 64;
    }
    const int64_t grid = ceil_div(n, per_block);
    const size_t shm = (size_t)per_block * k * sizeof(TV);
    hipLaunchKernelGGL((gemv_rows_lds_kernel<TM, TV>), dim3((unsigned)grid), dim3(block), shm,
                       s, m, v, o, n, (int)k, per_block, scale, gate);
  } else {
    const int64_t grid = ceil_div(n * k, 256);
    hipLaunchKernelGGL((gemv_rows_kernel<TM, TV>), dim3((unsigned)grid), dim3(256), 0, s, m, v,
                       o, n, k, scale, gate);
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

// ---------------------------------------------------------------------------------------
// Ragged local-inverse apply. One workgroup per (subdomain, tile of 64*V rows); lane l of
// every wave owns rows row0 + l*V .. row0 + l*V + V - 1 and reads them as one 16-byte
// vector per column (fully coalesced: a wave reads 1 KiB of one column per load). The
// WAVES waves of the workgroup split the columns round-robin; their partial sums are
// added in wave order through LDS. x_s is gathered (through sub_idx) into LDS once per
// workgroup, widened to the accumulator type.
// ---------------------------------------------------------------------------------------
template <typename TM, typename TV, typename TA, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void local_apply_kernel(
    const int32_t* __restrict__ tiles, const int64_t* __restrict__ sub_ptr,
    const int32_t* __restrict__ sub_idx, const int64_t* __restrict__ mat_off,
    const int32_t* __restrict__ ld, const float* __restrict__ scale,
    const TM* __restrict__ mats, const TV* __restrict__ x, TV* __restrict__ Y,
    const dds_pcg_state* __restrict__ gate) {
  if (gate != nullptr && gate->done) return;
  constexpr int V = Vec16<TM>::N;
  constexpr int ROWS = 64 * V;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TA* xs = reinterpret_cast<TA*>(smem_raw);

  const int s = tiles[2 * blockIdx.x];
  const int row0 = tiles[2 * blockIdx.x + 1];
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  const int ldm = ld[s];

  if (sub_idx != nullptr) {
    for (int c = threadIdx.x; c < k; c += WAVES * 64)
      xs[c] = static_cast<TA>(x[sub_idx[base + c]]);
  } else {
    for (int c = threadIdx.x; c < k; c += WAVES * 64) xs[c] = static_cast<TA>(x[base + c]);
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int r = row0 + lane * V;
  const bool active = r < k;
  TA acc[V];
#pragma unroll
  for (int j = 0; j < V; ++j) acc[j] = 0;

  if (active) {
    const TM* Ar = mats + mat_off[s] + r;
    int c = w;
    for (; c + 3 * WAVES < k; c += 4 * WAVES) {
      const Vec16<TM> a0 = load16_nt(Ar + (int64_t)c * ldm);
      const Vec16<TM> a1 = load16_nt(Ar + (int64_t)(c + WAVES) * ldm);
      const Vec16<TM> a2 = load16_nt(Ar + (int64_t)(c + 2 * WAVES) * ldm);
      const Vec16<TM> a3 = load16_nt(Ar + (int64_t)(c + 3 * WAVES) * ldm);
      const TA x0 = xs[c], x1 = xs[c + WAVES], x2 = xs[c + 2 * WAVES], x3 = xs[c + 3 * WAVES];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        acc[j] += static_cast<TA>(widen(a0.v[j])) * x0;
        acc[j] += static_cast<TA>(widen(a1.v[j])) * x1;
        acc[j] += static_cast<TA>(widen(a2.v[j])) * x2;
        acc[j] += static_cast<TA>(widen(a3.v[j])) * x3;
      }
    }
    for (; c < k; c += WAVES) {
      const Vec16<TM> a0 = load16_nt(Ar + (int64_t)c * ldm);
      const TA x0 = xs[c];
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] += static_cast<TA>(widen(a0.v[j])) * x0;
    }
  }

  if constexpr (WAVES > 1) {
    __syncthreads();  // xs is dead; reuse the LDS for the cross-wave sum
    TA* red = reinterpret_cast<TA*>(smem_raw);
    if (w > 0) {
#pragma unroll
      for (int j = 0; j < V; ++j) red[(w - 1) * ROWS + lane * V + j] = acc[j];
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int ww = 1; ww < WAVES; ++ww) {
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] += red[(ww - 1) * ROWS + lane * V + j];
    }
  }
  if (!active) return;
  const TA sc = scale != nullptr ? static_cast<TA>(scale[s]) : static_cast<TA>(1);
#pragma unroll
  for (int j = 0; j < V; ++j) {
    if (r + j < k) Y[base + r + j] = static_cast<TV>(acc[j] * sc);
  }
}

template <typename TM, typename TV, typename TA>
int launch_local(int64_t n_tiles, const int32_t* tiles, const int64_t* sub_ptr,
                 const int32_t* sub_idx, int64_t max_k, const int64_t* mat_off,
                 const int32_t* ld, const float* scale, const void* mats, const void* x,
                 void* Y, int waves, const dds_pcg_state* gate, hipStream_t s) {
  if (n_tiles == 0) return DDS_OK;
  constexpr int V = Vec16<TM>::N;
  constexpr int ROWS = 64 * V;
  if (waves <= 0) waves = 4;
  DDS_REQUIRE(waves == 1 || waves == 2 || waves == 4 || waves == 8,
              "dds_local_apply: waves must be 1, 2, 4 or 8 (got %d)", waves);
  const int64_t need = std::max<int64_t>(max_k, (int64_t)(waves - 1) * ROWS);
  const size_t shm = (size_t)need * sizeof(TA);
  DDS_REQUIRE(shm <= 64 * 1024, "dds_local_apply: max_k=%lld needs %zu bytes of LDS",
              (long long)max_k, shm);
  const TM* m = static_cast<const TM*>(mats);
  const TV* xv = static_cast<const TV*>(x);
  TV* yv = static_cast<TV*>(Y);
  const dim3 grid((unsigned)n_tiles);
#define DDS_LOCAL_LAUNCH(W)                                                                 \
  hipLaunchKernelGGL((local_apply_kernel<TM, TV, TA, W>), grid, dim3(W * 64), shm, s, tiles, \
                     sub_ptr, sub_idx, mat_off, ld, scale, m, xv, yv, gate)
  switch (waves) {
    case 1: DDS_LOCAL_LAUNCH(1); break;
    case 2: DDS_LOCAL_LAUNCH(2); break;
    case 4: DDS_LOCAL_LAUNCH(4); break;
    default: DDS_LOCAL_LAUNCH(8); break;
  }
#undef DDS_LOCAL_LAUNCH
  DDS_LAUNCHED();
  return DDS_OK;
}

int elem_size(int dt) {
  switch (dt) {
    case DDS_F64: return 8;
    case DDS_F32: return 4;
    case DDS_F16: return 2;
    case DDS_BF16: return 2;
    default: return 0;
  }
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_gemv_strided_batched(int mat_dtype, int vec_dtype, const void* mat,
                                        const void* vec, void* out, int64_t n, int64_t k,
                                        int use_lds, void* stream) {
  DDS_REQUIRE(n >= 0 && k >= 0, "gemv_strided_batched: negative size");
  DDS_REQUIRE(n * k <= (int64_t)INT32_MAX * 256, "gemv_strided_batched: n*k too large");
  if (n > 0 && k > 0) DDS_REQUIRE(mat && vec && out, "gemv_strided_batched: null pointer");
  hipStream_t s = as_stream(stream);
  if (vec_dtype == DDS_F64) {
    switch (mat_dtype) {
      case DDS_F64: return launch_gemv<double, double>(mat, vec, out, n, k, use_lds, s);
      case DDS_F32: return launch_gemv<float, double>(mat, vec, out, n, k, use_lds, s);
      case DDS_F16: return launch_gemv<_Float16, double>(mat, vec, out, n, k, use_lds, s);
      case DDS_BF16: return launch_gemv<bf16_t, double>(mat, vec, out, n, k, use_lds, s);
    }
  } else if (vec_dtype == DDS_F32) {
    switch (mat_dtype) {
      case DDS_F64:
        set_error("Double-Float mixed precision is not supported.");
        return DDS_EUNSUPPORTED;
      case DDS_F32: return launch_gemv<float, float>(mat, vec, out, n, k, use_lds, s);
      case DDS_F16: return launch_gemv<_Float16, float>(mat, vec, out, n, k, use_lds, s);
      case DDS_BF16: return launch_gemv<bf16_t, float>(mat, vec, out, n, k, use_lds, s);
    }
  }
  set_error("gemv_strided_batched: unsupported dtypes (mat %d, vec %d)", mat_dtype, vec_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_local_apply_batched(int mat_dtype, int vec_dtype, const void* mats,
                                       const void* x, void* y, int64_t n, int64_t k,
                                       int use_lds, const float* scale,
                                       const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n >= 0 && k >= 0, "dds_local_apply_batched: negative size");
  DDS_REQUIRE(n * k <= (int64_t)INT32_MAX * 256, "dds_local_apply_batched: n*k too large");
  DDS_REQUIRE(!use_lds || k <= 1024, "dds_local_apply_batched: LDS variant needs k <= 1024");
  if (n > 0 && k > 0) DDS_REQUIRE(mats && x && y, "dds_local_apply_batched: null pointer");
  hipStream_t s = as_stream(stream);
  if (vec_dtype == DDS_F64) {
    switch (mat_dtype) {
      case DDS_F64: return launch_gemv<double, double>(mats, x, y, n, k, use_lds, s, scale, gate);
      case DDS_F32: return launch_gemv<float, double>(mats, x, y, n, k, use_lds, s, scale, gate);
      case DDS_F16: return launch_gemv<_Float16, double>(mats, x, y, n, k, use_lds, s, scale, gate);
      case DDS_BF16: return launch_gemv<bf16_t, double>(mats, x, y, n, k, use_lds, s, scale, gate);
    }
  } else if (vec_dtype == DDS_F32) {
    switch (mat_dtype) {
      case DDS_F32: return launch_gemv<float, float>(mats, x, y, n, k, use_lds, s, scale, gate);
      case DDS_F16: return launch_gemv<_Float16, float>(mats, x, y, n, k, use_lds, s, scale, gate);
      case DDS_BF16: return launch_gemv<bf16_t, float>(mats, x, y, n, k, use_lds, s, scale, gate);
    }
  }
  set_error("dds_local_apply_batched: unsupported dtypes (mat %d, vec %d)", mat_dtype, vec_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_local_vec_width(int mat_dtype) {
  const int e = elem_size(mat_dtype);
  return e ? 16 / e : 0;
}

extern "C" int dds_local_tile_rows(int mat_dtype) { return 64 * dds_local_vec_width(mat_dtype); }

extern "C" int dds_local_apply(int mat_dtype, int vec_dtype, int acc_dtype, int64_t n_tiles,
                               const int32_t* tiles, const int64_t* sub_ptr,
                               const int32_t* sub_idx, int64_t max_k, const int64_t* mat_off,
                               const int32_t* ld, const float* scale, const void* mats,
                               const void* x, void* Y, int waves, const dds_pcg_state* gate,
                               void* stream) {
  DDS_REQUIRE(n_tiles >= 0 && max_k >= 0, "dds_local_apply: negative size");
  DDS_REQUIRE(n_tiles <= INT32_MAX, "dds_local_apply: too many tiles");
  if (n_tiles > 0)
    DDS_REQUIRE(tiles && sub_ptr && mat_off && ld && mats && x && Y,
                "dds_local_apply: null pointer");
  hipStream_t s = as_stream(stream);
#define DDS_ARGS n_tiles, tiles, sub_ptr, sub_idx, max_k, mat_off, ld, scale, mats, x, Y, waves, gate, s
  if (vec_dtype == DDS_F64 && acc_dtype == DDS_F64) {
    switch (mat_dtype) {
      case DDS_F64: return launch_local<double, double, double>(DDS_ARGS);
      case DDS_F32: return launch_local<float, double, double>(DDS_ARGS);
      case DDS_F16: return launch_local<_Float16, double, double>(DDS_ARGS);
      case DDS_BF16: return launch_local<bf16_t, double, double>(DDS_ARGS);
    }
  } else if (vec_dtype == DDS_F64 && acc_dtype == DDS_F32) {
    switch (mat_dtype) {
      case DDS_F32: return launch_local<float, double, float>(DDS_ARGS);
      case DDS_F16: return launch_local<_Float16, double, float>(DDS_ARGS);
      case DDS_BF16: return launch_local<bf16_t, double, float>(DDS_ARGS);
    }
  } else if (vec_dtype == DDS_F32 && acc_dtype == DDS_F32) {
    switch (mat_dtype) {
      case DDS_F32: return launch_local<float, float, float>(DDS_ARGS);
      case DDS_F16: return launch_local<_Float16, float, float>(DDS_ARGS);
      case DDS_BF16: return launch_local<bf16_t, float, float>(DDS_ARGS);
    }
  }
#undef DDS_ARGS
  set_error("dds_local_apply: unsupported dtypes (mat %d, vec %d, acc %d)", mat_dtype,
            vec_dtype, acc_dtype);
  return DDS_EUNSUPPORTED;
}
