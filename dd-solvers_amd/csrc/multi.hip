// Multi-right-hand-side kernels: blocks of 16 systems solved together.
//
// Vectors are (n, 16) row-major FP64 (the 16 values of one unknown are contiguous, so a
// gather of one unknown is one 128-byte access). Each column has its own dds_pcg_state;
// update kernels skip converged columns, and every gated kernel returns at once when all
// 16 are done. Reductions write per-block partials laid out [block][16].
//
// The local-inverse apply is a batched GEMM Y_s = A_s^-1 X_s (k_s x k_s times k_s x 16) on
// MFMA: A_s^-1 in FP16 or BF16, X_s split into 2 (FP16) or 3 (BF16) low-precision terms
// after per-(subdomain, column) scaling, FP32 accumulation, result scaled back to FP64.
#include "common.hpp"

namespace dds {
namespace {

constexpr int M = 16;  // right-hand sides per block
constexpr int kBlock = 256;

typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool all_done(const dds_pcg_state* st) {
  if (st == nullptr) return false;
  bool done = true;
#pragma unroll
  for (int j = 0; j < M; ++j) done = done && st[j].done != 0;
  return done;
}

// Threads own column pairs: in a grid-stride loop over the (n, 16) array taken as n * 8
// pairs of doubles, thread t handles pairs t, t + stride, ... (stride a multiple of 8), so
// its pair is always t % 8 = threadIdx.x % 8 and a wavefront reads or writes 1 KB of
// contiguous memory per instruction.
constexpr int kPairs = M / 2;

__device__ __forceinline__ double2 ld2(const double* p) { return *reinterpret_cast<const double2*>(p); }
__device__ __forceinline__ void st2(double* p, double2 v) { *reinterpret_cast<double2*>(p) = v; }

// Block sums per column of the pair each thread owns (pair threadIdx.x % 8) into
// out[blockIdx.x * M + j]; fixed order, so deterministic.
__device__ __forceinline__ void write_partials_pairs(double2 v, double* out,
                                                     double* scratch /* (kBlock/64) * M */) {
#pragma unroll
  for (int off = kPairs; off < 64; off <<= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  if (lane < kPairs) {
    scratch[w * M + 2 * lane] = v.x;
    scratch[w * M + 2 * lane + 1] = v.y;
  }
  __syncthreads();
  if (threadIdx.x < M) {
    double s = 0.0;
    for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) s += scratch[ww * M + threadIdx.x];
    out[(int64_t)blockIdx.x * M + threadIdx.x] = s;
  }
}

// Sum over the 8 lanes groups of a wavefront that share a pair (lanes l, l ^ 8, ..., l ^ 56).
__device__ __forceinline__ double2 wave_pair_sum(double2 v) {
#pragma unroll
  for (int off = kPairs; off < 64; off <<= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  return v;
}

int64_t vgrid(int64_t rows) {
  return std::min<int64_t>(ceil_div(std::max<int64_t>(rows * kPairs, 1), kBlock), kVecGrid);
}

// ---------------------------------------------------------------------------------------
// PCG vector kernels
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void residual_m_kernel(int64_t n, const double* __restrict__ b,
                                                            const double* __restrict__ q,
                                                            double* __restrict__ r,
                                                            double* __restrict__ partials) {
  __shared__ double scratch[(kBlock / 64) * M];
  double2 acc = {0.0, 0.0};
  const int64_t total = n * kPairs;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
    const double2 bv = ld2(b + 2 * t), qv = ld2(q + 2 * t);
    const double2 v = {bv.x - qv.x, bv.y - qv.y};
    st2(r + 2 * t, v);
    acc.x += v.x * v.x;
    acc.y += v.y * v.y;
  }
  write_partials_pairs(acc, partials, scratch);
}

__global__ __launch_bounds__(kBlock) void update_p_m_kernel(int64_t n, const double* __restrict__ z,
                                                            double* __restrict__ p,
                                                            const dds_pcg_state* __restrict__ st) {
  if (all_done(st)) return;
  const int c = 2 * (threadIdx.x % kPairs);
  const bool a0 = st[c].done == 0, a1 = st[c + 1].done == 0;
  if (!a0 && !a1) return;
  const bool f0 = st[c].it == 0, f1 = st[c + 1].it == 0;
  const double b0 = st[c].beta, b1 = st[c + 1].beta;
  const bool need_p = !(a0 && a1 && f0 && f1);
  const int64_t total = n * kPairs;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
    const double2 zv = ld2(z + 2 * t);
    double2 pv = {0.0, 0.0};
    if (need_p) pv = ld2(p + 2 * t);
    double2 o;
    o.x = a0 ? (f0 ? zv.x : pv.x * b0 + zv.x) : pv.x;
    o.y = a1 ? (f1 ? zv.y : pv.y * b1 + zv.y) : pv.y;
    st2(p + 2 * t, o);
  }
}

__global__ __launch_bounds__(kBlock) void update_xr_m_kernel(int64_t n, const double* __restrict__ p,
                                                             const double* __restrict__ q,
                                                             double* __restrict__ x,
                                                             double* __restrict__ r,
                                                             double* __restrict__ partials,
                                                             const dds_pcg_state* __restrict__ st) {
  __shared__ double scratch[(kBlock / 64) * M];
  if (all_done(st)) return;
  const int c = 2 * (threadIdx.x % kPairs);
  const bool a0 = st[c].done == 0, a1 = st[c + 1].done == 0;
  const double al0 = a0 ? st[c].alpha : 0.0, al1 = a1 ? st[c + 1].alpha : 0.0;
  double2 acc = {0.0, 0.0};
  if (a0 || a1) {
    const int64_t total = n * kPairs;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
      const double2 pv = ld2(p + 2 * t), qv = ld2(q + 2 * t);
      double2 xv = ld2(x + 2 * t), rv = ld2(r + 2 * t);
      if (a0) {
        xv.x += al0 * pv.x;
        rv.x -= al0 * qv.x;
      }
      if (a1) {
        xv.y += al1 * pv.y;
        rv.y -= al1 * qv.y;
      }
      st2(x + 2 * t, xv);
      st2(r + 2 * t, rv);
      acc.x += rv.x * rv.x;
      acc.y += rv.y * rv.y;
    }
  }
  write_partials_pairs(acc, partials, scratch);
}

__global__ __launch_bounds__(kBlock) void dot_m_kernel(int64_t n, const double* __restrict__ a,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ partials,
                                                       const dds_pcg_state* __restrict__ st) {
  __shared__ double scratch[(kBlock / 64) * M];
  if (all_done(st)) return;
  double2 acc = {0.0, 0.0};
  const int64_t total = n * kPairs;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
    const double2 av = ld2(a + 2 * t), bv = ld2(b + 2 * t);
    acc.x += av.x * bv.x;
    acc.y += av.y * bv.y;
  }
  write_partials_pairs(acc, partials, scratch);
}

// Fixed-order sum of partials[i][j] per column j, then the scalar update of column j.
// Thread t reads partials[t + 1024 k] (coalesced), so it always sees column t % 16; the
// four lanes of a wave with one column are added by two butterflies, the 16 waves through
// LDS in wave order, and thread j finishes column j.
__global__ __launch_bounds__(M * 64) void finalize_m_kernel(int op, const double* __restrict__ partials,
                                                            int count, dds_pcg_state* st,
                                                            double* norms, double* alphas,
                                                            double* betas) {
  __shared__ double red[M][M];  // [wave][column]
  const int t = threadIdx.x;
  const int lane = t & 63;
  double acc = 0.0;
  const int64_t total = (int64_t)count * M;
  for (int64_t i = t; i < total; i += M * 64) acc += partials[i];
  acc += __shfl_xor(acc, 16, 64);
  acc += __shfl_xor(acc, 32, 64);
  if (lane < M) red[t >> 6][lane] = acc;
  __syncthreads();
  if (t >= M) return;
  const int j = t;
  dds_pcg_state* s = st + j;
  if (s->done) return;
  double sum = 0.0;
#pragma unroll
  for (int w = 0; w < M; ++w) sum += red[w][j];
  auto check = [&](double rr) {
    s->rr = rr;
    const double res = sqrt(rr);
    norms[(int64_t)s->it * M + j] = res;
    if (res < s->atol) s->done = 1;
  };
  switch (op) {
    case DDS_FIN_CHECK: check(sum); break;
    case DDS_FIN_RHO:
      s->rho = sum;
      if (s->it > 0) s->beta = sum / s->rho_prev;
      break;
    case DDS_FIN_ALPHA:
      s->pq = sum;
      s->alpha = s->rho / sum;
      alphas[(int64_t)s->it * M + j] = s->alpha;
      betas[(int64_t)s->it * M + j] = s->it > 0 ? s->beta : 0.0;
      break;
    case DDS_FIN_NEXT:
      s->rho_prev = s->rho;
      s->it += 1;
      if (s->it >= s->maxiter) s->done = 2;
      else check(sum);
      break;
  }
}

// ---------------------------------------------------------------------------------------
// Operators
// ---------------------------------------------------------------------------------------

// SELL-64 (slot pairs, as dds_sellmv) times 16 vectors; optional per-column partials of x . y. One wavefront per
// slice: lane l owns column pair l % 8 of the rows 8u + l / 8 (u < 8), so each gather of
// x reads 8 whole 128-byte rows per instruction.
constexpr int kSellWaves = 16;  // wavefronts per SpMM workgroup

__global__ __launch_bounds__(64 * kSellWaves) void sell_m_kernel(int64_t n_rows, int64_t n_slices,
                                                        const int64_t* __restrict__ slice_ptr,
                                                        const int32_t* __restrict__ col,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y,
                                                        double* __restrict__ partials,
                                                        const dds_pcg_state* __restrict__ st) {
  __shared__ double scratch[kSellWaves * M];
  if (all_done(st)) return;
  const int lane = threadIdx.x & 63;
  const int cp = lane & (kPairs - 1);
  const int rg = lane >> 3;
  const int64_t wave = (int64_t)blockIdx.x * kSellWaves + (threadIdx.x >> 6);
  const int64_t n_waves = (int64_t)gridDim.x * kSellWaves;
  double2 dot = {0.0, 0.0};
  for (int64_t q = wave; q < n_slices; q += n_waves) {
    const int64_t s0 = slice_ptr[q];
    const int w = (int)((slice_ptr[q + 1] - s0) >> 6);
    double2 acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = double2{0.0, 0.0};
    for (int j = 0; j < w; ++j) {
      // slot pairs: entry (j, r) at 128 (j / 2) + 2 r + j % 2; an odd last slot at 64 j + r
      const bool paired = j < (w & ~1);
      const int64_t o = paired ? s0 + 128 * (j >> 1) + (j & 1) : s0 + 64 * j;
      const int sr = paired ? 2 : 1;
      const double* vj = val + o;
      const int32_t* cj = col + o;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = 8 * u + rg;
        const double v = vj[sr * r];
        const double2 xv = ld2(x + (int64_t)cj[sr * r] * M + 2 * cp);
        acc[u].x += v * xv.x;
        acc[u].y += v * xv.y;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t row = q * 64 + 8 * u + rg;
      if (row < n_rows) {
        st2(y + row * M + 2 * cp, acc[u]);
        if (partials != nullptr) {
          const double2 xr = ld2(x + row * M + 2 * cp);
          dot.x += acc[u].x * xr.x;
          dot.y += acc[u].y * xr.y;
        }
      }
    }
  }
  if (partials != nullptr) write_partials_pairs(dot, partials, scratch);
}

// CSR (long rows, e.g. R0A = R_C A, thousands of entries per row) times 16 vectors: one
// workgroup of 1024 threads per row, 128 groups of 8 lanes over the row's entries, then a
// fixed-order reduction (butterflies within waves, waves in order through LDS).
constexpr int kRowThreads = 1024;

__global__ __launch_bounds__(kRowThreads) void csr_rows_m_kernel(int64_t n_rows, const int64_t* __restrict__ row_ptr,
                                                                 const int32_t* __restrict__ col,
                                                                 const double* __restrict__ val,
                                                                 const double* __restrict__ x,
                                                                 double* __restrict__ y,
                                                                 const dds_pcg_state* __restrict__ st) {
  __shared__ double scratch[(kRowThreads / 64) * M];
  if (all_done(st)) return;
  const int64_t row = blockIdx.x;
  const int cp = threadIdx.x & (kPairs - 1);
  const int g = threadIdx.x >> 3;
  double2 acc = {0.0, 0.0};
  for (int64_t p = row_ptr[row] + g; p < row_ptr[row + 1]; p += kRowThreads / kPairs) {
    const double v = val[p];
    const double2 xv = ld2(x + (int64_t)col[p] * M + 2 * cp);
    acc.x += v * xv.x;
    acc.y += v * xv.y;
  }
  write_partials_pairs(acc, y, scratch);  // block `row` writes y[row * M + j]
}

// Y = S X for a dense symmetric n x n S: one wavefront per row.
__global__ __launch_bounds__(kBlock) void dense_symm_m_kernel(int64_t n, const double* __restrict__ S,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ y,
                                                              const dds_pcg_state* __restrict__ st) {
  if (all_done(st)) return;
  const int lane = threadIdx.x & 63;
  const int cp = lane & (kPairs - 1);
  const int g = lane >> 3;
  const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (row >= n) return;
  double2 acc = {0.0, 0.0};
  for (int64_t c = g; c < n; c += 8) {
    const double s = S[row * n + c];
    const double2 xv = ld2(x + c * M + 2 * cp);
    acc.x += s * xv.x;
    acc.y += s * xv.y;
  }
  acc = wave_pair_sum(acc);
  if (lane < kPairs) st2(y + row * M + 2 * cp, acc);
}

// xc[c][:] = sum of x[j][:] over the unknowns j of coarse cell c (one block per cell).
__global__ __launch_bounds__(kBlock) void restrict_m_kernel(const int64_t* __restrict__ coarse_ptr,
                                                            const double* __restrict__ x,
                                                            double* __restrict__ xc,
                                                            const dds_pcg_state* __restrict__ st) {
  __shared__ double scratch[(kBlock / 64) * M];
  if (all_done(st)) return;
  const int64_t s = coarse_ptr[blockIdx.x];
  const int64_t e = coarse_ptr[blockIdx.x + 1];
  const int cp = threadIdx.x & (kPairs - 1);
  double2 acc = {0.0, 0.0};
  for (int64_t j = s + (threadIdx.x >> 3); j < e; j += kBlock / kPairs) {
    const double2 v = ld2(x + j * M + 2 * cp);
    acc.x += v.x;
    acc.y += v.y;
  }
  write_partials_pairs(acc, xc, scratch);  // block c writes xc[c * M + j]
}

// z[j][:] = sum_p Y[dof_pos[p]][:] (or Y[j][:]) + sign * yc[cell[j]][:], with per-column
// partials of r . z.
__global__ __launch_bounds__(kBlock) void overlap_sum_m_kernel(
    int64_t n, const int64_t* __restrict__ dof_ptr, const int32_t* __restrict__ dof_pos,
    const double* __restrict__ Y, const int32_t* __restrict__ cell, const double* __restrict__ yc,
    double sign, double* z, const double* __restrict__ r, double* __restrict__ partials,
    const dds_pcg_state* __restrict__ st) {
  __shared__ double scratch[(kBlock / 64) * M];
  if (all_done(st)) return;
  const int cp = threadIdx.x & (kPairs - 1);
  double2 acc = {0.0, 0.0};
  const int64_t total = n * kPairs;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
    const int64_t j = t >> 3;
    double2 v;
    if (dof_ptr != nullptr) {
      v = double2{0.0, 0.0};
      for (int64_t p = dof_ptr[j]; p < dof_ptr[j + 1]; ++p) {
        const double2 yv = ld2(Y + (int64_t)dof_pos[p] * M + 2 * cp);
        v.x += yv.x;
        v.y += yv.y;
      }
    } else {
      v = ld2(Y + 2 * t);
    }
    if (yc != nullptr) {
      const double2 c = ld2(yc + (int64_t)cell[j] * M + 2 * cp);
      v.x += sign * c.x;
      v.y += sign * c.y;
    }
    st2(z + 2 * t, v);
    if (r != nullptr) {
      const double2 rv = ld2(r + 2 * t);
      acc.x += rv.x * v.x;
      acc.y += rv.y * v.y;
    }
  }
  if (partials != nullptr) write_partials_pairs(acc, partials, scratch);
}

// ---------------------------------------------------------------------------------------
// MFMA local apply
// ---------------------------------------------------------------------------------------
// A store of subdomain s (k = k_s, RB = ceil(k/16) row blocks, KP = ceil(k/32) column-block
// pairs): for row block R and pair K2 the 64 lanes' 8 elements are at
//   a_off[s] + ((R * KP + K2) * 64 + l) * 8,
// lane l holding A[16R + l%16][32K2 + 4(l/16) + i] (i < 4) then the same 16 columns on
// (the A operand of one 16x16x32 MFMA). Padding is zero. Measured at Q1 128^3 (16 right-hand
// sides, FP16, k <= 1000), against this order: groups of 4 row blocks stored pair by pair (the
// 4 operands of a step 4 KiB contiguous) 7% slower in this kernel, although a bare stream in
// that order (no MFMAs) ran 13% faster than one in this order.
// X pack of subdomain s: for pair K2 and term t (< NT) at b_off[s] + ((K2 * NT + t) * 64 + l) * 8,
// lane l holding term t of X[32K2 + 4(l/16) + i][l%16] / sigma[s][l%16] (i < 4), then the
// same 16 rows on (the B operands).

template <typename TM>
struct MfmaTraits;
template <>
struct MfmaTraits<_Float16> {
  static constexpr int NT = 2;  // hi, lo
  typedef _Float16 elem;
  __device__ static __forceinline__ elem cvt(float v) { return static_cast<_Float16>(v); }
  __device__ static __forceinline__ float back(elem v) { return static_cast<float>(v); }
};
template <>
struct MfmaTraits<bf16_t> {
  static constexpr int NT = 3;  // hi, mid, lo
  typedef bf16_t elem;
  __device__ static __forceinline__ elem cvt(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return bf16_t{static_cast<uint16_t>(u >> 16)};
  }
  __device__ static __forceinline__ float back(elem v) { return widen(v); }
};

// One workgroup per subdomain. Pass 1: column maxima of X_s (threads own column pairs of
// the gathered rows). Pass 2: one thread per B-operand chunk (pair K2, lane l): it gathers
// the 8 values X[32K2 + 16t2 + 4(l/16) + i][l%16] (t2 < 2, i < 4), scales them, splits
// each into NT terms and writes NT 16-byte chunks.
constexpr int kPackThreads = 1024;

template <typename TM>
__global__ __launch_bounds__(kPackThreads) void pack_x_kernel(const int64_t* __restrict__ sub_ptr,
                                                        const int32_t* __restrict__ sub_idx,
                                                        const int64_t* __restrict__ b_off,
                                                        const double* __restrict__ x,
                                                        TM* __restrict__ bp,
                                                        float* __restrict__ sigma,
                                                        const dds_pcg_state* __restrict__ st) {
  using Tr = MfmaTraits<TM>;
  constexpr int NT = Tr::NT;
  __shared__ float smax[(kPackThreads / 64) * M];
  __shared__ float sig[M];
  if (all_done(st)) return;
  const int s = blockIdx.x;
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  auto row_of = [&](int a) -> int64_t { return sub_idx != nullptr ? (int64_t)sub_idx[base + a] : base + a; };
  const int cp = threadIdx.x & (kPairs - 1);
  float m0 = 0.f, m1 = 0.f;
  for (int e = threadIdx.x; e < k * kPairs; e += kPackThreads) {
    const double2 v = ld2(x + row_of(e >> 3) * M + 2 * cp);
    m0 = fmaxf(m0, fabsf((float)v.x));
    m1 = fmaxf(m1, fabsf((float)v.y));
  }
#pragma unroll
  for (int off = kPairs; off < 64; off <<= 1) {
    m0 = fmaxf(m0, __shfl_xor(m0, off, 64));
    m1 = fmaxf(m1, __shfl_xor(m1, off, 64));
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < kPairs) {
    smax[w * M + 2 * lane] = m0;
    smax[w * M + 2 * lane + 1] = m1;
  }
  __syncthreads();
  if (threadIdx.x < M) {
    float v = 0.f;
    for (int ww = 0; ww < kPackThreads / 64; ++ww) v = fmaxf(v, smax[ww * M + threadIdx.x]);
    v = v > 0.f ? v : 1.f;
    sig[threadIdx.x] = v;
    sigma[(int64_t)s * M + threadIdx.x] = v;
  }
  __syncthreads();
  const int KP = (k + 31) / 32;
  TM* out = bp + b_off[s];
  for (int ch = threadIdx.x; ch < KP * 64; ch += kPackThreads) {
    const int K2 = ch >> 6, l = ch & 63;
    const int c = l & 15, r4 = 4 * (l >> 4);
    const double inv = 1.0 / (double)sig[c];
    Vec16<TM> terms[NT];
#pragma unroll
    for (int e8 = 0; e8 < 8; ++e8) {
      const int a = 32 * K2 + 16 * (e8 >> 2) + r4 + (e8 & 3);
      double rem = a < k ? x[row_of(a) * M + c] * inv : 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const typename Tr::elem q = Tr::cvt((float)rem);
        rem -= (double)Tr::back(q);
        terms[t].v[e8] = q;
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
      *reinterpret_cast<Vec16<TM>*>(out + ((int64_t)(K2 * NT + t) * 64 + l) * 8) = terms[t];
  }
}

// One v_mfma_f32_16x16x32 per 16-byte A and B operand: k runs over the lane group's 8
// elements. Both operands hold the same columns (32K2 + 4(l/16) + i, then 16 on, i < 4) in
// the same order, so the product sums over the right pairs; the two 16x16x16 instructions
// this replaces summed the same 32 products in two steps.
template <typename TM>
__device__ __forceinline__ float4_t mfma32(const Vec16<TM>& a, const Vec16<TM>& b, float4_t c);
template <>
__device__ __forceinline__ float4_t mfma32<_Float16>(const Vec16<_Float16>& a, const Vec16<_Float16>& b,
                                                     float4_t c) {
  typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
  h8_t av, bv;
  __builtin_memcpy(&av, &a, 16);
  __builtin_memcpy(&bv, &b, 16);
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ float4_t mfma32<bf16_t>(const Vec16<bf16_t>& a, const Vec16<bf16_t>& b,
                                                   float4_t c) {
  typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
  b8_t av, bv;
  __builtin_memcpy(&av, &a, 16);
  __builtin_memcpy(&bv, &b, 16);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
}

// Operand register buffers of the MFMA apply. Measured at Q1 128^3 (16 right-hand sides,
// FP16, best launch shape each, one device): 1 buffer 1.532 ms, 2 buffers 1.603 ms, 3 buffers
// 1.632 ms: more loads in flight per wavefront slow this stream down.
#ifndef DDS_MFMA_DEPTH
#define DDS_MFMA_DEPTH 1
#endif
constexpr int kMfmaDepth = DDS_MFMA_DEPTH;

// Row blocks R0 .. R0 + RPW - 1 of subdomain s (one wavefront): for each column pair K2 in
// increasing order, the NT B operands (packed x terms, reused across the RPW row blocks) and
// the RPW A operands are loaded, then one MFMA per (row block, term). Every row block's
// accumulation is the same sequence of instructions whatever RPW and whichever wavefront,
// so all launch shapes give the same bits.
template <typename TM, int RPW>
__device__ __forceinline__ void mfma_rows(int s, int R0, int lane, const int64_t* __restrict__ sub_ptr,
                                          const int64_t* __restrict__ a_off,
                                          const int64_t* __restrict__ b_off,
                                          const float* __restrict__ scale,
                                          const float* __restrict__ sigma, const TM* __restrict__ A,
                                          const TM* __restrict__ bp, double* __restrict__ Y) {
  using Tr = MfmaTraits<TM>;
  constexpr int NT = Tr::NT;
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  const int RB = (k + 15) / 16;
  const int KP = (k + 31) / 32;
  const int nrb = RB - R0 < RPW ? RB - R0 : RPW;
  float4_t c[RPW][NT];
#pragma unroll
  for (int r = 0; r < RPW; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) c[r][t] = float4_t{0.f, 0.f, 0.f, 0.f};
  const TM* As = A + a_off[s];
  const TM* Bs = bp + b_off[s];
  // the operand of row block R0 + r, pair K2: Ar[r] + K2 * 512
  const TM* Ar[RPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) Ar[r] = As + ((int64_t)(R0 + r) * KP) * 512 + lane * 8;
  if (nrb == RPW) {
    // full group, software-pipelined over D register buffers: the operands of pair K2 + D - 1
    // are loaded before the MFMAs of pair K2, and scheduling barriers keep the compiler from
    // sinking those loads below the MFMAs (which would leave one or two loads in flight).
    // Every step issues its loads (past the last pair: the last pair again), so the count of
    // loads in flight is the same on every path and the compiler's waits before the MFMAs
    // cover exactly the operands they use.
    constexpr int D = kMfmaDepth;
    Vec16<TM> bv[D][NT], av[D][RPW];
    auto load = [&](int K2, int j) __attribute__((always_inline)) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
        bv[j][t] = load16(Bs + ((int64_t)(K2 * NT + t) * 64 + lane) * 8);
#pragma unroll
      for (int r = 0; r < RPW; ++r) av[j][r] = load16_nt(Ar[r] + (int64_t)K2 * 512);
    };
#pragma unroll
    for (int j = 0; j < D - 1; ++j) load(j < KP ? j : KP - 1, j);
    for (int K2 = 0; K2 < KP; K2 += D) {
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const int Kn = K2 + j + D - 1;
        load(Kn < KP ? Kn : KP - 1, (j + D - 1) % D);
        __builtin_amdgcn_sched_barrier(0);
        if (K2 + j < KP) {
#pragma unroll
          for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int t = 0; t < NT; ++t) c[r][t] = mfma32<TM>(av[j][r], bv[j][t], c[r][t]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  } else {
    for (int K2 = 0; K2 < KP; ++K2) {
      Vec16<TM> bv[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) bv[t] = load16(Bs + ((int64_t)(K2 * NT + t) * 64 + lane) * 8);
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        if (r < nrb) {
          const Vec16<TM> av = load16_nt(Ar[r] + (int64_t)K2 * 512);
#pragma unroll
          for (int t = 0; t < NT; ++t) c[r][t] = mfma32<TM>(av, bv[t], c[r][t]);
        }
      }
    }
  }
  // D tile: lane l holds rows 4(l/16) + i, column l % 16
  const int col = lane % 16;
  const double sg = (double)sigma[(int64_t)s * M + col] * (scale != nullptr ? (double)scale[s] : 1.0);
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    if (r < nrb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * (R0 + r) + 4 * (lane / 16) + i;
        double v = 0.0;
#pragma unroll
        for (int t = NT - 1; t >= 0; --t) v += (double)c[r][t][i];
        if (row < k) __builtin_nontemporal_store(v * sg, Y + (base + row) * M + col);
      }
    }
  }
}

// Persistent: work item i = (s, g) covers row blocks 4g .. 4g+3 of subdomain s; wave w of
// block b takes items b*W + w, b*W + w + gridDim*W, ...
template <typename TM>
__global__ __launch_bounds__(512) void mfma_apply_kernel(int64_t n_items, const int32_t* __restrict__ items,
                                  const int64_t* __restrict__ sub_ptr, const int64_t* __restrict__ a_off,
                                  const int64_t* __restrict__ b_off, const float* __restrict__ scale,
                                  const float* __restrict__ sigma, const TM* __restrict__ A,
                                  const TM* __restrict__ bp, double* __restrict__ Y,
                                  const dds_pcg_state* __restrict__ st) {
  if (all_done(st)) return;
  const int lane = threadIdx.x & 63;
  const int W = blockDim.x >> 6;
  const int64_t stride = (int64_t)gridDim.x * W;
  for (int64_t it = (int64_t)blockIdx.x * W + (threadIdx.x >> 6); it < n_items; it += stride)
    mfma_rows<TM, 4>(items[2 * it], 4 * items[2 * it + 1], lane, sub_ptr, a_off, b_off, scale,
                     sigma, A, bp, Y);
}

// One workgroup per subdomain (order[b]: longest first): its W wavefronts take the row-block
// groups g = w, w + W, ... of RPW row blocks each. The waves of a subdomain run on one CU
// together, so its packed x (B operands) comes from HBM once and the other waves' reads of
// it hit the XCD's L2, where the persistent item grid spreads a subdomain's groups over
// XCDs that each fetch it again.
template <typename TM, int RPW>
__global__ __launch_bounds__(512) void mfma_sub_kernel(const int32_t* __restrict__ order, const int64_t* __restrict__ sub_ptr,
                                const int64_t* __restrict__ a_off, const int64_t* __restrict__ b_off,
                                const float* __restrict__ scale, const float* __restrict__ sigma,
                                const TM* __restrict__ A, const TM* __restrict__ bp,
                                double* __restrict__ Y, const dds_pcg_state* __restrict__ st) {
  if (all_done(st)) return;
  const int lane = threadIdx.x & 63;
  const int W = blockDim.x >> 6;
  const int s = order[blockIdx.x];
  const int k = (int)(sub_ptr[s + 1] - sub_ptr[s]);
  const int NG = ((k + 15) / 16 + RPW - 1) / RPW;
  for (int g = threadIdx.x >> 6; g < NG; g += W)
    mfma_rows<TM, RPW>(s, RPW * g, lane, sub_ptr, a_off, b_off, scale, sigma, A, bp, Y);
}

}  // namespace
}  // namespace dds

using namespace dds;

#define DDS_GATE_M(st) const dds_pcg_state* gate_ = (st)

extern "C" int dds_pcg_residual_m(int64_t n, const double* b, const double* q, double* r,
                                  double* partials, int* n_partials, void* stream) {
  DDS_REQUIRE(n >= 0 && b && q && r && partials, "dds_pcg_residual_m: bad arguments");
  const int64_t g = vgrid(n);
  hipLaunchKernelGGL(residual_m_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n, b,
                     q, r, partials);
  DDS_LAUNCHED();
  if (n_partials) *n_partials = (int)g;
  return DDS_OK;
}

extern "C" int dds_pcg_update_p_m(int64_t n, const double* z, double* p,
                                  const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n >= 0 && z && p && st, "dds_pcg_update_p_m: bad arguments");
  hipLaunchKernelGGL(update_p_m_kernel, dim3((unsigned)vgrid(n)), dim3(kBlock), 0,
                     as_stream(stream), n, z, p, st);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_pcg_update_xr_m(int64_t n, const double* p, const double* q, double* x,
                                   double* r, double* partials, int* n_partials,
                                   const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n >= 0 && p && q && x && r && partials && st, "dds_pcg_update_xr_m: bad arguments");
  const int64_t g = vgrid(n);
  hipLaunchKernelGGL(update_xr_m_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n,
                     p, q, x, r, partials, st);
  DDS_LAUNCHED();
  if (n_partials) *n_partials = (int)g;
  return DDS_OK;
}

extern "C" int dds_dot_m(int64_t n, const double* a, const double* b, double* partials,
                         int* n_partials, const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n >= 0 && a && b && partials, "dds_dot_m: bad arguments");
  const int64_t g = vgrid(n);
  hipLaunchKernelGGL(dot_m_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n, a, b,
                     partials, st);
  DDS_LAUNCHED();
  if (n_partials) *n_partials = (int)g;
  return DDS_OK;
}

extern "C" int dds_pcg_finalize_m(int op, const double* partials, int count, dds_pcg_state* st,
                                  double* norms, double* alphas, double* betas, void* stream) {
  DDS_REQUIRE(op >= DDS_FIN_CHECK && op <= DDS_FIN_NEXT, "dds_pcg_finalize_m: bad op %d", op);
  DDS_REQUIRE(count >= 0 && count <= kMaxPartials, "dds_pcg_finalize_m: bad count %d", count);
  DDS_REQUIRE(st && norms && alphas && betas && (partials || count == 0),
              "dds_pcg_finalize_m: null pointer");
  hipLaunchKernelGGL(finalize_m_kernel, dim3(1), dim3(M * 64), 0, as_stream(stream), op, partials,
                     count, st, norms, alphas, betas);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_sellmm(int64_t n_rows, const int64_t* slice_ptr, const int32_t* col,
                          const double* values, const double* x, double* y, double* partials,
                          int* n_partials, const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n_rows >= 0 && slice_ptr && col && values && x && y, "dds_sellmm: bad arguments");
  const int64_t n_slices = ceil_div(n_rows, 64);
  const int64_t grid = std::min<int64_t>(kVecGrid, ceil_div(n_slices, kSellWaves));
  hipLaunchKernelGGL(sell_m_kernel, dim3((unsigned)std::max<int64_t>(grid, 1)), dim3(64 * kSellWaves), 0,
                     as_stream(stream), n_rows, n_slices, slice_ptr, col, values, x, y, partials, st);
  DDS_LAUNCHED();
  if (n_partials) *n_partials = partials ? (int)std::max<int64_t>(grid, 1) : 0;
  return DDS_OK;
}

extern "C" int dds_csrmm_rows(int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                              const double* values, const double* x, double* y,
                              const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n_rows >= 0 && row_ptr && col && values && x && y, "dds_csrmm_rows: bad arguments");
  if (n_rows == 0) return DDS_OK;
  DDS_REQUIRE(n_rows <= INT32_MAX, "dds_csrmm_rows: too many rows");
  hipLaunchKernelGGL(csr_rows_m_kernel, dim3((unsigned)n_rows), dim3(kRowThreads), 0,
                     as_stream(stream), n_rows, row_ptr, col, values, x, y, st);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_dense_symm_m(int64_t n, const double* S, const double* x, double* y,
                                const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n >= 0 && S && x && y, "dds_dense_symm_m: bad arguments");
  if (n == 0) return DDS_OK;
  hipLaunchKernelGGL(dense_symm_m_kernel, dim3((unsigned)ceil_div(n, kBlock / 64)), dim3(kBlock), 0,
                     as_stream(stream), n, S, x, y, st);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_restrict_m(int64_t n_c, const int64_t* coarse_ptr, const double* x, double* xc,
                              const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n_c >= 0 && n_c <= INT32_MAX && coarse_ptr && x && xc, "dds_restrict_m: bad arguments");
  if (n_c == 0) return DDS_OK;
  hipLaunchKernelGGL(restrict_m_kernel, dim3((unsigned)n_c), dim3(kBlock), 0, as_stream(stream),
                     coarse_ptr, x, xc, st);
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_overlap_sum_m(int64_t n, const int64_t* dof_ptr, const int32_t* dof_pos,
                                 const double* Y, const int32_t* cell, const double* yc,
                                 double sign, double* z, const double* r, double* partials,
                                 int* n_partials, const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n >= 0 && Y && z, "dds_overlap_sum_m: bad arguments");
  DDS_REQUIRE(!dof_ptr || dof_pos, "dds_overlap_sum_m: dof_ptr without dof_pos");
  DDS_REQUIRE(!yc || cell, "dds_overlap_sum_m: yc without cell");
  DDS_REQUIRE(!partials || r, "dds_overlap_sum_m: partials need r");
  const int64_t g = vgrid(n);
  hipLaunchKernelGGL(overlap_sum_m_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), n,
                     dof_ptr, dof_pos, Y, cell, yc, sign, z, r, partials, st);
  DDS_LAUNCHED();
  if (n_partials) *n_partials = partials ? (int)g : 0;
  return DDS_OK;
}

extern "C" int dds_mfma_pack_x(int mat_dtype, int64_t n_s, const int64_t* sub_ptr,
                               const int32_t* sub_idx, const int64_t* b_off, const double* x,
                               void* bp, float* sigma, const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_s <= INT32_MAX && sub_ptr && b_off && x && bp && sigma,
              "dds_mfma_pack_x: bad arguments");
  if (n_s == 0) return DDS_OK;
  hipStream_t s = as_stream(stream);
  if (mat_dtype == DDS_F16) {
    hipLaunchKernelGGL(pack_x_kernel<_Float16>, dim3((unsigned)n_s), dim3(kPackThreads), 0, s, sub_ptr,
                       sub_idx, b_off, x, static_cast<_Float16*>(bp), sigma, st);
  } else if (mat_dtype == DDS_BF16) {
    hipLaunchKernelGGL(pack_x_kernel<bf16_t>, dim3((unsigned)n_s), dim3(kPackThreads), 0, s, sub_ptr,
                       sub_idx, b_off, x, static_cast<bf16_t*>(bp), sigma, st);
  } else {
    set_error("dds_mfma_pack_x: storage must be F16 or BF16 (got %d)", mat_dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_mfma_apply(int mat_dtype, int64_t n_items, const int32_t* items,
                              const int64_t* sub_ptr, const int64_t* a_off, const int64_t* b_off,
                              const float* scale, const float* sigma, const void* A,
                              const void* bp, double* Y, int grid, int waves,
                              const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n_items >= 0 && items && sub_ptr && a_off && b_off && sigma && A && bp && Y,
              "dds_mfma_apply: bad arguments");
  DDS_REQUIRE(waves == 1 || waves == 2 || waves == 4 || waves == 8,
              "dds_mfma_apply: waves must be 1, 2, 4 or 8 (got %d)", waves);
  if (n_items == 0) return DDS_OK;
  const int64_t full = ceil_div(n_items, waves);
  const int64_t g = grid > 0 ? std::min<int64_t>(grid, full) : full;
  hipStream_t s = as_stream(stream);
  if (mat_dtype == DDS_F16) {
    hipLaunchKernelGGL(mfma_apply_kernel<_Float16>, dim3((unsigned)g), dim3(64 * waves), 0, s,
                       n_items, items, sub_ptr, a_off, b_off, scale, sigma,
                       static_cast<const _Float16*>(A), static_cast<const _Float16*>(bp), Y, st);
  } else if (mat_dtype == DDS_BF16) {
    hipLaunchKernelGGL(mfma_apply_kernel<bf16_t>, dim3((unsigned)g), dim3(64 * waves), 0, s,
                       n_items, items, sub_ptr, a_off, b_off, scale, sigma,
                       static_cast<const bf16_t*>(A), static_cast<const bf16_t*>(bp), Y, st);
  } else {
    set_error("dds_mfma_apply: storage must be F16 or BF16 (got %d)", mat_dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_mfma_apply_sub(int mat_dtype, int64_t n_s, const int32_t* order,
                                  const int64_t* sub_ptr, const int64_t* a_off,
                                  const int64_t* b_off, const float* scale, const float* sigma,
                                  const void* A, const void* bp, double* Y, int waves, int rpw,
                                  const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_s <= INT32_MAX && order && sub_ptr && a_off && b_off && sigma && A &&
                  bp && Y,
              "dds_mfma_apply_sub: bad arguments");
  DDS_REQUIRE(waves == 4 || waves == 8,
              "dds_mfma_apply_sub: waves must be 4 or 8 (got %d)", waves);
  DDS_REQUIRE(rpw == 2 || rpw == 4 || rpw == 8,
              "dds_mfma_apply_sub: rpw must be 2, 4 or 8 (got %d)", rpw);
  if (n_s == 0) return DDS_OK;
  hipStream_t s = as_stream(stream);
#define DDS_SUB(TM, R)                                                                     \
  hipLaunchKernelGGL((mfma_sub_kernel<TM, R>), dim3((unsigned)n_s), dim3(64 * waves), 0, s, \
                     order, sub_ptr, a_off, b_off, scale, sigma, static_cast<const TM*>(A),  \
                     static_cast<const TM*>(bp), Y, st)
  if (mat_dtype == DDS_F16) {
    if (rpw == 2) DDS_SUB(_Float16, 2);
    else if (rpw == 4) DDS_SUB(_Float16, 4);
    else DDS_SUB(_Float16, 8);
  } else if (mat_dtype == DDS_BF16) {
    if (rpw == 2) DDS_SUB(bf16_t, 2);
    else if (rpw == 4) DDS_SUB(bf16_t, 4);
    else DDS_SUB(bf16_t, 8);
  } else {
    set_error("dds_mfma_apply_sub: storage must be F16 or BF16 (got %d)", mat_dtype);
    return DDS_EUNSUPPORTED;
  }
#undef DDS_SUB
  DDS_LAUNCHED();
  return DDS_OK;
}
