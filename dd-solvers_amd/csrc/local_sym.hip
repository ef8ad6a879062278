// Local-inverse apply on symmetric, tile-packed storage: each subdomain inverse
// A_s^-1 (symmetric) is stored as its lower triangle of 64 x 64 tiles, so every
// off-diagonal tile is read from HBM once and used twice (y_I += T x_J and y_J += T^T x_I).
//
// Storage of subdomain s (k = k_s unknowns, T = ceil(k/64) tile rows, last block rL rows),
// in chunks of V = 16/sizeof elements (NCH = 64/V chunk columns per 64-wide tile):
//   off-diagonal tile (I, J), J < I, R_I rows (64, or rL for the last tile row): chunk j of
//     row r holds columns [jV, jV+V) at tile_off + (j*R_I + r)*V, so a wave loads one chunk
//     of all its rows as one coalesced 1 KiB (64 x 16 B) access;
//   diagonal tile (I, I): only its lower triangle, with the diagonal halved: chunk column j
//     holds rows r = jV .. R_I-1 at tile_off + V*(j*R_I - V*j*(j-1)/2) + (r - jV)*V, the
//     entries above the diagonal zero (ceil(rL/V) chunk columns in the last tile). Its row
//     product gives (L + D/2) x and its transposed product (L^T + D/2) x, which add to A x.
//   D64 = 2048 + 32 V elements (a full diagonal tile), and
//   tile_off(I, J) = 4096*I(I-1)/2 + I*D64 + 4096*J                       for I < T-1
//                  = 4096*(T-1)(T-2)/2 + (T-1)*D64 + 64*rL*J                for I = T-1
//   Padding entries are zero.
//
// Work item (s, p): one wavefront handles tile rows I1 = p and I2 = T-1-p (one row if they
// coincide), so every item touches T+1 tiles. It writes a partial result vector Yp[s][p]
// (length 64*T, blocks 0..I2 written); sym_combine_kernel adds the partial vectors of a
// subdomain in item order, which keeps the result deterministic without atomics.
#include "common.hpp"

namespace dds {
namespace {

template <typename TM>
struct SymShape {
  static constexpr int V = 16 / (int)sizeof(TM);  // elements per 16-byte chunk
  static constexpr int NCH = 64 / V;              // chunks per 64-wide tile row
  static constexpr int NH = NCH / 2;              // chunks per column half (32 columns)
  static constexpr int G = NH >= 16 ? 1 : 16 / NH;
  static constexpr int64_t D64 = 2048 + 32 * V;   // elements of a full diagonal tile
};

// LDS position (in chunks) of chunk j (of a column half) of tile row r: rows padded by one
// chunk, so the chunk writes (lane = row) and the column reads (lane = column) spread over
// the banks, and a lane's column read steps by a constant stride from row to row
// (immediate offsets, no per-row address arithmetic).
template <typename TM>
__device__ __forceinline__ int sym_slot(int r, int j) {
  using S = SymShape<TM>;
  return r * (S::NH + 1) + j;
}

// Workgroups are one wavefront, and the LDS operations of a wavefront are performed in
// order, so a wavefront-scope fence (a compiler barrier, no hardware wait) orders the
// lanes' LDS writes before other lanes' reads. A workgroup barrier would also wait for the
// global loads in flight, which are the prefetched tile halves.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Issues the NH chunk loads of column half h of the tile at tp (R rows, nch chunk
// columns; diag: packed lower-triangular layout, rows r >= jV of chunk j) and returns the
// lane's mask of chunks that hold data. Every load is issued on every path (lanes with
// nothing to load read the tile's first chunk), so the count of loads in flight is the same
// on all paths and the compiler waits for exactly the loads a computation needs; the
// unused values are zeroed where they are consumed, in compute_half.
template <typename TM>
__device__ __forceinline__ uint32_t load_half(const TM* tp, int R, int nch, bool diag, int h,
                                              bool dummy, int lane,
                                              Vec16<TM> (&c)[SymShape<TM>::NH]) {
  using S = SymShape<TM>;
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < S::NH; ++j) {
    const int jj = h * S::NH + j;
    const bool ok = !dummy && jj < nch && lane < R && (!diag || lane >= jj * S::V);
    const int64_t ub =
        diag ? (int64_t)S::V * ((int64_t)jj * R - (int64_t)S::V * jj * (jj - 1) / 2) -
                   (int64_t)jj * S::V * S::V
             : (int64_t)jj * R * S::V;
    c[j] = load16_nt(tp + (ok ? (int64_t)lane * S::V + ub : 0));
    m |= ok ? (1u << j) : 0u;
  }
  return m;
}

// Column half h (32 columns) of a tile, loaded by load_half with mask m (full: every chunk
// holds data):
//   row += sum_c T[lane][c] * xjs[c]                          (lane = row)
//   returns sum_r T[r][c] * xis[r] for column c = 32 h + lane % 32 (every lane)
// For the transposed product the half is staged in LDS, lanes 0-31 sum rows 0-31 and lanes
// 32-63 rows 32-63 of their column, and one cross-lane exchange adds the two row halves.
template <typename TM, typename TA>
__device__ __forceinline__ TA compute_half(Vec16<TM> (&c)[SymShape<TM>::NH], uint32_t m,
                                           bool full, int R, int nch, int h, const TA* xjs,
                                           const TA* xis, TM* lds, int lane, TA& row) {
  using S = SymShape<TM>;
  if (!full) {
#pragma unroll
    for (int j = 0; j < S::NH; ++j)
      if (!((m >> j) & 1u)) {
#pragma unroll
        for (int e = 0; e < S::V; ++e) c[j].v[e] = TM{};
      }
  }
  // four independent accumulators (shorter dependent FMA chains), added in a fixed order
  TA ra[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < S::NH; ++j) {
    const int jj = h * S::NH + j;
    if (jj < nch) {
#pragma unroll
      for (int e = 0; e < S::V; ++e)
        ra[(j * S::V + e) & 3] += static_cast<TA>(widen(c[j].v[e])) * xjs[jj * S::V + e];
    }
  }
  row += (ra[0] + ra[1]) + (ra[2] + ra[3]);
  wave_lds_fence();  // previous readers of lds are done
#pragma unroll
  for (int j = 0; j < S::NH; ++j)
    *reinterpret_cast<Vec16<TM>*>(lds + sym_slot<TM>(lane, j) * S::V) = c[j];
  wave_lds_fence();
  const int cc = lane & 31;
  const int jc = cc / S::V;
  const int ec = cc % S::V;
  const int r0 = lane & 32;
  const int nr = R - r0 < 32 ? R - r0 : 32;
  TA q0 = 0, q1 = 0, q2 = 0, q3 = 0;
  int t = 0;
  for (; t + 3 < nr; t += 4) {
    const TA a0 = static_cast<TA>(widen(lds[sym_slot<TM>(r0 + t, jc) * S::V + ec]));
    const TA a1 = static_cast<TA>(widen(lds[sym_slot<TM>(r0 + t + 1, jc) * S::V + ec]));
    const TA a2 = static_cast<TA>(widen(lds[sym_slot<TM>(r0 + t + 2, jc) * S::V + ec]));
    const TA a3 = static_cast<TA>(widen(lds[sym_slot<TM>(r0 + t + 3, jc) * S::V + ec]));
    q0 += a0 * xis[r0 + t];
    q1 += a1 * xis[r0 + t + 1];
    q2 += a2 * xis[r0 + t + 2];
    q3 += a3 * xis[r0 + t + 3];
  }
  for (; t < nr; ++t)
    q0 += static_cast<TA>(widen(lds[sym_slot<TM>(r0 + t, jc) * S::V + ec])) * xis[r0 + t];
  TA acc = (q0 + q1) + (q2 + q3);
  acc += __shfl_xor(acc, 32, 64);
  return acc;
}

template <typename TM, typename TV, typename TA>
__global__ __launch_bounds__(64) void sym_apply_kernel(
    const int32_t* __restrict__ items, const int64_t* __restrict__ sub_ptr,
    const int32_t* __restrict__ sub_idx, const int64_t* __restrict__ mat_off,
    const int64_t* __restrict__ ypos, const float* __restrict__ scale,
    const TM* __restrict__ mats, const TV* __restrict__ x, TV* __restrict__ Yp,
    const dds_pcg_state* __restrict__ gate) {
  using S = SymShape<TM>;
  if (gate != nullptr && gate->done) return;
  __shared__ __align__(16) TM lds_tile[64 * (32 + SymShape<TM>::V)];  // padded rows
  __shared__ __align__(16) TA xjs[64];
  __shared__ __align__(16) TA xi1s[64];
  __shared__ __align__(16) TA xi2s[64];

  const int lane = threadIdx.x;
  const int s = items[2 * blockIdx.x];
  const int p = items[2 * blockIdx.x + 1];
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  const int T = (k + 63) / 64;
  const int rL = k - 64 * (T - 1);
  const int I1 = p;
  const int I2 = T - 1 - p;
  const bool two = I1 != I2;
  const TM* M = mats + mat_off[s];
  TV* Y = Yp + ypos[s] + (int64_t)p * 64 * T;
  const TA sc = scale != nullptr ? static_cast<TA>(scale[s]) : static_cast<TA>(1);

  auto gather = [&](int J) -> TA {
    const int a = 64 * J + lane;
    if (a >= k) return static_cast<TA>(0);
    const int64_t g = sub_idx != nullptr ? (int64_t)sub_idx[base + a] : base + a;
    return static_cast<TA>(x[g]);
  };
  auto tile_off = [&](int I, int J) -> int64_t {
    if (I < T - 1) return 4096 * ((int64_t)I * (I - 1) / 2) + (int64_t)I * S::D64 + 4096 * (int64_t)J;
    return 4096 * ((int64_t)(T - 1) * (T - 2) / 2) + (int64_t)(T - 1) * S::D64 + (int64_t)64 * rL * J;
  };
  auto rows_of = [&](int I) { return I < T - 1 ? 64 : rL; };
  auto chunks_of = [&](int J) { return J < T - 1 ? S::NCH : (rL + S::V - 1) / S::V; };
  auto store = [&](int J, TA v) {
    if (64 * J + lane < k) Y[64 * J + lane] = static_cast<TV>(v * sc);
  };

  xi1s[lane] = gather(I1);
  xi2s[lane] = gather(I2);
  TA row1 = 0, row2 = 0, hold1 = 0;
  // Two-stage gather: the index of block J + 2 and the value of block J + 1 are in flight
  // while block J is processed. The index is kept as loaded (32 bits; the local position
  // when sub_idx is null): widening it right after the load would wait for every load in
  // flight.
  auto index_of = [&](int J) -> int32_t {
    const int a = 64 * J + lane;
    if (J > I2 || a >= k) return -1;
    return sub_idx != nullptr ? sub_idx[base + a] : a;
  };
  TA xj = gather(0);
  int32_t g_next = index_of(1);
  int gJ = 1;  // block whose index is in g_next
  auto start_block = [&]() {  // x of the block about to start into xjs; gather ahead
    wave_lds_fence();         // readers of xjs from the previous block are done
    xjs[lane] = xj;
    const int32_t g_cur = g_next;
    ++gJ;
    g_next = index_of(gJ);
    xj = g_cur >= 0 ? static_cast<TA>(x[sub_idx != nullptr ? (int64_t)g_cur : base + g_cur])
                    : static_cast<TA>(0);
    wave_lds_fence();
  };
  // Tasks, in order: for J = 0 .. I2, tile (I2, J), then tile (I1, J) if J <= I1, each in one
  // or two column halves. The loads of the next task are issued before the current one is
  // computed (two register buffers), so a wave always has a tile half in flight; after the
  // last task a dummy issue keeps the load count equal on every path.
  auto halves_of = [&](int J) { return chunks_of(J) > S::NH ? 2 : 1; };
  auto next_task = [&](int& J, int& w, int& h) -> bool {
    if (h + 1 < halves_of(J)) {
      ++h;
      return true;
    }
    h = 0;
    if (w == 0 && two && J <= I1) {
      w = 1;
      return true;
    }
    w = 0;
    ++J;
    return J <= I2;
  };
  auto issue = [&](int J, int w, int h, bool dummy, Vec16<TM> (&c)[S::NH]) -> uint32_t {
    const int I = w == 0 ? I2 : I1;
    return load_half<TM>(M + tile_off(I, J), rows_of(I), chunks_of(J), J == I, h, dummy, lane,
                         c);
  };
  int J = 0, w = 0, h = 0;  // current task: tile (w == 0 ? I2 : I1, J), column half h
  TA contrib = 0, tc0 = 0;
  Vec16<TM> cA[S::NH], cB[S::NH];
  uint32_t mA = issue(0, 0, 0, false, cA), mB = 0;
  start_block();
  auto step = [&](Vec16<TM> (&cur)[S::NH], uint32_t mc, Vec16<TM> (&nxt)[S::NH],
                  uint32_t& mn) -> bool {
    int J2 = J, w2 = w, h2 = h;
    const bool more = next_task(J2, w2, h2);
    mn = issue(more ? J2 : J, more ? w2 : w, more ? h2 : h, !more, nxt);
    const int I = w == 0 ? I2 : I1;
    const bool diag = J == I;
    const int R = rows_of(I), nch = chunks_of(J);
    const bool full = !diag && R == 64 && nch >= (h + 1) * S::NH;
    TA r = w == 0 ? row2 : row1;
    const TA col = compute_half<TM, TA>(cur, mc, full, R, nch, h, xjs, w == 0 ? xi2s : xi1s,
                                        lds_tile, lane, r);
    if (h == 0) tc0 = col;
    if (h + 1 == halves_of(J)) {  // tile complete: column lane of block J is tv
      const TA tv = lane < 32 ? tc0 : (h == 0 ? static_cast<TA>(0) : col);
      if (diag)  // (L^T + D/2) x_I belongs to the tile's own row block
        r += tv;
      else
        contrib += tv;
    }
    if (w == 0)
      row2 = r;
    else
      row1 = r;
    if (!more || J2 != J) {  // block J is complete
      if (J < I2) {
        if (two && J == I1)
          hold1 = contrib;  // block I1 also receives row1, stored at the end
        else
          store(J, contrib);
      }
      contrib = 0;
      if (more) start_block();
    }
    J = J2;
    w = w2;
    h = h2;
    return more;
  };
  while (step(cA, mA, cB, mB) && step(cB, mB, cA, mA)) {
  }
  if (two) store(I1, row1 + hold1);
  store(I2, row2);
}

// y[sub_ptr[s] + a] = sum over the items p of subdomain s that cover local row a (blocks
// 0..T-1-p) of Yp[ypos[s] + p*64*T + a], in item order. Grid (n_s, ceil(max_k/256)):
// subdomains on x, which has no 65535 limit.
template <typename TV>
__global__ __launch_bounds__(256) void sym_combine_kernel(const int64_t* __restrict__ sub_ptr,
                                                          const int64_t* __restrict__ ypos,
                                                          const TV* __restrict__ Yp,
                                                          TV* __restrict__ y,
                                                          const dds_pcg_state* __restrict__ gate) {
  if (gate != nullptr && gate->done) return;
  const int s = blockIdx.x;
  const int a = blockIdx.y * 256 + threadIdx.x;
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  if (a >= k) return;
  const int T = (k + 63) / 64;
  const int P = (T + 1) / 2;
  const int last = min(P - 1, T - 1 - a / 64);
  const TV* src = Yp + ypos[s] + a;
  TV acc = 0;
  for (int p = 0; p <= last; ++p) acc += src[(int64_t)p * 64 * T];
  y[base + a] = acc;
}

template <typename TM, typename TV, typename TA>
int launch_sym(int64_t n_items, const int32_t* items, const int64_t* sub_ptr,
               const int32_t* sub_idx, const int64_t* mat_off, const int64_t* ypos,
               const float* scale, const void* mats, const void* x, void* Yp,
               const dds_pcg_state* gate, hipStream_t s) {
  if (n_items == 0) return DDS_OK;
  hipLaunchKernelGGL((sym_apply_kernel<TM, TV, TA>), dim3((unsigned)n_items), dim3(64), 0, s,
                     items, sub_ptr, sub_idx, mat_off, ypos, scale,
                     static_cast<const TM*>(mats), static_cast<const TV*>(x),
                     static_cast<TV*>(Yp), gate);
  DDS_LAUNCHED();
  return DDS_OK;
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_local_apply_sym(int mat_dtype, int vec_dtype, int acc_dtype, int64_t n_items,
                                   const int32_t* items, const int64_t* sub_ptr,
                                   const int32_t* sub_idx, const int64_t* mat_off,
                                   const int64_t* ypos, const float* scale, const void* mats,
                                   const void* x, void* Yp, const dds_pcg_state* gate,
                                   void* stream) {
  DDS_REQUIRE(n_items >= 0 && n_items <= INT32_MAX, "dds_local_apply_sym: bad item count");
  if (n_items > 0)
    DDS_REQUIRE(items && sub_ptr && mat_off && ypos && mats && x && Yp,
                "dds_local_apply_sym: null pointer");
  hipStream_t s = as_stream(stream);
#define DDS_ARGS n_items, items, sub_ptr, sub_idx, mat_off, ypos, scale, mats, x, Yp, gate, s
  if (vec_dtype == DDS_F64 && acc_dtype == DDS_F64) {
    switch (mat_dtype) {
      case DDS_F64: return launch_sym<double, double, double>(DDS_ARGS);
      case DDS_F32: return launch_sym<float, double, double>(DDS_ARGS);
      case DDS_F16: return launch_sym<_Float16, double, double>(DDS_ARGS);
      case DDS_BF16: return launch_sym<bf16_t, double, double>(DDS_ARGS);
    }
  } else if (vec_dtype == DDS_F64 && acc_dtype == DDS_F32) {
    switch (mat_dtype) {
      case DDS_F32: return launch_sym<float, double, float>(DDS_ARGS);
      case DDS_F16: return launch_sym<_Float16, double, float>(DDS_ARGS);
      case DDS_BF16: return launch_sym<bf16_t, double, float>(DDS_ARGS);
    }
  } else if (vec_dtype == DDS_F32 && acc_dtype == DDS_F32) {
    switch (mat_dtype) {
      case DDS_F32: return launch_sym<float, float, float>(DDS_ARGS);
      case DDS_F16: return launch_sym<_Float16, float, float>(DDS_ARGS);
      case DDS_BF16: return launch_sym<bf16_t, float, float>(DDS_ARGS);
    }
  }
#undef DDS_ARGS
  set_error("dds_local_apply_sym: unsupported dtypes (mat %d, vec %d, acc %d)", mat_dtype,
            vec_dtype, acc_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_sym_combine(int dtype, int64_t n_s, int64_t max_k, const int64_t* sub_ptr,
                               const int64_t* ypos, const void* Yp, void* y,
                               const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_s <= INT32_MAX && max_k >= 0 && ceil_div(max_k, 256) <= 65535,
              "dds_sym_combine: bad sizes");
  if (n_s == 0 || max_k == 0) return DDS_OK;
  DDS_REQUIRE(sub_ptr && ypos && Yp && y, "dds_sym_combine: null pointer");
  const dim3 grid((unsigned)n_s, (unsigned)ceil_div(max_k, 256));
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(sym_combine_kernel<double>, grid, dim3(256), 0, s, sub_ptr, ypos,
                       static_cast<const double*>(Yp), static_cast<double*>(y), gate);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(sym_combine_kernel<float>, grid, dim3(256), 0, s, sub_ptr, ypos,
                       static_cast<const float*>(Yp), static_cast<float*>(y), gate);
  } else {
    set_error("dds_sym_combine: unsupported dtype %d", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}
