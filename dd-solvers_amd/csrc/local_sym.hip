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

#include <mutex>

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

// y[sub_ptr[s] + a] = sum over the items q of subdomain s that cover local row a (item q:
// pairs q G .. q G + G - 1, blocks 0..T-1-qG) of Yp[ypos[s] + q*64*T + a], in item order.
// Grid (n_s, ceil(max_k/256)): subdomains on x, which has no 65535 limit.
template <typename TV>
__global__ __launch_bounds__(256) void sym_combine_kernel(int G, const int64_t* __restrict__ sub_ptr,
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
  const int Q = ((T + 1) / 2 + G - 1) / G;
  const int last = min(Q - 1, (T - 1 - a / 64) / G);
  const TV* src = Yp + ypos[s] + a;
  TV acc = 0;
  for (int q = 0; q <= last; ++q) acc += src[(int64_t)q * 64 * T];
  y[base + a] = acc;
}

// ---------------------------------------------------------------------------------------
// Symmetric apply on the matrix cores (16-bit storage, FP32 accumulation).
//
// Storage of subdomain s: the same lower triangle of 64 x 64 tiles (tile rows I, tiles
// J = 0..I in order), each tile kept as its non-zero 16 x 32 blocks (Rb = 0..3 row blocks,
// K2 = 0..1 column halves) of 512 elements: lane l of block (Rb, K2) holds the 8 elements
//   T[16 Rb + l%16][32 K2 + 8(l/16) + e],  e < 8,
// the A operand of a v_mfma_f32_16x16x32 (row m = l%16, k = 8(l/16) + e). Off-diagonal
// tiles keep blocks q = 2 Rb + K2 for Rb < nb (nb = ceil(R/16) for a tile row of R rows);
// diagonal tiles keep L + D/2 of the tile (zero above the diagonal), without the two zero
// blocks (Rb < 2, K2 = 1) and, on the last tile, without blocks past its rL columns:
//   block index (Rb, K2) = Rb + (nk == 2 ? max(Rb - 2, 0) : 0) + K2,  nk = ceil(C/32).
//   tile_off(I, J) = 512 (8 I(I-1)/2 + 6 I + 8 J)                          for I < T-1
//                  = 512 (8 (T-1)(T-2)/2 + 6 (T-1) + 2 nb(rL) J)            for I = T-1
// Padding is zero.
//
// x is packed once per apply (sym_pack_x_kernel): each 64-row block J of a subdomain's x is
// taken relative to sig_J = 2^e > max |x| over the block and split into NT 16-bit terms
// (FP16: hi + lo; BF16: hi + mid + lo), which keeps about 22 (FP16) or 24 (BF16) bits of
// x / sig_J; the apply loads the terms and scales of its blocks into LDS with its first tiles.
//   Row product y_I += T x_J: D = A(T block) x B, B column n < NT holding term n of x_J / sig_J;
//   y_I += sig_J D.
//   Transposed product y_J += T^T x_I: D = A' x B', A' row m < NT holding term m of
//   x_I / sig_I, B' the tile block in column-major order, read back from the LDS copy of the
//   tile with ds_read_tr16_b64 (in each group of 16 lanes, lane (r, q) names 4 consecutive
//   elements of row r; lane c receives element c of the 4 rows); y_J += sig_I D.
// The products accumulate in FP32 (MFMA per tile, then the scaled sums); the NT term columns
// (rows) are added, and scale[s] applied when the result is written.
// ---------------------------------------------------------------------------------------
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef short i16x4_t __attribute__((__vector_size__(4 * sizeof(short))));
typedef short i16x8_t __attribute__((ext_vector_type(8)));

template <typename TM>
struct SymMfma;
template <>
struct SymMfma<_Float16> {
  static constexpr int NT = 2;
  typedef f16x8_t op_t;
  __device__ static __forceinline__ f32x4_t mfma(op_t a, op_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  // term t of v: the t-th 16-bit piece of the split v = hi + lo
  __device__ static __forceinline__ uint16_t term(float v, int t) {
    const _Float16 hi = static_cast<_Float16>(v);
    const _Float16 q = t == 0 ? hi : static_cast<_Float16>(v - static_cast<float>(hi));
    uint16_t u;
    __builtin_memcpy(&u, &q, 2);
    return u;
  }
};
template <>
struct SymMfma<bf16_t> {
  static constexpr int NT = 3;
  typedef bf16x8_t op_t;
  __device__ static __forceinline__ f32x4_t mfma(op_t a, op_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ uint32_t rne(float v) {  // bf16 bits of v, widened
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u & 0xFFFF0000u;
  }
  __device__ static __forceinline__ uint16_t term(float v, int t) {
    const uint32_t hi = rne(v);
    if (t == 0) return static_cast<uint16_t>(hi >> 16);
    const float r1 = v - __uint_as_float(hi);
    const uint32_t mid = rne(r1);
    if (t == 1) return static_cast<uint16_t>(mid >> 16);
    return static_cast<uint16_t>(rne(r1 - __uint_as_float(mid)) >> 16);
  }
};

template <typename TM>
__device__ __forceinline__ typename SymMfma<TM>::op_t as_op(const Vec16<TM>& v) {
  typename SymMfma<TM>::op_t o;
  __builtin_memcpy(&o, &v, 16);
  return o;
}

// Pack of x, once per apply, per 64-row block b (blk_owner[b] = s, the block's index in its
// subdomain J = b - blk_ptr[s]; lane = row 64 J + lane): sig[b] = 2^e with max |x| over the
// block's rows < 2^e (1 when they are zero; e clamped to the normal float range) and
// x / sig[b] split into NT 16-bit terms at xt[b NT 64 + (2 (lane/32) NT + t) 32 + lane % 32]
// (rows past k zero). A wavefront packs kPackBlocks consecutive blocks; four wavefronts per
// workgroup.
constexpr int kPackBlocks = 8;

// The kPackBlocks blocks from b0 of one wavefront (IDX: rows gathered through sub_idx). Every
// load is unconditional, at clamped in-range indices (a block's subdomain has k >= 1), so
// the blocks' dependent loads are in flight together; out-of-range rows are selected away.
template <typename TM, typename TV, bool IDX>
__device__ __forceinline__ void sym_pack_blocks(int64_t b0, int64_t n_blk, int lane,
                                                const int32_t* __restrict__ blk_owner,
                                                const int64_t* __restrict__ blk_ptr,
                                                const int64_t* __restrict__ sub_ptr,
                                                const int32_t* __restrict__ sub_idx,
                                                const TV* __restrict__ x, float* __restrict__ sig,
                                                uint16_t* __restrict__ xt) {
  using S = SymMfma<TM>;
  constexpr int NT = S::NT;
  int own[kPackBlocks];
#pragma unroll
  for (int u = 0; u < kPackBlocks; ++u)
    own[u] = __builtin_amdgcn_readfirstlane(blk_owner[b0 + u < n_blk ? b0 + u : n_blk - 1]);
  int64_t row[kPackBlocks];
  int ok[kPackBlocks];
#pragma unroll
  for (int u = 0; u < kPackBlocks; ++u) {
    const int64_t b = b0 + u < n_blk ? b0 + u : n_blk - 1;
    const int64_t base = sub_ptr[own[u]];
    const int k = (int)(sub_ptr[own[u] + 1] - base);
    const int a = 64 * (int)(b - blk_ptr[own[u]]) + lane;
    ok[u] = a < k;
    row[u] = base + (a < k ? a : k - 1);
  }
  if constexpr (IDX) {
#pragma unroll
    for (int u = 0; u < kPackBlocks; ++u) row[u] = sub_idx[row[u]];
  }
  double v[kPackBlocks];
#pragma unroll
  for (int u = 0; u < kPackBlocks; ++u) {
    const double xv = static_cast<double>(x[row[u]]);
    v[u] = ok[u] ? xv : 0.0;
  }
#pragma unroll
  for (int u = 0; u < kPackBlocks; ++u) {
    const int64_t b = b0 + u;
    if (b < n_blk) {
      double m = fabs(v[u]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
      int e = 0;
      if (m > 0.0) frexp(m, &e);
      e = e < -126 ? -126 : (e > 127 ? 127 : e);
      const float vs = static_cast<float>(v[u] * ldexp(1.0, -e));  // exact scaling
      uint16_t* out = xt + b * NT * 64;
#pragma unroll
      for (int t = 0; t < NT; ++t) out[((lane >> 5) * NT + t) * 32 + (lane & 31)] = S::term(vs, t);
      if (lane == 0) sig[b] = ldexpf(1.0f, e);
    }
  }
}

template <typename TM, typename TV>
__global__ __launch_bounds__(256) void sym_pack_x_kernel(int64_t n_blk,
                                                         const int32_t* __restrict__ blk_owner,
                                                         const int64_t* __restrict__ blk_ptr,
                                                         const int64_t* __restrict__ sub_ptr,
                                                         const int32_t* __restrict__ sub_idx,
                                                         const TV* __restrict__ x,
                                                         float* __restrict__ sig,
                                                         uint16_t* __restrict__ xt,
                                                         const dds_pcg_state* __restrict__ gate) {
  if (gate != nullptr && gate->done) return;
  const int64_t b0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kPackBlocks;
  if (b0 >= n_blk) return;
  const int lane = threadIdx.x & 63;
  if (sub_idx != nullptr)
    sym_pack_blocks<TM, TV, true>(b0, n_blk, lane, blk_owner, blk_ptr, sub_ptr, sub_idx, x, sig, xt);
  else
    sym_pack_blocks<TM, TV, false>(b0, n_blk, lane, blk_owner, blk_ptr, sub_ptr, sub_idx, x, sig, xt);
}

constexpr uint32_t kOob = 0x80000000u;  // a buffer offset past every range: loads zeros
// Largest subdomain of the matrix-core apply: its result (64 floats per 64-row block) and
// block scales sit in LDS beside the tile stages.
constexpr int kSymMfmaMaxK = 4096;

__host__ __device__ inline int64_t symm_store_elems(int k) {  // elements of one subdomain
  const int T = (k + 63) / 64, rL = k - 64 * (T - 1);
  const int nbL = (rL + 15) / 16, nkL = (rL + 31) / 32;
  int ndl = 0;
  for (int Rb = 0; Rb < nbL; ++Rb)
    for (int K2 = 0; K2 < nkL; ++K2) ndl += (K2 == 1 && Rb < 2) ? 0 : 1;
  return 512 * (8 * ((int64_t)(T - 1) * (T - 2) / 2) + 6 * (int64_t)(T - 1) +
                2 * (int64_t)nbL * (T - 1) + ndl);
}

// Element offset of tile (I, J) in a subdomain's matrix-core symmetric store.
__device__ __forceinline__ int64_t symm_tile_off(int I, int J, int T, int nbL) {
  if (I < T - 1) return 512 * (8 * ((int64_t)I * (I - 1) / 2) + 6 * (int64_t)I + 8 * (int64_t)J);
  return 512 * (8 * ((int64_t)(T - 1) * (T - 2) / 2) + 6 * (int64_t)(T - 1) + 2 * (int64_t)nbL * J);
}

// LDS arrangement of a 1 KiB block: slot s (16 bytes) holds chunk s ^ (4 (s/16 mod 2)) of the
// block (chunk l = lane l's operand), so that both the row-product reads (lane l reads its
// chunk) and the transposed reads (16 chunks r + 4h + 8a + 16b + 32n at a time) cover 16
// distinct 16-byte bank groups. The map is its own inverse.
__device__ __forceinline__ int symm_slot(int c) { return c ^ (((c >> 4) & 1) << 2); }

// The x terms and scales of blocks 0 .. nb-1 of a subdomain (k unknowns from base, gathered
// through sub_idx when it is given) computed in place into LDS, in the layout and with the
// arithmetic of sym_pack_blocks (the same bits): per block, the power-of-two scale above its
// largest magnitude, and the NT 16-bit terms of each scaled value.
template <typename TM, typename TX>
__device__ __forceinline__ void symm_pack_lds(int nb, int k, int64_t base,
                                              const int32_t* __restrict__ sub_idx,
                                              const TX* __restrict__ x, int lane,
                                              uint16_t* xtl, float* sig) {
  using S = SymMfma<TM>;
  constexpr int NT = S::NT;
  constexpr int U = 8;  // blocks whose loads are in flight together
  for (int b0 = 0; b0 < nb; b0 += U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int a = 64 * (b0 + u) + lane;
      const bool ok = b0 + u < nb && a < k;
      int64_t row = base + (ok ? a : 0);
      if (sub_idx != nullptr) row = sub_idx[row];
      const double xv = static_cast<double>(x[row]);
      v[u] = ok ? xv : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u;
      if (b < nb) {
        double m = fabs(v[u]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
        int e = 0;
        if (m > 0.0) frexp(m, &e);
        e = e < -126 ? -126 : (e > 127 ? 127 : e);
        const float vs = static_cast<float>(v[u] * ldexp(1.0, -e));  // exact scaling
        uint16_t* out = xtl + b * NT * 64;
#pragma unroll
        for (int t = 0; t < NT; ++t) out[((lane >> 5) * NT + t) * 32 + (lane & 31)] = S::term(vs, t);
        if (lane == 0) sig[b] = ldexpf(1.0f, e);
      }
    }
  }
}

// The transposed operand of column block nb, rows 32 kr2 .. 32 kr2 + 31, of the tile in LDS
// at sb (lane-ordered blocks in symm_slot order): lane l gets T[32 kr2 + 8 (l/16) + e][16 nb
// + l % 16], e < 8 -- the B operand of x^T T and, read as an A operand, the block of T^T.
// Each lane reads rows 32 kr2 + 8 g + r (+4 h), columns 16 nb + 4 q4 with ds_read_tr16_b64.
template <typename op_t>
__device__ __forceinline__ op_t symm_tr_operand(const unsigned char* sb, int kr2, int nb, int lane) {
  const int g = lane >> 4, r = (lane >> 2) & 3, q4 = lane & 3;
  i16x4_t hv[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int rho = 32 * kr2 + 8 * g + r + 4 * h;
    const int blk = 2 * (rho >> 4) + (nb >> 1);
    const int ln = (rho & 15) + 16 * (2 * (nb & 1) + (q4 >> 1));
    hv[h] = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) i16x4_t*)(
        sb + blk * 1024 + symm_slot(ln) * 16 + 8 * (q4 & 1)));
  }
  const i16x8_t v = {hv[0][0], hv[0][1], hv[0][2], hv[0][3], hv[1][0], hv[1][1], hv[1][2], hv[1][3]};
  op_t o;
  __builtin_memcpy(&o, &v, 16);
  return o;
}

// A stage of the tile pipeline: the tile (8 blocks of 1 KiB), then, when the x terms travel
// with the tiles (PER_TILE), the terms of block J (the B operands of its row products) and of
// block I (the A operands of its transposed products, loaded with the first tile of a row;
// zeros otherwise), XL x 256 bytes each. Otherwise the item's terms are loaded into LDS once,
// before its tiles (Tmax XB bytes beside the stages). Measured at Q1 128^3 (max_k = 1000, two
// stages): FP16 (2 terms) 659.8 us with the item's terms in LDS (25.2 KiB per workgroup, six
// per CU), 670.8 us with the terms per tile (21.5 KiB, seven per CU); BF16 (3 terms) 686.3 us
// with the item's terms (27.3 KiB, five per CU), 661.7 us per tile (22.5 KiB, six per CU).
template <int NT>
struct SymmStage {
  static constexpr bool PER_TILE = NT > 2;
  static constexpr int XB = NT * 128;            // bytes of x terms per 64-row block
  static constexpr int XL = PER_TILE ? (XB + 255) / 256 : 0;  // 4-byte loads per block of terms
  static constexpr int XJ = 8192;                // offset of the terms of block J
  static constexpr int XI = XJ + 256 * XL;       // offset of the terms of block I
  static constexpr int BYTES = XI + 256 * XL;
  static constexpr int LOADS = 8 + 2 * XL;       // loads issued per tile, on every path
  // LDS of one workgroup: stages, 16 zero bytes, yl (64 floats), the item's terms (when not
  // PER_TILE), yv (64 Tmax floats), block scales (64 ceil(Tmax / 64) floats)
  static constexpr size_t lds_bytes(int NS, int Tmax) {
    return (size_t)NS * BYTES + 16 + 256 + (PER_TILE ? 0 : (size_t)Tmax * XB + 256) +
           (size_t)Tmax * 256 + (size_t)256 * ((Tmax + 63) / 64);
  }
};

// The loads of tile (I, J) into LDS stage `st`: its 8 blocks (zeros for absent blocks; every
// block zero when `real` is false, so every path issues the same count of loads), then the
// x terms of blocks J and I. voff: this lane's chunk of a block in bytes (16 symm_slot(lane)).
template <int NT>
__device__ __forceinline__ void symm_issue(uint32_t st, bool real, int I, int J, int T, int rL,
                                           int nbL, i32x4_t rm, uint32_t voff, i32x4_t rx,
                                           int lane) {
  using SS = SymmStage<NT>;
  const bool diag = J == I;
  const int nb = I < T - 1 ? 4 : nbL;
  const uint32_t toff = real ? (uint32_t)(symm_tile_off(I, J, T, nbL) * 2) : 0u;
  if (real && !diag && nb == 4) {  // a whole off-diagonal tile: 8 consecutive blocks
    dma16x4(st, voff, rm, toff);
    dma16x4(st + 4096u, voff, rm, toff + 4096u);
  } else {
    const int nk = diag && I == T - 1 ? (rL + 31) / 32 : 2;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int Rb = q >> 1, K2 = q & 1;
      const bool ok = real && Rb < nb && K2 < nk && !(diag && K2 == 1 && Rb < 2);
      const int idx = diag ? Rb + (nk == 2 ? (Rb > 2 ? Rb - 2 : 0) : 0) + K2 : q;
      dma16(st + (uint32_t)q * 1024u, voff, rm, ok ? toff + (uint32_t)idx * 1024u : kOob);
    }
  }
  const uint32_t l4 = 4u * (uint32_t)lane;
#pragma unroll
  for (int u = 0; u < SS::XL; ++u)
    dma4(st + (uint32_t)(SS::XJ + 256 * u), l4, rx,
         real ? (uint32_t)(J * SS::XB + 256 * u) : kOob);
#pragma unroll
  for (int u = 0; u < SS::XL; ++u)
    dma4(st + (uint32_t)(SS::XI + 256 * u), l4, rx,
         real && J == 0 ? (uint32_t)(I * SS::XB + 256 * u) : kOob);
}

// The tile rows of an item: the pairs p0 .. p_end-1 cover rows p and T-1-p, taken in
// increasing order (rows p0 .. p_end-1, then max(T - p_end, p_end) .. T-1-p0; the middle row
// of an odd T once), so a whole subdomain (p0 = 0, p_end = ceil(T/2)) reads its tiles in
// storage order. Row i of the item:
__device__ __forceinline__ int symm_row(int T, int p0, int p_end, int i) {
  const int nlow = p_end - p0;
  return i < nlow ? p0 + i : max(T - p_end, p_end) + (i - nlow);
}

// A cursor over the item's tiles: row index i (tile row I), tile J <= I.
struct SymmCursor {
  int i, I, J;
};
__device__ __forceinline__ SymmCursor symm_advance(SymmCursor c, int T, int p0, int p_end) {
  if (++c.J > c.I) {
    c.J = 0;
    c.I = symm_row(T, p0, p_end, ++c.i);
  }
  return c;
}

// The tile under the cursor into LDS stage `st`; zeros when the cursor is past the item.
template <int NT>
__device__ __forceinline__ void symm_issue_at(uint32_t st, SymmCursor c, int nrows, int T,
                                              int rL, int nbL, i32x4_t rm, uint32_t voff,
                                              i32x4_t rx, int lane) {
  const bool real = c.i < nrows;
  symm_issue<NT>(st, real, real ? c.I : 0, real ? c.J : 0, T, rL, nbL, rm, voff, rx, lane);
}

// LDS stages of the tile pipeline (a tile and its x terms, 8.5 KiB for FP16): NS - 1 tiles
// in flight per wavefront. Measured at Q1 128^3 FP16 with the x terms of the whole item in
// LDS instead (25 KiB per workgroup at max_k = 1000, six per CU): 2 stages 660.6 us, 3
// stages 676.5 us (four workgroups per CU), 4 stages 728 us. Half-tile stages (4 KiB, twice
// the waits per tile) were slower at every depth: 952 us with 3 stages, 991 us with 4,
// 1106 us with 5.
#ifndef DDS_SYMM_STAGES
#define DDS_SYMM_STAGES 2
#endif
constexpr int kSymmStages = DDS_SYMM_STAGES;
static_assert(kSymmStages >= 2 && SymmStage<3>::LOADS * (kSymmStages - 1) <= 63, "vmcnt range");

// One workgroup (one wavefront) per item (s, q): the tile-row pairs p = q G .. q G + G - 1
// of subdomain s (G = pairs_per_item, the last group of a subdomain possibly shorter), tile
// row by tile row. The tiles and their x terms stream through NS LDS stages (NS - 1 in
// flight while one is computed). The item's result accumulates in LDS in FP32 (yv: each tile's transposed part
// T_IJ^T x_I into block J, each row's product sum_J T_IJ x_J into block I), and every row the
// item touches is written once, at its end: Yp[ypos[s] + q 64 T + a] for the local rows
// a < min(k, 64 (T - q G)).
// FUSE: the x terms and scales are computed by each item from x itself (TX, gathered through
// sub_idx) into LDS (symm_pack_lds), instead of read from the pack (sigx, xt).
template <typename TM, typename TV, bool FUSE = false, typename TX = double>
__global__ __launch_bounds__(64) void sym_mfma_apply_kernel(
    int G, int Tmax, const int32_t* __restrict__ items, const int64_t* __restrict__ sub_ptr,
    const int64_t* __restrict__ blk_ptr, const int64_t* __restrict__ mat_off,
    const int64_t* __restrict__ ypos, const float* __restrict__ scale,
    const TM* __restrict__ mats, const float* __restrict__ sigx,
    const uint16_t* __restrict__ xt, TV* __restrict__ Yp,
    const dds_pcg_state* __restrict__ gate, const TX* __restrict__ xin = nullptr,
    const int32_t* __restrict__ sub_idx = nullptr) {
  using S = SymMfma<TM>;
  using op_t = typename S::op_t;
  constexpr int NT = S::NT;
  using SS = SymmStage<NT>;
  constexpr int NS = kSymmStages;
  constexpr int SB = SS::BYTES;    // one stage
  constexpr int LPT = SS::LOADS;   // loads per tile
  constexpr int NF = NS - 1;       // tiles in flight
  if (gate != nullptr && gate->done) return;
  extern __shared__ __align__(16) unsigned char sym_smem[];
  // [NS stages x SB][16 zero bytes][yl: 64 floats][x terms: Tmax XB + 256, when not
  // PER_TILE][yv: 64 Tmax floats][sig: 64 ceil(Tmax / 64) floats]
  unsigned char* stages = sym_smem;
  unsigned char* zero16 = sym_smem + NS * SB;
  float* yl = reinterpret_cast<float*>(zero16 + 16);
  unsigned char* xtl = zero16 + 16 + 256;
  float* yv = reinterpret_cast<float*>(xtl + (SS::PER_TILE ? 0 : Tmax * SS::XB + 256));
  float* sig = yv + 64 * Tmax;

  const int lane = threadIdx.x;
  const int s = items[2 * blockIdx.x];
  const int q = items[2 * blockIdx.x + 1];
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  const int T = (k + 63) / 64, rL = k - 64 * (T - 1), nbL = (rL + 15) / 16;
  const int P = (T + 1) / 2;
  const int p0 = q * G, p_end = min(P, p0 + G);
  const int I2max = T - 1 - p0;  // the item touches blocks 0..I2max
  const int nrows = (p_end - p0) + (T - p0) - max(T - p_end, p_end);
  int ntot = 0;
  for (int i = 0; i < nrows; ++i) ntot += symm_row(T, p0, p_end, i) + 1;
  const double out_sc = scale != nullptr ? static_cast<double>(scale[s]) : 1.0;
  const i32x4_t rm = buffer_rsrc(mats + mat_off[s], (uint32_t)(symm_store_elems(k) * 2));
  const uint32_t stage0 = lds_addr(stages);
  const uint32_t voff = 16u * (uint32_t)symm_slot(lane);  // the chunk this lane's slot receives

  if (lane < 4) reinterpret_cast<float*>(zero16)[lane] = 0.f;
  for (int a = lane; a < 64 * (I2max + 1); a += 64) yv[a] = 0.f;
  // the x terms (unless they travel with the tiles) and the scales of blocks 0..I2max
  // (zeros past the subdomain), then the first tiles; fused, the first tiles, then the terms
  // and scales computed from x
  static_assert(!(FUSE && SS::PER_TILE), "fused x terms are kept in LDS for the whole item");
  const int64_t b0x = FUSE ? 0 : blk_ptr[s];
  const i32x4_t rx = buffer_rsrc(xt + b0x * NT * 64, (uint32_t)(T * SS::XB));
  if constexpr (!FUSE) {
    const i32x4_t rs = buffer_rsrc(sigx + b0x, (uint32_t)(T * 4));
    if (!SS::PER_TILE) {
      const int xdw = (I2max + 1) * SS::XB / 4;  // dwords
      for (int d0 = 0; d0 < xdw; d0 += 64)
        dma4(lds_addr(xtl + 4 * d0), (uint32_t)lane * 4u, rx, 4u * (uint32_t)d0);
    }
    for (int d0 = 0; d0 <= I2max; d0 += 64)
      dma4(lds_addr(sig + d0), (uint32_t)lane * 4u, rs, 4u * (uint32_t)d0);
  }
  SymmCursor ic = {0, symm_row(T, p0, p_end, 0), 0};  // issue cursor
#pragma unroll
  for (int t = 0; t < NF; ++t) {
    symm_issue_at<NT>(stage0 + (uint32_t)(t % NS) * SB, ic, nrows, T, rL, nbL, rm, voff, rx, lane);
    if (ic.i < nrows) ic = symm_advance(ic, T, p0, p_end);
  }
  if constexpr (FUSE)
    symm_pack_lds<TM, TX>(I2max + 1, k, base, sub_idx, xin, lane, reinterpret_cast<uint16_t*>(xtl), sig);
  wait_vm<LPT * NF>();  // the terms and scales (issued before every tile) are in
  wave_lds_fence();

  const int tm = lane & 15;
  // operand of x rows 64 J + 32 h + 8 (l/16) + e, term l % 16 (zero past the NT terms), from
  // the terms of block J at xb: the B operand of a row product and the A' operand of a
  // transposed product
  auto x_op = [&](const unsigned char* xb, int h) __attribute__((always_inline)) -> op_t {
    const unsigned char* src = tm < NT ? xb + (h * NT + tm) * 64 + 16 * (lane >> 4) : zero16;
    op_t o;
    const uint4 v = *reinterpret_cast<const uint4*>(src);
    __builtin_memcpy(&o, &v, 16);
    return o;
  };
  const f32x4_t z4 = {0.f, 0.f, 0.f, 0.f};
  f32x4_t d[4];   // row product of the current tile row, row blocks 0..3
  f32x4_t dt[4];  // transposed product of the current tile, column blocks 0..3
  op_t a0, a1, b0, b1;
  float sI = 1.f;  // sig of the current tile row

  auto tile = [&](const unsigned char* sb, float sJ) __attribute__((always_inline)) {
#pragma unroll
    for (int Rb = 0; Rb < 4; ++Rb) {
      const uint4 c0 = *reinterpret_cast<const uint4*>(sb + (2 * Rb) * 1024 + voff);
      const uint4 c1 = *reinterpret_cast<const uint4*>(sb + (2 * Rb + 1) * 1024 + voff);
      op_t o0, o1;
      __builtin_memcpy(&o0, &c0, 16);
      __builtin_memcpy(&o1, &c1, 16);
      f32x4_t p = S::mfma(o0, b0, z4);
      p = S::mfma(o1, b1, p);
      d[Rb] += sJ * p;  // T x_J = sig[J] (T (x_J / sig[J]))
    }
#pragma unroll
    for (int kr2 = 0; kr2 < 2; ++kr2)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb)
        dt[nb] = S::mfma(kr2 == 0 ? a0 : a1, symm_tr_operand<op_t>(sb, kr2, nb, lane), dt[nb]);
  };
  // yv[64 J + 16 nb + lane] += sI x the transposed part (rows 16 nb + lane of D, the NT term
  // rows added), by lanes 0..15 themselves: the four reads, then the four writes (a wavefront's
  // LDS operations run in order, so the row sums of add_row see these)
  auto add_transposed = [&](int J, float sI) __attribute__((always_inline)) {
    if (lane < 16) {
      float* yJ = yv + 64 * J + lane;
      float o[4];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) o[nb] = yJ[16 * nb];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        float v = dt[nb][0];
#pragma unroll
        for (int t = 1; t < NT; ++t) v += dt[nb][t];
        yJ[16 * nb] = o[nb] + sI * v;
      }
    }
  };
  // yv[64 I + lane] += the row product (the NT term columns of D, lanes 16 g + t, added;
  // lane 16 g holds rows 16 Rb + 4 g + i)
  auto add_row = [&](int I) __attribute__((always_inline)) {
#pragma unroll
    for (int Rb = 0; Rb < 4; ++Rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = d[Rb][i];
        float v1 = __shfl_down(v, 1, 64);
        float v2 = NT > 2 ? __shfl_down(v, 2, 64) : 0.f;
        if ((lane & 15) == 0) yl[16 * Rb + 4 * (lane >> 4) + i] = v + v1 + v2;
      }
    }
    wave_lds_fence();
    yv[64 * I + lane] += yl[lane];
    wave_lds_fence();
  };

  SymmCursor cc = {0, symm_row(T, p0, p_end, 0), 0};  // compute cursor
  for (int t = 0; t < ntot; ++t) {
    // into the stage of task t - 1, read completely before
    symm_issue_at<NT>(stage0 + (uint32_t)((t + NF) % NS) * SB, ic, nrows, T, rL, nbL, rm, voff,
                      rx, lane);
    if (ic.i < nrows) ic = symm_advance(ic, T, p0, p_end);
    wait_vm<LPT * NF>();  // task t's loads are in (vector loads complete in issue order)
    const unsigned char* sb = stages + (t % NS) * SB;
    if (cc.J == 0) {  // a new tile row
      const unsigned char* xi = SS::PER_TILE ? sb + SS::XI : xtl + cc.I * SS::XB;
      a0 = x_op(xi, 0);
      a1 = x_op(xi, 1);
      sI = sig[cc.I];
#pragma unroll
      for (int i = 0; i < 4; ++i) d[i] = z4;
    }
    const unsigned char* xj = SS::PER_TILE ? sb + SS::XJ : xtl + cc.J * SS::XB;
    b0 = x_op(xj, 0);
    b1 = x_op(xj, 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) dt[i] = z4;
    tile(sb, sig[cc.J]);
    add_transposed(cc.J, sI);
    if (cc.J == cc.I) add_row(cc.I);  // the row is complete
    cc = symm_advance(cc, T, p0, p_end);
  }
  wait_vm<0>();  // no load into LDS outlives the wavefront

  TV* Y = Yp + ypos[s] + (int64_t)q * 64 * T;
  const int rows = min(k, 64 * (I2max + 1));
  // non-temporal: the result is read once, by the next kernel; measured on a stream of this
  // kernel's shape, its 31 MB of plain stores cost 60 us of the 3.9 GB read stream (each store
  // batch disturbs the reads' DRAM pages), non-temporal ones 50 us
  for (int a = lane; a < rows; a += 64)
    __builtin_nontemporal_store(static_cast<TV>(static_cast<double>(yv[a]) * out_sc), Y + a);
}

// ---- 16 right-hand sides on the symmetric tiles -------------------------------------------
// Y_s = scale[s] A_s^-1 X_s for an (n, 16) block X, from the matrix-core symmetric store
// (lower tiles read once, used for Y_I += T X_J and Y_J += T^T X_I), with X_s packed per
// 64-row block b = blk_ptr[s] + J and column c by dds_sym_mfma_pack16: sigma[b][c] = 2^e >
// max |X_b[:, c]| and X_b / sigma[b] (exact) in NT terms; for half h (rows 32 h ..) and term
// t the B operand (lane l: rows 32 h + 8 (l/16) + e, column l%16) at ((2 b + h) NT + t) 512 +
// 8 l elements.
//
// One workgroup of 4 wavefronts per subdomain (order[b], longest first), k <= 1024. Wavefront
// w owns rows 16 w .. 16 w + 15 of every 64-row block of the result and keeps them in
// registers (yacc[J], the MFMA D layout: lane l rows 4 (l/16) + i, column l % 16): from each
// tile (I, J) it takes row block w (T[16w.., :] X_J into row block I) and column block w
// (T[:, 16w..]^T X_I into row block J). The tiles stream through NS = NF + 2 LDS stages,
// each with the tile (8 blocks, symm_slot order), the packed X_J (2 NT KiB) and, with the
// first tile of a row, X_I (2 NT KiB); each wavefront issues a quarter of a stage's loads
// and a barrier per tile publishes them. A stage is refilled two barriers after its last
// read.
constexpr int kSym16MaxT = 16;  // 64-row blocks of the largest subdomain (k <= 1024)
#ifndef DDS_SYM16_AHEAD
#define DDS_SYM16_AHEAD 1
#endif
constexpr int kSym16Ahead = DDS_SYM16_AHEAD;  // tiles in flight
constexpr int kSym16Stages = kSym16Ahead + 2;

template <int NT>
struct Sym16Stage {
  static constexpr int XJ = 8192;                // packed X of block J (2 NT chunks of 1 KiB)
  static constexpr int XI = XJ + 2 * NT * 1024;  // packed X of block I
  static constexpr int BYTES = XI + 2 * NT * 1024;
  static constexpr int LOADS = 2 + NT;           // per wavefront and tile, on every path
};
static_assert(Sym16Stage<3>::LOADS * kSym16Ahead <= 63, "vmcnt range");

// Byte offset of block q (row block q >> 1, column half q & 1) of tile (I, J) in a
// subdomain's matrix-core symmetric store; kOob for a block the store does not hold.
__device__ __forceinline__ uint32_t symm_block_src(int q, bool real, int I, int J, int T, int rL,
                                                   int nbL) {
  if (!real) return kOob;
  const bool diag = J == I;
  const int nb = I < T - 1 ? 4 : nbL;
  const int Rb = q >> 1, K2 = q & 1;
  const uint32_t toff = (uint32_t)(symm_tile_off(I, J, T, nbL) * 2);
  if (!diag) return Rb < nb ? toff + (uint32_t)q * 1024u : kOob;
  const int nk = I == T - 1 ? (rL + 31) / 32 : 2;
  const bool ok = Rb < nb && K2 < nk && !(K2 == 1 && Rb < 2);
  const int idx = Rb + (nk == 2 ? (Rb > 2 ? Rb - 2 : 0) : 0) + K2;
  return ok ? toff + (uint32_t)idx * 1024u : kOob;
}

// The four wavefronts' LDS reads are complete and their awaited loads landed.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int N>
__device__ __forceinline__ void add_at(f32x4_t (&a)[N], int j, f32x4_t v) {
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (i == j) a[i] += v;
}

__device__ __forceinline__ bool sym16_done(const dds_pcg_state* st) {
  if (st == nullptr) return false;
  bool done = true;
#pragma unroll
  for (int j = 0; j < 16; ++j) done = done && st[j].done != 0;
  return done;
}

// One wavefront per 64-row block b of the (n, 16) block X (gathered through sub_idx): the
// column scales and the NT terms of X_b / sigma[b] in B-operand order.
template <typename TM>
__global__ __launch_bounds__(256) void sym16_pack_kernel(
    int64_t n_blk, const int32_t* __restrict__ blk_owner, const int64_t* __restrict__ blk_ptr,
    const int64_t* __restrict__ sub_ptr, const int32_t* __restrict__ sub_idx,
    const double* __restrict__ x, float* __restrict__ sigma, uint16_t* __restrict__ bp,
    const dds_pcg_state* __restrict__ st) {
  using S = SymMfma<TM>;
  constexpr int NT = S::NT;
  if (sym16_done(st)) return;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= n_blk) return;
  const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const int s = blk_owner[b];
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  const int J = (int)(b - blk_ptr[s]);
  double v[2][8];
  double m = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int a = 64 * J + 32 * h + 8 * g + e;
      const int ac = a < k ? a : 0;
      const int64_t row = sub_idx != nullptr ? (int64_t)sub_idx[base + ac] : base + ac;
      v[h][e] = a < k ? x[row * 16 + c] : 0.0;
      m = fmax(m, fabs(v[h][e]));
    }
  m = fmax(m, __shfl_xor(m, 16, 64));
  m = fmax(m, __shfl_xor(m, 32, 64));
  int ex = 0;
  if (m > 0.0) frexp(m, &ex);
  ex = ex < -126 ? -126 : (ex > 127 ? 127 : ex);
  const double inv = ldexp(1.0, -ex);  // exact scaling
  if (g == 0) sigma[b * 16 + c] = ldexpf(1.0f, ex);
  uint16_t* out = bp + (int64_t)b * 2 * NT * 512 + 8 * lane;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      uint16_t q[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) q[e] = S::term(static_cast<float>(v[h][e] * inv), t);
      uint4 w;
      __builtin_memcpy(&w, q, 16);
      *reinterpret_cast<uint4*>(out + (h * NT + t) * 512) = w;
    }
}

template <typename TM>
__global__ __launch_bounds__(256) void sym16_apply_kernel(
    const int32_t* __restrict__ order, const int64_t* __restrict__ sub_ptr,
    const int64_t* __restrict__ blk_ptr, const int64_t* __restrict__ mat_off,
    const float* __restrict__ scale, const float* __restrict__ sigma,
    const TM* __restrict__ mats, const uint16_t* __restrict__ bp, double* __restrict__ Y,
    const dds_pcg_state* __restrict__ st) {
  using S = SymMfma<TM>;
  using op_t = typename S::op_t;
  constexpr int NT = S::NT;
  using SS = Sym16Stage<NT>;
  constexpr int NS = kSym16Stages;
  constexpr int NF = kSym16Ahead;
  constexpr int SB = SS::BYTES;
  extern __shared__ __align__(16) unsigned char s16_smem[];
  if (sym16_done(st)) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int s = order[blockIdx.x];
  const int64_t base = sub_ptr[s];
  const int k = (int)(sub_ptr[s + 1] - base);
  const int T = (k + 63) / 64, rL = k - 64 * (T - 1), nbL = (rL + 15) / 16;
  const int ntot = T * (T + 1) / 2;
  const i32x4_t rm = buffer_rsrc(mats + mat_off[s], (uint32_t)(symm_store_elems(k) * 2));
  const int64_t b0 = blk_ptr[s];
  const i32x4_t rx = buffer_rsrc(bp + b0 * 2 * NT * 512, (uint32_t)(2 * T * NT * 1024));
  const uint32_t stage0 = lds_addr(s16_smem);
  // the column scales of the subdomain's blocks (16 floats each) after the stages, a quarter
  // per wavefront, issued before every tile
  const float* sgl = reinterpret_cast<const float*>(s16_smem + NS * SB);
  dma4(stage0 + (uint32_t)(NS * SB + 256 * w), 4u * (uint32_t)lane,
       buffer_rsrc(sigma + b0 * 16, (uint32_t)(T * 64)), 256u * (uint32_t)w);
  const uint32_t vsw = 16u * (uint32_t)symm_slot(lane);  // tile blocks: this lane's slot
  const uint32_t vnat = 16u * (uint32_t)lane;            // packed X: lane order

  // Tile rows in increasing order, even rows J = 0 .. I and odd rows J = I .. 0, so that a
  // row starts with the blocks of X the previous one used last (still in L2).
  auto row_first = [](int I, int J) { return J == ((I & 1) ? I : 0); };
  auto row_last = [](int I, int J) { return J == ((I & 1) ? 0 : I); };
  auto advance = [](int& I, int& J) __attribute__((always_inline)) {
    if (I & 1) {
      if (--J < 0) {
        ++I;
        J = 0;
      }
    } else if (++J > I) {
      J = ++I;
    }
  };
  // wavefront w's quarter of the loads of tile (I, J) into stage t % NS: tile blocks 2w and
  // 2w + 1, then chunks c = w NT .. w NT + NT - 1 of [X_J (2 NT chunks), X_I (2 NT chunks,
  // with the first tile of a row)]
  auto issue = [&](int t, int I, int J, bool real) __attribute__((always_inline)) {
    const uint32_t sa = stage0 + (uint32_t)(t % NS) * SB;
#pragma unroll
    for (int qq = 0; qq < 2; ++qq) {
      const int q = 2 * w + qq;
      dma16(sa + (uint32_t)q * 1024u, vsw, rm, symm_block_src(q, real, I, J, T, rL, nbL));
    }
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      const int c = w * NT + u;
      if (c < 2 * NT)
        dma16_keep(sa + (uint32_t)(SS::XJ + c * 1024), vnat, rx,
                   real ? (uint32_t)((2 * J * NT + c) * 1024) : kOob);
      else
        dma16_keep(sa + (uint32_t)(SS::XI + (c - 2 * NT) * 1024), vnat, rx,
                   real && row_first(I, J) ? (uint32_t)((2 * I * NT + c - 2 * NT) * 1024) : kOob);
    }
  };
  int iI = 0, iJ = 0;  // issue cursor
#pragma unroll
  for (int t = 0; t < NF; ++t) {
    issue(t, iI, iJ, t < ntot);
    advance(iI, iJ);
  }

  const f32x4_t z4 = {0.f, 0.f, 0.f, 0.f};
  f32x4_t yacc[kSym16MaxT];
#pragma unroll
  for (int j = 0; j < kSym16MaxT; ++j) yacc[j] = z4;
  f32x4_t racc = z4;  // row block w of the current tile row
  op_t xi[2][NT];     // X_I operands of the current tile row
  float sI = 1.f;     // their column scale (this lane's column)
  const int c = lane & 15;
  int cI = 0, cJ = 0;
  for (int t = 0; t < ntot; ++t) {
    issue(t + NF, iI, iJ, t + NF < ntot);  // into the stage read two barriers ago
    advance(iI, iJ);
    wait_vm<SS::LOADS * NF>();  // this wavefront's loads of tile t are in
    lds_barrier();              // and every wavefront's
    const unsigned char* sb = s16_smem + (t % NS) * SB;
    if (row_first(cI, cJ)) {
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < NT; ++u)
          xi[h][u] = *reinterpret_cast<const op_t*>(sb + SS::XI + (h * NT + u) * 1024 + vnat);
      racc = z4;
      sI = sgl[cI * 16 + c];
    }
    // row block w: T[16w.., 32h..] X_J[32h.., :] = sigma_J (T X_J / sigma_J)
    f32x4_t p = z4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const op_t a = *reinterpret_cast<const op_t*>(sb + (2 * w + h) * 1024 + vsw);
#pragma unroll
      for (int u = 0; u < NT; ++u)
        p = S::mfma(a, *reinterpret_cast<const op_t*>(sb + SS::XJ + (h * NT + u) * 1024 + vnat), p);
    }
    racc += sgl[cJ * 16 + c] * p;
    // column block w: T[32h.., 16w..]^T X_I[32h.., :]
    f32x4_t d = z4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const op_t a = symm_tr_operand<op_t>(sb, h, w, lane);
#pragma unroll
      for (int u = 0; u < NT; ++u) d = S::mfma(a, xi[h][u], d);
    }
    add_at(yacc, cJ, sI * d);
    if (row_last(cI, cJ)) add_at(yacc, cI, racc);
    advance(cI, cJ);
  }
  wait_vm<0>();  // no load into LDS outlives the workgroup

  const double sg = scale != nullptr ? (double)scale[s] : 1.0;
  double* Ys = Y + base * 16;
#pragma unroll
  for (int J = 0; J < kSym16MaxT; ++J) {
    if (J < T) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 64 * J + 16 * w + 4 * (lane >> 4) + i;
        if (row < k) __builtin_nontemporal_store((double)yacc[J][i] * sg, Ys + (int64_t)row * 16 + c);
      }
    }
  }
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

extern "C" int dds_sym_mfma_pack(int mat_dtype, int vec_dtype, int64_t n_s, int64_t n_blk,
                                 const int32_t* blk_owner, const int64_t* blk_ptr,
                                 const int64_t* sub_ptr, const int32_t* sub_idx, const void* x,
                                 float* sig, void* xt, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_blk >= 0 && ceil_div(n_blk, 4 * kPackBlocks) <= INT32_MAX,
              "dds_sym_mfma_pack: bad sizes");
  if (n_blk == 0 || n_s == 0) return DDS_OK;
  DDS_REQUIRE(blk_owner && blk_ptr && sub_ptr && x && sig && xt, "dds_sym_mfma_pack: null pointer");
  hipStream_t st = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n_blk, 4 * kPackBlocks));
#define DDS_PACK(TM, TV)                                                                      \
  do {                                                                                        \
    hipLaunchKernelGGL((sym_pack_x_kernel<TM, TV>), grid, dim3(256), 0, st, n_blk, blk_owner,  \
                       blk_ptr, sub_ptr, sub_idx, static_cast<const TV*>(x), sig,             \
                       static_cast<uint16_t*>(xt), gate);                                     \
    DDS_LAUNCHED();                                                                           \
    return DDS_OK;                                                                            \
  } while (0)
  if (mat_dtype == DDS_F16 && vec_dtype == DDS_F64) DDS_PACK(_Float16, double);
  if (mat_dtype == DDS_F16 && vec_dtype == DDS_F32) DDS_PACK(_Float16, float);
  if (mat_dtype == DDS_BF16 && vec_dtype == DDS_F64) DDS_PACK(bf16_t, double);
  if (mat_dtype == DDS_BF16 && vec_dtype == DDS_F32) DDS_PACK(bf16_t, float);
#undef DDS_PACK
  set_error("dds_sym_mfma_pack: unsupported dtypes (mat %d, vec %d)", mat_dtype, vec_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_sym_mfma_apply_packed(int mat_dtype, int vec_dtype, int64_t n_s,
                                         int64_t n_items, int64_t max_k, int64_t pairs_per_item,
                                         const int32_t* items, const int64_t* sub_ptr,
                                         const int64_t* blk_ptr, const int64_t* mat_off,
                                         const int64_t* ypos, const float* scale,
                                         const void* mats, const float* sig, const void* xt,
                                         void* Yp, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_items >= 0 && n_items <= INT32_MAX,
              "dds_sym_mfma_apply_packed: bad sizes");
  DDS_REQUIRE(max_k >= 0 && max_k <= kSymMfmaMaxK,
              "dds_sym_mfma_apply_packed: subdomains of at most %d unknowns (got %lld)",
              kSymMfmaMaxK, (long long)max_k);
  DDS_REQUIRE(pairs_per_item >= 1 && pairs_per_item <= kSymMfmaMaxK,
              "dds_sym_mfma_apply_packed: pairs_per_item %lld out of range",
              (long long)pairs_per_item);
  if (n_items == 0 || n_s == 0) return DDS_OK;
  DDS_REQUIRE(items && sub_ptr && blk_ptr && mat_off && ypos && mats && sig && xt && Yp,
              "dds_sym_mfma_apply_packed: null pointer");
  hipStream_t st = as_stream(stream);
  const int Tmax = (int)ceil_div(max_k, 64);
  const int G = (int)pairs_per_item;
  int dev = 0, max_lds = 0;
  DDS_HIP(hipGetDevice(&dev));
  DDS_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
#define DDS_SYMM(TM, TV)                                                                       \
  do {                                                                                         \
    const size_t lds = SymmStage<SymMfma<TM>::NT>::lds_bytes(kSymmStages, Tmax);               \
    DDS_REQUIRE(lds <= (size_t)max_lds,                                                        \
                "dds_sym_mfma_apply_packed: %zu bytes of LDS for max_k %lld exceed the %d "    \
                "per workgroup", lds, (long long)max_k, max_lds);                              \
    if (lds > 65536)                                                                           \
      DDS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&sym_mfma_apply_kernel<TM, TV>), \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));      \
    hipLaunchKernelGGL((sym_mfma_apply_kernel<TM, TV>), dim3((unsigned)n_items), dim3(64), lds,  \
                       st, G, Tmax, items, sub_ptr, blk_ptr, mat_off, ypos, scale,             \
                       static_cast<const TM*>(mats), sig, static_cast<const uint16_t*>(xt),    \
                       static_cast<TV*>(Yp), gate);                                            \
    DDS_LAUNCHED();                                                                            \
    return DDS_OK;                                                                             \
  } while (0)
  if (mat_dtype == DDS_F16 && vec_dtype == DDS_F64) DDS_SYMM(_Float16, double);
  if (mat_dtype == DDS_F16 && vec_dtype == DDS_F32) DDS_SYMM(_Float16, float);
  if (mat_dtype == DDS_BF16 && vec_dtype == DDS_F64) DDS_SYMM(bf16_t, double);
  if (mat_dtype == DDS_BF16 && vec_dtype == DDS_F32) DDS_SYMM(bf16_t, float);
#undef DDS_SYMM
  set_error("dds_sym_mfma_apply_packed: unsupported dtypes (mat %d, vec %d)", mat_dtype,
            vec_dtype);
  return DDS_EUNSUPPORTED;
}

// The pack, then the apply. pairs_per_item: G as above.
extern "C" int dds_local_apply_sym_mfma(int mat_dtype, int vec_dtype, int64_t n_s,
                                        int64_t n_items, int64_t max_k, int64_t pairs_per_item,
                                        const int32_t* items, const int64_t* sub_ptr,
                                        const int32_t* sub_idx, const int64_t* mat_off,
                                        const int64_t* ypos, const float* scale,
                                        const void* mats, const void* x, int64_t n_blk,
                                        const int32_t* blk_owner, const int64_t* blk_ptr,
                                        float* sig, void* xt, void* Yp,
                                        const dds_pcg_state* gate, void* stream) {
  if (n_items == 0 || n_s == 0) return DDS_OK;
  const int rc = dds_sym_mfma_pack(mat_dtype, vec_dtype, n_s, n_blk, blk_owner, blk_ptr,
                                   sub_ptr, sub_idx, x, sig, xt, gate, stream);
  if (rc != DDS_OK) return rc;
  return dds_sym_mfma_apply_packed(mat_dtype, vec_dtype, n_s, n_items, max_k, pairs_per_item,
                                   items, sub_ptr, blk_ptr, mat_off, ypos, scale, mats, sig, xt,
                                   Yp, gate, stream);
}

// The apply with the x pack done by each item in its own LDS (FP16 tiles only: their two
// terms per block fit beside the stages), so no pack pass and no packed x in HBM. The same
// result bits as the pack followed by the apply.
extern "C" int dds_local_apply_sym_mfma_fused(int mat_dtype, int vec_dtype, int y_dtype,
                                              int64_t n_s, int64_t n_items, int64_t max_k,
                                              int64_t pairs_per_item, const int32_t* items,
                                              const int64_t* sub_ptr, const int32_t* sub_idx,
                                              const int64_t* mat_off, const int64_t* ypos,
                                              const float* scale, const void* mats,
                                              const void* x, void* Yp,
                                              const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_items >= 0 && n_items <= INT32_MAX,
              "dds_local_apply_sym_mfma_fused: bad sizes");
  DDS_REQUIRE(max_k >= 0 && max_k <= kSymMfmaMaxK,
              "dds_local_apply_sym_mfma_fused: subdomains of at most %d unknowns (got %lld)",
              kSymMfmaMaxK, (long long)max_k);
  DDS_REQUIRE(pairs_per_item >= 1 && pairs_per_item <= kSymMfmaMaxK,
              "dds_local_apply_sym_mfma_fused: pairs_per_item %lld out of range",
              (long long)pairs_per_item);
  if (n_items == 0 || n_s == 0) return DDS_OK;
  DDS_REQUIRE(items && sub_ptr && mat_off && ypos && mats && x && Yp,
              "dds_local_apply_sym_mfma_fused: null pointer");
  hipStream_t st = as_stream(stream);
  const int Tmax = (int)ceil_div(max_k, 64);
  const int G = (int)pairs_per_item;
  int dev = 0, max_lds = 0;
  DDS_HIP(hipGetDevice(&dev));
  DDS_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
#define DDS_SYMF(TV, TX)                                                                       \
  do {                                                                                         \
    const size_t lds = SymmStage<SymMfma<_Float16>::NT>::lds_bytes(kSymmStages, Tmax);         \
    DDS_REQUIRE(lds <= (size_t)max_lds,                                                        \
                "dds_local_apply_sym_mfma_fused: %zu bytes of LDS for max_k %lld exceed the "  \
                "%d per workgroup", lds, (long long)max_k, max_lds);                           \
    const void* fn = reinterpret_cast<const void*>(&sym_mfma_apply_kernel<_Float16, TV, true, TX>); \
    if (lds > 65536)                                                                           \
      DDS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));  \
    hipLaunchKernelGGL((sym_mfma_apply_kernel<_Float16, TV, true, TX>), dim3((unsigned)n_items), \
                       dim3(64), lds, st, G, Tmax, items, sub_ptr, nullptr, mat_off, ypos,     \
                       scale, static_cast<const _Float16*>(mats), nullptr, nullptr,            \
                       static_cast<TV*>(Yp), gate, static_cast<const TX*>(x), sub_idx);        \
    DDS_LAUNCHED();                                                                            \
    return DDS_OK;                                                                             \
  } while (0)
  if (mat_dtype == DDS_F16) {
    if (vec_dtype == DDS_F64 && y_dtype == DDS_F64) DDS_SYMF(double, double);
    if (vec_dtype == DDS_F64 && y_dtype == DDS_F32) DDS_SYMF(float, double);
    if (vec_dtype == DDS_F32 && y_dtype == DDS_F32) DDS_SYMF(float, float);
  }
#undef DDS_SYMF
  set_error("dds_local_apply_sym_mfma_fused: unsupported dtypes (mat %d, vec %d, y %d)",
            mat_dtype, vec_dtype, y_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_sym_mfma_pack16(int mat_dtype, int64_t n_blk, const int32_t* blk_owner,
                                   const int64_t* blk_ptr, const int64_t* sub_ptr,
                                   const int32_t* sub_idx, const double* x, float* sigma,
                                   void* bp, const dds_pcg_state* st, void* stream) {
  DDS_REQUIRE(n_blk >= 0, "dds_sym_mfma_pack16: bad block count");
  if (n_blk == 0) return DDS_OK;
  DDS_REQUIRE(blk_owner && blk_ptr && sub_ptr && x && sigma && bp,
              "dds_sym_mfma_pack16: null pointer");
  hipStream_t s = as_stream(stream);
  const dim3 grid((unsigned)ceil_div(n_blk, 4));
  if (mat_dtype == DDS_F16) {
    hipLaunchKernelGGL(sym16_pack_kernel<_Float16>, grid, dim3(256), 0, s, n_blk, blk_owner,
                       blk_ptr, sub_ptr, sub_idx, x, sigma, static_cast<uint16_t*>(bp), st);
  } else if (mat_dtype == DDS_BF16) {
    hipLaunchKernelGGL(sym16_pack_kernel<bf16_t>, grid, dim3(256), 0, s, n_blk, blk_owner,
                       blk_ptr, sub_ptr, sub_idx, x, sigma, static_cast<uint16_t*>(bp), st);
  } else {
    set_error("dds_sym_mfma_pack16: storage must be F16 or BF16 (got %d)", mat_dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_sym_mfma_apply16(int mat_dtype, int64_t n_s, int64_t max_k,
                                    const int32_t* order, const int64_t* sub_ptr,
                                    const int64_t* blk_ptr, const int64_t* mat_off,
                                    const float* scale, const float* sigma, const void* mats,
                                    const void* bp, double* Y, const dds_pcg_state* st,
                                    void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_s <= INT32_MAX, "dds_sym_mfma_apply16: bad subdomain count");
  DDS_REQUIRE(max_k >= 0 && max_k <= 64 * kSym16MaxT,
              "dds_sym_mfma_apply16: max_k %lld exceeds %d", (long long)max_k, 64 * kSym16MaxT);
  if (n_s == 0 || max_k == 0) return DDS_OK;
  DDS_REQUIRE(order && sub_ptr && blk_ptr && mat_off && sigma && mats && bp && Y,
              "dds_sym_mfma_apply16: null pointer");
  hipStream_t s = as_stream(stream);
#define DDS_S16(TM)                                                                             \
  do {                                                                                          \
    constexpr size_t lds = (size_t)kSym16Stages * Sym16Stage<SymMfma<TM>::NT>::BYTES +           \
                           (size_t)kSym16MaxT * 64;                                             \
    static_assert(lds <= 160 * 1024, "LDS");                                                    \
    if (lds > 65536)                                                                            \
      DDS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&sym16_apply_kernel<TM>),        \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));       \
    hipLaunchKernelGGL((sym16_apply_kernel<TM>), dim3((unsigned)n_s), dim3(256), lds, s, order,  \
                       sub_ptr, blk_ptr, mat_off, scale, sigma, static_cast<const TM*>(mats),   \
                       static_cast<const uint16_t*>(bp), Y, st);                                 \
  } while (0)
  if (mat_dtype == DDS_F16) {
    DDS_S16(_Float16);
  } else if (mat_dtype == DDS_BF16) {
    DDS_S16(bf16_t);
  } else {
    set_error("dds_sym_mfma_apply16: storage must be F16 or BF16 (got %d)", mat_dtype);
    return DDS_EUNSUPPORTED;
  }
#undef DDS_S16
  DDS_LAUNCHED();
  return DDS_OK;
}

extern "C" int dds_sym_combine(int dtype, int64_t n_s, int64_t max_k, int64_t pairs_per_item,
                               const int64_t* sub_ptr, const int64_t* ypos, const void* Yp,
                               void* y, const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_s >= 0 && n_s <= INT32_MAX && max_k >= 0 && ceil_div(max_k, 256) <= 65535,
              "dds_sym_combine: bad sizes");
  DDS_REQUIRE(pairs_per_item >= 1 && pairs_per_item <= INT32_MAX,
              "dds_sym_combine: pairs_per_item %lld out of range", (long long)pairs_per_item);
  const int G = (int)pairs_per_item;
  if (n_s == 0 || max_k == 0) return DDS_OK;
  DDS_REQUIRE(sub_ptr && ypos && Yp && y, "dds_sym_combine: null pointer");
  const dim3 grid((unsigned)n_s, (unsigned)ceil_div(max_k, 256));
  hipStream_t s = as_stream(stream);
  if (dtype == DDS_F64) {
    hipLaunchKernelGGL(sym_combine_kernel<double>, grid, dim3(256), 0, s, G, sub_ptr, ypos,
                       static_cast<const double*>(Yp), static_cast<double*>(y), gate);
  } else if (dtype == DDS_F32) {
    hipLaunchKernelGGL(sym_combine_kernel<float>, grid, dim3(256), 0, s, G, sub_ptr, ypos,
                       static_cast<const float*>(Yp), static_cast<float*>(y), gate);
  } else {
    set_error("dds_sym_combine: unsupported dtype %d", dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  return DDS_OK;
}

// ---------------------------------------------------------------------------------------
// The matrix-core symmetric apply without device allocation: its workspace (the block
// scales and 16-bit terms of the x pack) lives in `scratch`, a file-scope device array of
// kSymmScratchBytes that every GPU the library runs on holds until the process ends. Its
// device address is looked up once per device (hipGetSymbolAddress, that device current) on
// the first call there; kernels get that address, never the array's host-side name.
// ---------------------------------------------------------------------------------------
static constexpr size_t kSymmScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kSymmScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kScratchDevices] = {};

int scratch_for_current_device(unsigned char** out) {
  int dev = 0;
  DDS_HIP(hipGetDevice(&dev));
  DDS_REQUIRE(dev >= 0 && dev < kScratchDevices, "device %d beyond the %d supported", dev,
              kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    DDS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DDS_REQUIRE(p != nullptr, "scratch: no device address on device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return DDS_OK;
}
}  // namespace

extern "C" int dds_local_apply_sym_mfma_noalloc(
    int mat_dtype, int vec_dtype, int64_t n_s, int64_t n_items, int64_t max_k,
    int64_t pairs_per_item, const int32_t* items, const int64_t* sub_ptr, const int32_t* sub_idx,
    const int64_t* mat_off, const int64_t* ypos, const float* scale, const void* mats,
    const void* x, int64_t n_blk, const int32_t* blk_owner, const int64_t* blk_ptr, void* Yp,
    const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(mat_dtype == DDS_F16 || mat_dtype == DDS_BF16,
              "dds_local_apply_sym_mfma_noalloc: unsupported matrix dtype %d", mat_dtype);
  DDS_REQUIRE(n_blk >= 0, "dds_local_apply_sym_mfma_noalloc: bad block count");
  const int nt = mat_dtype == DDS_F16 ? 2 : 3;
  const size_t sig_bytes = (size_t)ceil_div(n_blk * 4, 256) * 256;
  const size_t need = sig_bytes + (size_t)n_blk * nt * 128;
  DDS_REQUIRE(need <= kSymmScratchBytes,
              "dds_local_apply_sym_mfma_noalloc: %lld blocks of 64 rows need %zu bytes of "
              "workspace, beyond the %zu of the static scratch (use dds_local_apply_sym_mfma "
              "with a caller workspace)",
              (long long)n_blk, need, kSymmScratchBytes);
  if (n_items == 0 || n_s == 0) return DDS_OK;
  unsigned char* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != DDS_OK) return rc;
  return dds_local_apply_sym_mfma(mat_dtype, vec_dtype, n_s, n_items, max_k, pairs_per_item,
                                  items, sub_ptr, sub_idx, mat_off, ypos, scale, mats, x, n_blk,
                                  blk_owner, blk_ptr, reinterpret_cast<float*>(ws),
                                  ws + sig_bytes, Yp, gate, stream);
}

