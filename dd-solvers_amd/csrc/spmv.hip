// Sparse matrix-vector products: BSR (column-major blocks) and CSR, with an optional
// x . y partial-sum epilogue so the PCG step gets p . q without another pass.
#include "common.hpp"

namespace dds {
namespace {

constexpr int kBlock = 256;

// two-element vectors for 8- and 16-byte loads
template <typename T>
struct Pair;
template <>
struct Pair<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Pair<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
typedef int32_t i32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void write_partial(double dot, double* partials, double* scratch) {
  const double t = block_sum(dot, scratch);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// ---------------------------------------------------------------------------------------
// CSR: G lanes per row (grid-stride over rows), strided partial sums then a butterfly.
// Accumulates in double whatever the value type.
// ---------------------------------------------------------------------------------------
template <typename T, typename I, int G>
__global__ __launch_bounds__(kBlock) void csr_kernel(int64_t n_rows, const I* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col,
                                                      const T* __restrict__ val,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      double* __restrict__ partials,
                                                      const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kBlock / 64];
  if (gate != nullptr && gate->done) return;
  const int lg = threadIdx.x & (G - 1);
  const int64_t group = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G;
  const int64_t n_groups = (int64_t)gridDim.x * kBlock / G;
  double dot = 0.0;
  for (int64_t row = group; row < n_rows; row += n_groups) {
    const int64_t s = row_ptr[row];
    const int64_t e = row_ptr[row + 1];
    double acc = 0.0;
    for (int64_t p = s + lg; p < e; p += G)
      acc += static_cast<double>(load_nt(val + p)) * static_cast<double>(x[load_nt(col + p)]);
    acc = group_sum<G>(acc);
    if (lg == 0) {
      y[row] = static_cast<T>(acc);
      if (partials != nullptr) dot += acc * static_cast<double>(x[row]);
    }
  }
  if (partials != nullptr) write_partial(dot, partials, scratch);
}

// ---------------------------------------------------------------------------------------
// CSR with long rows (lanes = BT > 64, e.g. R_C A of a coarse space: few rows of hundreds
// of entries): one workgroup of BT threads per row (grid-stride over rows), so a row costs
// one or two dependent load steps instead of nnz/64; the waves' butterfly sums are added in
// wave order through LDS (deterministic). The x . y partial of a workgroup is written by its
// first thread.
// ---------------------------------------------------------------------------------------
template <typename T, typename I, int BT>
__global__ __launch_bounds__(BT) void csr_rows_kernel(int64_t n_rows, const I* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col,
                                                      const T* __restrict__ val,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      double* __restrict__ partials,
                                                      const dds_pcg_state* __restrict__ gate) {
  constexpr int NW = BT / 64;
  __shared__ double wsum[2][NW];
  if (gate != nullptr && gate->done) return;
  const int wid = threadIdx.x >> 6;
  double dot = 0.0;
  int buf = 0;
  for (int64_t row = blockIdx.x; row < n_rows; row += gridDim.x) {
    const int64_t s = row_ptr[row];
    const int64_t e = row_ptr[row + 1];
    double acc = 0.0;
    for (int64_t p = s + threadIdx.x; p < e; p += BT)
      acc += static_cast<double>(load_nt(val + p)) * static_cast<double>(x[load_nt(col + p)]);
    acc = group_sum<64>(acc);
    if ((threadIdx.x & 63) == 0) wsum[buf][wid] = acc;
    __syncthreads();  // (wsum alternates between two buffers, so one barrier per row suffices)
    if (threadIdx.x == 0) {
      double r = wsum[buf][0];
#pragma unroll
      for (int w = 1; w < NW; ++w) r += wsum[buf][w];
      y[row] = static_cast<T>(r);
      if (partials != nullptr) dot += r * static_cast<double>(x[row]);
    }
    buf ^= 1;
  }
  if (partials != nullptr && threadIdx.x == 0) partials[blockIdx.x] = dot;
}

// ---------------------------------------------------------------------------------------
// BSR, general block size: one lane per scalar row (the row-per-thread order of the
// reference's RPT-I kernel: blocks of the row in stored order, then c ascending).
// ---------------------------------------------------------------------------------------
template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void bsr_row_kernel(int64_t n_block_rows, int b,
                                                          const I* __restrict__ crow,
                                                          const int32_t* __restrict__ col,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ x,
                                                          T* __restrict__ y,
                                                          double* __restrict__ partials,
                                                          const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kBlock / 64];
  if (gate != nullptr && gate->done) return;
  const int64_t n_rows = n_block_rows * b;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const int64_t bb = (int64_t)b * b;
  double dot = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < n_rows; row += stride) {
    const int64_t br = row / b;
    const int r = (int)(row - br * b);
    double acc = 0.0;
    for (int64_t blk = crow[br]; blk < crow[br + 1]; ++blk) {
      const T* v = val + blk * bb + r;
      const int64_t off = (int64_t)col[blk] * b;
      for (int c = 0; c < b; ++c)
        acc += static_cast<double>(v[c * b]) * static_cast<double>(x[off + c]);
    }
    y[row] = static_cast<T>(acc);
    if (partials != nullptr) dot += acc * static_cast<double>(x[row]);
  }
  if (partials != nullptr) write_partial(dot, partials, scratch);
}

// ---------------------------------------------------------------------------------------
// BSR with a block size B known at compile time (the DG sizes: (p+1)(p+2)/2 on triangles,
// (p+1)(p+2)(p+3)/6 on tetrahedra): one lane per scalar row as above (blocks in stored
// order, c ascending), with the column loop unrolled so that a block's B values and B
// entries of x are all in flight at once, and the next block's column index loaded ahead.
// ---------------------------------------------------------------------------------------
template <typename T, typename I, int B>
__global__ __launch_bounds__(kBlock) void bsr_rowb_kernel(int64_t n_block_rows,
                                                          const I* __restrict__ crow,
                                                          const int32_t* __restrict__ col,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ x,
                                                          T* __restrict__ y,
                                                          double* __restrict__ partials,
                                                          const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kBlock / 64];
  if (gate != nullptr && gate->done) return;
  constexpr int BB = B * B;
  const int64_t n_rows = n_block_rows * B;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double dot = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < n_rows; row += stride) {
    const int64_t br = row / B;
    const int r = (int)(row - br * B);
    double acc = 0.0;
    const int64_t e0 = (int64_t)crow[br], e1 = (int64_t)crow[br + 1];
    int32_t cn = e0 < e1 ? col[e0] : 0;
    for (int64_t blk = e0; blk < e1; ++blk) {
      const int32_t cc = cn;
      if (blk + 1 < e1) cn = col[blk + 1];
      const T* v = val + blk * BB + r;
      const T* xs = x + (int64_t)cc * B;
      T vv[B], xx[B];
#pragma unroll
      for (int c = 0; c < B; ++c) vv[c] = load_nt(v + c * B);
#pragma unroll
      for (int c = 0; c < B; ++c) xx[c] = xs[c];
#pragma unroll
      for (int c = 0; c < B; ++c) acc += static_cast<double>(vv[c]) * static_cast<double>(xx[c]);
    }
    y[row] = static_cast<T>(acc);
    if (partials != nullptr) dot += acc * static_cast<double>(x[row]);
  }
  if (partials != nullptr) write_partial(dot, partials, scratch);
}

// ---------------------------------------------------------------------------------------
// BSR with a small block size B (3 .. 7: DG elements, and b = 4): one lane per block. The G lanes
// of a block row take its blocks lg, lg + G, ...; a lane reads its block's B*B values
// (contiguous) and the B entries of x it multiplies, keeps B row sums, and the G lanes'
// sums are combined by a butterfly; lane r mod G writes row r. Far fewer load instructions
// than one lane per scalar row (which reads one value per instruction per row).
// ---------------------------------------------------------------------------------------
template <typename T, typename I, int B, int G>
__global__ __launch_bounds__(kBlock) void bsr_blk_kernel(int64_t n_block_rows,
                                                          const I* __restrict__ crow,
                                                          const int32_t* __restrict__ col,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ x,
                                                          T* __restrict__ y,
                                                          double* __restrict__ partials,
                                                          const dds_pcg_state* __restrict__ gate) {
  static_assert((G & (G - 1)) == 0 && G <= 64, "bad BSR group");
  __shared__ double scratch[kBlock / 64];
  if (gate != nullptr && gate->done) return;
  const int lg = threadIdx.x & (G - 1);
  const int64_t group = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G;
  const int64_t n_groups = (int64_t)gridDim.x * kBlock / G;
  double dot = 0.0;
  for (int64_t br = group; br < n_block_rows; br += n_groups) {
    double acc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] = 0.0;
    for (int64_t blk = (int64_t)crow[br] + lg; blk < (int64_t)crow[br + 1]; blk += G) {
      const T* v = val + blk * (B * B);
      const T* xs = x + (int64_t)col[blk] * B;
      double xv[B];
#pragma unroll
      for (int c = 0; c < B; ++c) xv[c] = static_cast<double>(xs[c]);
#pragma unroll
      for (int c = 0; c < B; ++c)
#pragma unroll
        for (int r = 0; r < B; ++r) acc[r] += static_cast<double>(v[c * B + r]) * xv[c];
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
      for (int r = 0; r < B; ++r) acc[r] += __shfl_xor(acc[r], off, G);
#pragma unroll
    for (int r = 0; r < B; ++r) {
      if ((r & (G - 1)) == lg) {
        const int64_t row = br * B + r;
        y[row] = static_cast<T>(acc[r]);
        if (partials != nullptr) dot += acc[r] * static_cast<double>(x[row]);
      }
    }
  }
  if (partials != nullptr) write_partial(dot, partials, scratch);
}

// ---------------------------------------------------------------------------------------
// BSR, runs of block rows streamed by a wavefront: the values of NR consecutive block rows
// are one contiguous run (column-major blocks: element e of the run is in block e / B^2,
// column (e mod B^2) / B, row e mod B). G lanes read it V elements at a time (one 8- or
// 16-byte load) at V l + S t, S = V G, t = 0, 1, ...; S is a multiple of B, so a lane's
// elements always lie in the rows (V l + k) mod B and it keeps V running sums. Its elements
// come in increasing order, so the block row they belong to only moves forward: when it
// changes the lane parks its sums in LDS (slot [block row][V l + k], zero if the lane had
// nothing there) and starts again. The x entries the run multiplies (B per block) are
// gathered into LDS first, so the stream issues one global load per step instead of a
// value load plus a dependent gather; a run with more than NR * MAXB blocks reads x from
// global memory instead. U steps are loaded together. Then row r of block row j is the sum
// of the slots [j][q], q = r, r + B, ... < S, in that order. Each wavefront takes a
// contiguous range of runs, contiguous per XCD; no workgroup barrier (wavefronts run
// different counts).
// ---------------------------------------------------------------------------------------
constexpr int kRunWaves = 4;

template <typename T, int V>
struct VecLoad;
template <typename T>
struct VecLoad<T, 1> {
  __device__ static __forceinline__ void load(const T* p, T (&v)[1]) { v[0] = load_nt(p); }
};
template <typename T>
struct VecLoad<T, 2> {
  __device__ static __forceinline__ void load(const T* p, T (&v)[2]) {
    const typename Pair<T>::type t =
        __builtin_nontemporal_load(reinterpret_cast<const typename Pair<T>::type*>(p));
    v[0] = t.x;
    v[1] = t.y;
  }
};

// LDS ordering among the lanes of one wavefront
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, typename I, int B, int V, int G, int NR, int U, int MAXB>
__global__ __launch_bounds__(64 * kRunWaves) void bsr_run_kernel(
    int64_t n_block_rows, const I* __restrict__ crow, const int32_t* __restrict__ col,
    const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y,
    double* __restrict__ partials, const dds_pcg_state* __restrict__ gate) {
  constexpr int BB = B * B;
  constexpr int S = V * G;
  static_assert(G <= 64 && NR <= 64 && S % B == 0 && (V == 1 || (V == 2 && B % 2 == 0)),
                "bad BSR run tiling");
  __shared__ double red[kRunWaves][NR * S];
  __shared__ int32_t bst[kRunWaves][NR + 1];
  __shared__ T xsh[kRunWaves][NR * MAXB * B];
  __shared__ double scratch[kRunWaves];
  if (gate != nullptr && gate->done) return;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t n_runs = (n_block_rows + NR - 1) / NR;
  const int64_t wave = xcd_block_id() * kRunWaves + wid;
  const int64_t n_waves = (int64_t)gridDim.x * kRunWaves;
  const int64_t per = (n_runs + n_waves - 1) / n_waves;
  const int64_t q0 = wave * per;
  const int64_t q1 = q0 + per < n_runs ? q0 + per : n_runs;
  double* R = red[wid];
  int32_t* bs = bst[wid];
  double dot = 0.0;
  for (int64_t q = q0; q < q1; ++q) {
    const int64_t br0 = q * NR;
    const int nr = (int)(n_block_rows - br0 < NR ? n_block_rows - br0 : NR);
    const int64_t b0 = (int64_t)crow[br0];
    if (lane <= nr) bs[lane] = (int32_t)(((int64_t)crow[br0 + lane] - b0) * BB);
    for (int i = lane; i < NR * S; i += 64) R[i] = 0.0;
    wave_sync();
    const int len = bs[nr];
    const int nbk = len / BB;
    const bool staged = nbk <= NR * MAXB;
    T* X = xsh[wid];
    if (staged) {
      for (int i = lane; i < nbk * B; i += 64) {
        const int kb = i / B, c = i - kb * B;
        X[i] = x[(int64_t)col[b0 + kb] * B + c];
      }
    }
    wave_sync();
    if (lane < G) {
      const T* seg = val + b0 * BB;
      const int32_t* cb = col + b0;
      int cur = 0;
      int nxt = bs[1];
      double acc[V];
#pragma unroll
      for (int k = 0; k < V; ++k) acc[k] = 0.0;
      for (int e0 = V * lane; e0 < len; e0 += U * S) {
        T vv[U][V];
        T xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = e0 + u * S;
          if (e < len) {
            const int blk = e / BB;
            const int c = (e - blk * BB) / B;
            VecLoad<T, V>::load(seg + e, vv[u]);
            xv[u] = staged ? X[blk * B + c] : x[(int64_t)cb[blk] * B + c];
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = e0 + u * S;
          if (e < len) {
            while (e >= nxt) {  // the next element is in a later block row: park the sums
#pragma unroll
              for (int k = 0; k < V; ++k) {
                R[cur * S + V * lane + k] = acc[k];
                acc[k] = 0.0;
              }
              nxt = bs[++cur + 1];
            }
#pragma unroll
            for (int k = 0; k < V; ++k)
              acc[k] += static_cast<double>(vv[u][k]) * static_cast<double>(xv[u]);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < V; ++k) R[cur * S + V * lane + k] = acc[k];
    }
    wave_sync();
    for (int o = lane; o < nr * B; o += 64) {
      const int j = o / B, r = o - j * B;
      const double* Rj = R + j * S;
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < S / B; ++i) sum += Rj[r + i * B];
      const int64_t row = (br0 + j) * B + r;
      y[row] = static_cast<T>(sum);
      if (partials != nullptr) dot += sum * static_cast<double>(x[row]);
    }
    wave_sync();
  }
  if (partials != nullptr) write_partial(dot, partials, scratch);
}

// ---------------------------------------------------------------------------------------
// BSR with a power-of-two block size B (used for b = 2): G lanes stream the block row's
// contiguous value segment (nnzb_row * B * B elements), so consecutive lanes read
// consecutive values. With
// column-major blocks element q of the segment sits in row q % B, which is lane % B for
// every lane because G is a multiple of B; lanes with equal lane % B are summed by a
// butterfly over xor offsets G/2 .. B, leaving row r's result in lane r.
// ---------------------------------------------------------------------------------------
template <typename T, typename I, int B, int G>
__global__ __launch_bounds__(kBlock) void bsr_seg_kernel(int64_t n_block_rows,
                                                          const I* __restrict__ crow,
                                                          const int32_t* __restrict__ col,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ x,
                                                          T* __restrict__ y,
                                                          double* __restrict__ partials,
                                                          const dds_pcg_state* __restrict__ gate) {
  static_assert((B & (B - 1)) == 0 && G % B == 0 && G <= 64, "bad BSR tiling");
  constexpr int BB = B * B;
  __shared__ double scratch[kBlock / 64];
  if (gate != nullptr && gate->done) return;
  const int lg = threadIdx.x & (G - 1);
  const int64_t group = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G;
  const int64_t n_groups = (int64_t)gridDim.x * kBlock / G;
  double dot = 0.0;
  for (int64_t br = group; br < n_block_rows; br += n_groups) {
    const int64_t q0 = (int64_t)crow[br] * BB;
    const int64_t q1 = (int64_t)crow[br + 1] * BB;
    double acc = 0.0;
    for (int64_t q = q0 + lg; q < q1; q += G) {
      const int64_t blk = q / BB;
      const int c = (int)((q / B) & (B - 1));
      acc += static_cast<double>(val[q]) * static_cast<double>(x[(int64_t)col[blk] * B + c]);
    }
#pragma unroll
    for (int off = G / 2; off >= B; off >>= 1) acc += __shfl_xor(acc, off, G);
    if (lg < B) {
      const int64_t row = br * B + lg;
      y[row] = static_cast<T>(acc);
      if (partials != nullptr) dot += acc * static_cast<double>(x[row]);
    }
  }
  if (partials != nullptr) write_partial(dot, partials, scratch);
}

// ---------------------------------------------------------------------------------------
// Sliced ELLPACK with 64-row slices (SELL-64), slots stored in pairs: slice q holds rows
// 64q .. 64q+63 as w_q slots (entry j of each row), padded with (col = own row, value 0).
// Slots 2p and 2p+1 of lane l are adjacent, at slice_ptr[q] + 128p + 2l (+1); an odd last
// slot j is stored alone at slice_ptr[q] + 64j + l. One wavefront per slice, lane = row:
// a slot pair is one coalesced 16-byte load per lane (1 KiB per wavefront for FP64) and one
// 4-byte load of its two compact column offsets, so the loads in flight carry twice the
// bytes of one-slot loads; no row pointers are read, the padding costs only its bytes, and
// each row still sums its entries in slot order.
// Compact column indices: a slice whose slots each span < 65536 columns stores 16-bit
// offsets from a per-slot base (2 bytes per entry instead of 4); other slices keep int32
// columns, in the same pair order. Contiguous runs of slices per wave, contiguous runs of
// waves per XCD. Two wavefronts per workgroup: a workgroup retires when its slowest wave
// does, and short workgroups pack the CUs more evenly.
// ---------------------------------------------------------------------------------------
constexpr int kSellWaves = 2;


// Row sum of one lane over its w slots; pair(p) gives the two column indices of slot pair
// p, single(j) the column of the odd last slot.
template <typename T, typename PairF, typename SingleF>
__device__ __forceinline__ double sell_slice(const T* __restrict__ v, int lane, int w, PairF pair,
                                             SingleF single, const T* __restrict__ x) {
  using T2 = typename Pair<T>::type;
  const T2* vp = reinterpret_cast<const T2*>(v) + lane;
  const int np = w >> 1;
  double acc = 0.0;
  int p = 0;
  for (; p + 3 < np; p += 4) {
    i32x2_t cc[4];
    T2 vv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) cc[u] = pair(p + u);
#pragma unroll
    for (int u = 0; u < 4; ++u) vv[u] = __builtin_nontemporal_load(vp + 64 * (p + u));
    T xx[8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xx[2 * u] = x[cc[u].x];
      xx[2 * u + 1] = x[cc[u].y];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc += static_cast<double>(vv[u].x) * static_cast<double>(xx[2 * u]);
      acc += static_cast<double>(vv[u].y) * static_cast<double>(xx[2 * u + 1]);
    }
  }
  for (; p < np; ++p) {
    const i32x2_t cc = pair(p);
    const T2 vv = __builtin_nontemporal_load(vp + 64 * p);
    acc += static_cast<double>(vv.x) * static_cast<double>(x[cc.x]);
    acc += static_cast<double>(vv.y) * static_cast<double>(x[cc.y]);
  }
  if (w & 1) {
    const int j = w - 1;
    acc += static_cast<double>(load_nt(v + 64 * j + lane)) * static_cast<double>(x[single(j)]);
  }
  return acc;
}

template <typename T>
__global__ __launch_bounds__(64 * kSellWaves) void sell_kernel(
    int64_t n_rows, int64_t n_slices, const int64_t* __restrict__ slice_ptr,
    const int64_t* __restrict__ col_ptr, const uint16_t* __restrict__ col16,
    const int32_t* __restrict__ col32, const int32_t* __restrict__ slot_base,
    const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y,
    double* __restrict__ partials, const dds_pcg_state* __restrict__ gate) {
  __shared__ double scratch[kSellWaves];
  if (gate != nullptr && gate->done) return;
  const int lane = threadIdx.x & 63;
  // contiguous slice ranges per wave, contiguous wave ranges per XCD (grid % 8 == 0)
  const int64_t wave = xcd_block_id() * kSellWaves + (threadIdx.x >> 6);
  const int64_t n_waves = (int64_t)gridDim.x * kSellWaves;
  const int64_t per = (n_slices + n_waves - 1) / n_waves;
  const int64_t q0 = wave * per;
  const int64_t q1 = q0 + per < n_slices ? q0 + per : n_slices;
  double dot = 0.0;
  for (int64_t q = q0; q < q1; ++q) {
    const int64_t s0 = slice_ptr[q];
    const int w = (int)((slice_ptr[q + 1] - s0) >> 6);
    const int64_t cp = col_ptr[q];
    double acc;
    if (cp >= 0) {
      // 16-bit offsets: slot pair p of this lane is one 32-bit word (cp is a multiple of 64)
      const uint32_t* d2 = reinterpret_cast<const uint32_t*>(col16 + cp) + lane;
      const uint16_t* d1 = col16 + cp + lane;
      const int32_t* b = slot_base + (s0 >> 6);
      acc = sell_slice<T>(
          val + s0, lane, w,
          [&](int p) {
            const uint32_t o = load_nt(d2 + 64 * p);
            return i32x2_t{b[2 * p] + (int32_t)(uint16_t)(o & 0xffffu),
                           b[2 * p + 1] + (int32_t)(uint16_t)(o >> 16)};
          },
          [&](int j) { return b[j] + (int32_t)load_nt(d1 + 64 * j); }, x);
    } else {
      const int32_t* c = col32 + (~cp);
      const i32x2_t* c2 = reinterpret_cast<const i32x2_t*>(c) + lane;
      acc = sell_slice<T>(
          val + s0, lane, w, [&](int p) { return __builtin_nontemporal_load(c2 + 64 * p); },
          [&](int j) { return load_nt(c + 64 * j + lane); }, x);
    }
    const int64_t row = q * 64 + lane;
    if (row < n_rows) {
      y[row] = static_cast<T>(acc);
      if (partials != nullptr) dot += acc * static_cast<double>(x[row]);
    }
  }
  if (partials != nullptr) write_partial(dot, partials, scratch);
}

int64_t grid_for(int64_t work_items, bool capped) {
  int64_t g = ceil_div(std::max<int64_t>(work_items, 1), kBlock);
  const int64_t cap = capped ? kMaxPartials : (int64_t)1 << 20;
  return std::min(g, cap);
}

template <typename T, typename I>
int launch_csr(int64_t n_rows, const void* row_ptr, const int32_t* col, const void* values,
               const void* x, void* y, int lanes, double* partials, int* n_partials,
               const dds_pcg_state* gate, hipStream_t s) {
  const I* rp = static_cast<const I*>(row_ptr);
  const T* v = static_cast<const T*>(values);
  const T* xv = static_cast<const T*>(x);
  T* yv = static_cast<T*>(y);
  if (lanes > 64) {
    const int64_t grid = std::min<int64_t>(n_rows, partials != nullptr ? kMaxPartials : 1 << 20);
    if (n_partials) *n_partials = partials ? (int)grid : 0;
#define DDS_CSR_ROWS(BT)                                                                    \
  hipLaunchKernelGGL((csr_rows_kernel<T, I, BT>), dim3((unsigned)grid), dim3(BT), 0, s, n_rows, \
                     rp, col, v, xv, yv, partials, gate)
    switch (lanes) {
      case 128: DDS_CSR_ROWS(128); break;
      case 256: DDS_CSR_ROWS(256); break;
      case 512: DDS_CSR_ROWS(512); break;
      default: DDS_CSR_ROWS(1024); break;
    }
#undef DDS_CSR_ROWS
    DDS_LAUNCHED();
    return DDS_OK;
  }
  const int64_t grid = grid_for(n_rows * lanes, partials != nullptr);
  if (n_partials) *n_partials = partials ? (int)grid : 0;
#define DDS_CSR(G)                                                                            \
  hipLaunchKernelGGL((csr_kernel<T, I, G>), dim3((unsigned)grid), dim3(kBlock), 0, s, n_rows, \
                     rp, col, v, xv, yv, partials, gate)
  switch (lanes) {
    case 1: DDS_CSR(1); break;
    case 2: DDS_CSR(2); break;
    case 4: DDS_CSR(4); break;
    case 8: DDS_CSR(8); break;
    case 16: DDS_CSR(16); break;
    case 32: DDS_CSR(32); break;
    default: DDS_CSR(64); break;
  }
#undef DDS_CSR
  DDS_LAUNCHED();
  return DDS_OK;
}

template <typename T, typename I>
int launch_bsr(int64_t nbr, int64_t b, const void* crow, const int32_t* col, const void* values,
               const void* x, void* y, double* partials, int* n_partials,
               const dds_pcg_state* gate, hipStream_t s) {
  const I* cr = static_cast<const I*>(crow);
  const T* v = static_cast<const T*>(values);
  const T* xv = static_cast<const T*>(x);
  T* yv = static_cast<T*>(y);
  const bool capped = partials != nullptr;
  int64_t grid;
#define DDS_SEG(B, G)                                                                    \
  do {                                                                                   \
    grid = grid_for(nbr * G, capped);                                                    \
    hipLaunchKernelGGL((bsr_seg_kernel<T, I, B, G>), dim3((unsigned)grid), dim3(kBlock), \
                       0, s, nbr, cr, col, v, xv, yv, partials, gate);                   \
  } while (0)
#define DDS_BLK(B, G)                                                                    \
  do {                                                                                   \
    grid = grid_for(nbr * G, capped);                                                    \
    hipLaunchKernelGGL((bsr_blk_kernel<T, I, B, G>), dim3((unsigned)grid), dim3(kBlock), \
                       0, s, nbr, cr, col, v, xv, yv, partials, gate);                   \
  } while (0)
#define DDS_ROWB(B)                                                                      \
  do {                                                                                   \
    grid = grid_for(nbr * B, capped);                                                    \
    hipLaunchKernelGGL((bsr_rowb_kernel<T, I, B>), dim3((unsigned)grid), dim3(kBlock), 0, \
                       s, nbr, cr, col, v, xv, yv, partials, gate);                      \
  } while (0)
#define DDS_RUN(B, V, G, NR, U, MAXB)                                                      \
  do {                                                                                     \
    grid = std::min<int64_t>(capped ? kMaxPartials : (int64_t)1 << 20,                     \
                             ceil_div(ceil_div(ceil_div(nbr, NR), kRunWaves), kXcds) *      \
                                 kXcds);                                                   \
    hipLaunchKernelGGL((bsr_run_kernel<T, I, B, V, G, NR, U, MAXB>), dim3((unsigned)grid), \
                       dim3(64 * kRunWaves), 0, s, nbr, cr, col, v, xv, yv, partials, gate); \
  } while (0)
  // measured on the DG shapes of the reference's BSR benchmark (2D and 3D): runs of block
  // rows streamed (x staged in LDS) for the DG sizes from 6 up, one lane per block for b =
  // 3 .. 7, the value segment of a block row streamed for b = 2, one lane per row with the
  // block size compiled in for b = 8 (the Q1 block size sweep)
  // two-element loads need values aligned to two elements (always so for a fresh tensor)
  const bool pairs_ok = (reinterpret_cast<uintptr_t>(values) % (2 * sizeof(T))) == 0;
  switch (pairs_ok ? b : -b) {
    case 2: case -2: DDS_SEG(2, 16); break;
    case 3: case -3: DDS_BLK(3, 4); break;
    case 4: case -4: DDS_BLK(4, 8); break;
    case 5: case -5: DDS_BLK(5, 4); break;
    case 6: DDS_RUN(6, 2, 63, 4, 3, 6); break;
    case 7: case -7: DDS_BLK(7, 8); break;
    case 8: case -8: DDS_ROWB(8); break;
    case 10: DDS_RUN(10, 2, 60, 4, 4, 6); break;
    case 15: case -15: DDS_RUN(15, 1, 60, 2, 8, 6); break;
    case 20: DDS_RUN(20, 2, 60, 1, 6, 6); break;
    case 21: case -21: DDS_RUN(21, 1, 63, 2, 4, 6); break;
    case 28: DDS_RUN(28, 2, 56, 1, 7, 6); break;
    default:
      grid = grid_for(nbr * b, capped);
      hipLaunchKernelGGL((bsr_row_kernel<T, I>), dim3((unsigned)grid), dim3(kBlock), 0, s, nbr,
                         (int)b, cr, col, v, xv, yv, partials, gate);
  }
#undef DDS_SEG
#undef DDS_BLK
#undef DDS_ROWB
#undef DDS_RUN
  if (n_partials) *n_partials = partials ? (int)grid : 0;
  DDS_LAUNCHED();
  return DDS_OK;
}

}  // namespace
}  // namespace dds

using namespace dds;

extern "C" int dds_csrmv(int val_dtype, int idx_dtype, int64_t n_rows, const void* row_ptr,
                         const int32_t* col, const void* values, const void* x, void* y,
                         int lanes, double* dot_partials, int* n_partials,
                         const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_rows >= 0, "dds_csrmv: negative size");
  if (n_partials) *n_partials = 0;
  if (n_rows == 0) return DDS_OK;
  DDS_REQUIRE(row_ptr && x && y, "dds_csrmv: null pointer");
  DDS_REQUIRE(lanes >= 1 && lanes <= 1024 && (lanes & (lanes - 1)) == 0,
              "dds_csrmv: lanes must be a power of two in [1, 1024] (got %d)", lanes);
  hipStream_t s = as_stream(stream);
  if (val_dtype == DDS_F64 && idx_dtype == DDS_I64)
    return launch_csr<double, int64_t>(n_rows, row_ptr, col, values, x, y, lanes, dot_partials,
                                       n_partials, gate, s);
  if (val_dtype == DDS_F64 && idx_dtype == DDS_I32)
    return launch_csr<double, int32_t>(n_rows, row_ptr, col, values, x, y, lanes, dot_partials,
                                       n_partials, gate, s);
  if (val_dtype == DDS_F32 && idx_dtype == DDS_I64)
    return launch_csr<float, int64_t>(n_rows, row_ptr, col, values, x, y, lanes, dot_partials,
                                      n_partials, gate, s);
  if (val_dtype == DDS_F32 && idx_dtype == DDS_I32)
    return launch_csr<float, int32_t>(n_rows, row_ptr, col, values, x, y, lanes, dot_partials,
                                      n_partials, gate, s);
  set_error("dds_csrmv: unsupported dtypes (values %d, row_ptr %d)", val_dtype, idx_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_bsrmv(int val_dtype, int idx_dtype, int64_t n_block_rows, int64_t block_size,
                         const void* crow, const int32_t* col, const void* values, const void* x,
                         void* y, double* dot_partials, int* n_partials,
                         const dds_pcg_state* gate, void* stream) {
  DDS_REQUIRE(n_block_rows >= 0 && block_size >= 1, "dds_bsrmv: bad sizes");
  if (n_partials) *n_partials = 0;
  if (n_block_rows == 0) return DDS_OK;
  DDS_REQUIRE(crow && x && y, "dds_bsrmv: null pointer");
  hipStream_t s = as_stream(stream);
  if (val_dtype == DDS_F64 && idx_dtype == DDS_I64)
    return launch_bsr<double, int64_t>(n_block_rows, block_size, crow, col, values, x, y,
                                       dot_partials, n_partials, gate, s);
  if (val_dtype == DDS_F64 && idx_dtype == DDS_I32)
    return launch_bsr<double, int32_t>(n_block_rows, block_size, crow, col, values, x, y,
                                       dot_partials, n_partials, gate, s);
  if (val_dtype == DDS_F32 && idx_dtype == DDS_I64)
    return launch_bsr<float, int64_t>(n_block_rows, block_size, crow, col, values, x, y,
                                      dot_partials, n_partials, gate, s);
  if (val_dtype == DDS_F32 && idx_dtype == DDS_I32)
    return launch_bsr<float, int32_t>(n_block_rows, block_size, crow, col, values, x, y,
                                      dot_partials, n_partials, gate, s);
  set_error("dds_bsrmv: unsupported dtypes (values %d, crow %d)", val_dtype, idx_dtype);
  return DDS_EUNSUPPORTED;
}

extern "C" int dds_sellmv(int val_dtype, int64_t n_rows, const int64_t* slice_ptr,
                          const int64_t* col_ptr, const uint16_t* col16, const int32_t* col32,
                          const int32_t* slot_base, const void* values, const void* x, void* y,
                          double* dot_partials, int* n_partials, const dds_pcg_state* gate,
                          void* stream) {
  DDS_REQUIRE(n_rows >= 0, "dds_sellmv: negative size");
  if (n_partials) *n_partials = 0;
  if (n_rows == 0) return DDS_OK;
  DDS_REQUIRE(slice_ptr && col_ptr && slot_base && values && x && y, "dds_sellmv: null pointer");
  DDS_REQUIRE(col16 || col32, "dds_sellmv: no column indices");
  const int64_t n_slices = ceil_div(n_rows, 64);
  // one slice per wave up to the partial-sum capacity, contiguous runs of slices beyond;
  // a multiple of the XCD count (so every XCD gets an equal contiguous share)
  const int64_t grid = std::min<int64_t>(
      kMaxPartials, ceil_div(ceil_div(n_slices, kSellWaves), kXcds) * kXcds);
  hipStream_t s = as_stream(stream);
  if (val_dtype == DDS_F64) {
    hipLaunchKernelGGL(sell_kernel<double>, dim3((unsigned)grid), dim3(64 * kSellWaves), 0, s, n_rows,
                       n_slices, slice_ptr, col_ptr, col16, col32, slot_base, static_cast<const double*>(values),
                       static_cast<const double*>(x), static_cast<double*>(y), dot_partials, gate);
  } else if (val_dtype == DDS_F32) {
    hipLaunchKernelGGL(sell_kernel<float>, dim3((unsigned)grid), dim3(64 * kSellWaves), 0, s, n_rows,
                       n_slices, slice_ptr, col_ptr, col16, col32, slot_base, static_cast<const float*>(values),
                       static_cast<const float*>(x), static_cast<float*>(y), dot_partials, gate);
  } else {
    set_error("dds_sellmv: unsupported dtype %d", val_dtype);
    return DDS_EUNSUPPORTED;
  }
  DDS_LAUNCHED();
  if (n_partials) *n_partials = dot_partials ? (int)grid : 0;
  return DDS_OK;
}
