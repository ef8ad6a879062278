// Shared device/host helpers for the dd_solvers HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "dd_solvers_hip.h"

namespace dds {

// ---------------------------------------------------------------------------------------
// Error reporting: every entry point returns a dds_status; the message of the last failure
// on the calling thread is kept for dds_last_error().
// ---------------------------------------------------------------------------------------
void set_error(const char* fmt, ...);

#define DDS_REQUIRE(cond, ...)                \
  do {                                        \
    if (!(cond)) {                            \
      ::dds::set_error(__VA_ARGS__);          \
      return DDS_EINVAL;                      \
    }                                         \
  } while (0)

#define DDS_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::dds::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                       __FILE__, __LINE__);                                        \
      return DDS_EHIP;                                                             \
    }                                                                              \
  } while (0)

#define DDS_LAUNCHED()                                                             \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) {                                                        \
      ::dds::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),  \
                       __FILE__, __LINE__);                                        \
      return DDS_EHIP;                                                             \
    }                                                                              \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------------------
// Storage types. fp16 is clang's _Float16; bf16 is kept as raw 16-bit words and widened
// by a shift (exact).
// ---------------------------------------------------------------------------------------
struct bf16_t {
  uint16_t bits;
};

__device__ __forceinline__ float widen(bf16_t v) {
  return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
}
__device__ __forceinline__ float widen(_Float16 v) { return static_cast<float>(v); }
__device__ __forceinline__ float widen(float v) { return v; }
__device__ __forceinline__ double widen(double v) { return v; }

template <typename T>
struct Widened {
  using type = float;
};
template <>
struct Widened<double> {
  using type = double;
};

// 16-byte vector of storage elements (one global_load_dwordx4 per lane).
template <typename T>
struct alignas(16) Vec16 {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
};

template <typename T>
__device__ __forceinline__ Vec16<T> load16(const T* p) {
  return *reinterpret_cast<const Vec16<T>*>(p);
}

// Streaming loads for data read once per launch (matrix values and indices): the
// non-temporal hint keeps them from evicting the vectors that are reused from L2 / MALL.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ Vec16<T> load16_nt(const T* p) {
  const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
  Vec16<T> r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

template <typename T>
__device__ __forceinline__ T load_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X. For a grid whose size
// is a multiple of 8, this renumbers them so that each XCD runs a contiguous range of
// logical block ids (neighbouring rows share that XCD's L2).
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_block_id() {
  const int64_t b = blockIdx.x;
  const int64_t per = gridDim.x / kXcds;
  return (b % kXcds) * per + b / kXcds;
}

// ---------------------------------------------------------------------------------------
// Wave64 / block reductions (fixed order: butterfly within a wave, then waves in order).
// ---------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the block; the result is valid in every thread. `scratch` holds >= nwaves T.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  T s = 0;
  for (int i = 0; i < nw; ++i) s += scratch[i];
  return s;
}

// Maximum number of per-block partial sums a reduction-producing kernel writes.
constexpr int kMaxPartials = DDS_MAX_PARTIALS;
// Grid of the grid-stride vector kernels (their partial counts stay at or below this):
// 2048 workgroups of 256 fill the 256 CUs 8 deep.
constexpr int kVecGrid = 2048;


// ---- loads straight into LDS ------------------------------------------------------------
typedef int i32x4_t __attribute__((ext_vector_type(4)));

// Buffer resource (gfx9 family): 64-bit base, byte range; loads past the range return zero.
__device__ __forceinline__ i32x4_t buffer_rsrc(const void* base, uint32_t bytes) {
  const uint64_t b = reinterpret_cast<uint64_t>(base);
  i32x4_t r;
  r[0] = static_cast<int>(static_cast<uint32_t>(b));
  r[1] = static_cast<int>(static_cast<uint32_t>(b >> 32));
  r[2] = static_cast<int>(bytes);
  r[3] = 0x00020000;
  return r;
}

__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return static_cast<uint32_t>(
      reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) const void*)(p)));
}

__device__ __forceinline__ i32x4_t sgpr4(i32x4_t v) {  // wave-uniform: SGPR operands
  i32x4_t r;
  r[0] = __builtin_amdgcn_readfirstlane(v[0]);
  r[1] = __builtin_amdgcn_readfirstlane(v[1]);
  r[2] = __builtin_amdgcn_readfirstlane(v[2]);
  r[3] = __builtin_amdgcn_readfirstlane(v[3]);
  return r;
}

// Loads straight into LDS (buffer_load ... lds): lane l's bytes land at m0 + offset + l *
// width, read from base + soffset + voffset + offset (the instruction offset applies to
// both addresses). The compiler does not track these loads, so their completion is awaited
// explicitly with s_waitcnt vmcnt (vector loads complete in issue order). The LDS reads
// issued before are awaited first (lgkmcnt(0)), so a stage is never overwritten while it is
// being read. The 16-byte forms stream matrix tiles, read once per apply: non-temporal.
// Four consecutive 1 KiB blocks: LDS m0 + 1024 q, memory soffset + voff + 1024 q, q < 4.
__device__ __forceinline__ void dma16x4(uint32_t m0, uint32_t voff, i32x4_t rsrc, uint32_t soff) {
  m0 = __builtin_amdgcn_readfirstlane(m0);
  soff = __builtin_amdgcn_readfirstlane(soff);
  rsrc = sgpr4(rsrc);
  asm volatile(
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, %0\n\t"
      "s_nop 0\n\t"
      "buffer_load_dwordx4 %1, %2, %3 offen nt lds\n\t"
      "buffer_load_dwordx4 %1, %2, %3 offen offset:1024 nt lds\n\t"
      "buffer_load_dwordx4 %1, %2, %3 offen offset:2048 nt lds\n\t"
      "buffer_load_dwordx4 %1, %2, %3 offen offset:3072 nt lds" ::"s"(m0),
      "v"(voff), "s"(rsrc), "s"(soff)
      : "memory");
}
// One 1 KiB block: LDS m0, memory soffset + voff.
__device__ __forceinline__ void dma16(uint32_t m0, uint32_t voff, i32x4_t rsrc, uint32_t soff) {
  m0 = __builtin_amdgcn_readfirstlane(m0);
  soff = __builtin_amdgcn_readfirstlane(soff);
  rsrc = sgpr4(rsrc);
  asm volatile(
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, %0\n\t"
      "s_nop 0\n\t"
      "buffer_load_dwordx4 %1, %2, %3 offen nt lds" ::"s"(m0),
      "v"(voff), "s"(rsrc), "s"(soff)
      : "memory");
}
// One 1 KiB block kept in the caches (data read again by other wavefronts): LDS m0, memory
// soffset + voff.
__device__ __forceinline__ void dma16_keep(uint32_t m0, uint32_t voff, i32x4_t rsrc, uint32_t soff) {
  m0 = __builtin_amdgcn_readfirstlane(m0);
  soff = __builtin_amdgcn_readfirstlane(soff);
  rsrc = sgpr4(rsrc);
  asm volatile(
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, %0\n\t"
      "s_nop 0\n\t"
      "buffer_load_dwordx4 %1, %2, %3 offen lds" ::"s"(m0),
      "v"(voff), "s"(rsrc), "s"(soff)
      : "memory");
}
// 4 bytes per lane: LDS m0 + 4 l, memory soffset + voff.
__device__ __forceinline__ void dma4(uint32_t m0, uint32_t voff, i32x4_t rsrc, uint32_t soff) {
  m0 = __builtin_amdgcn_readfirstlane(m0);
  soff = __builtin_amdgcn_readfirstlane(soff);
  rsrc = sgpr4(rsrc);
  asm volatile(
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, %0\n\t"
      "s_nop 0\n\t"
      "buffer_load_dword %1, %2, %3 offen lds" ::"s"(m0),
      "v"(voff), "s"(rsrc), "s"(soff)
      : "memory");
}
template <int N>
__device__ __forceinline__ void wait_vm() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

}  // namespace dds
