// device_common.hpp -- wave64 group primitives shared by the gfx950 kernels.
//
// A "group" is G consecutive lanes that own one CSR row.  For G <= 64 a group
// lives inside one wavefront, so group-level ordering of LDS traffic needs only
// a compiler fence (LDS executes one wave's instructions in issue order); for
// G > 64 the group is the whole workgroup and uses s_barrier.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace appt {

constexpr int kWave = 64;

template <int G>
__device__ __forceinline__ void gsync() {
  if constexpr (G <= kWave) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// Sum over the G lanes of a group (G <= 64: shuffles; G > 64: whole block via
// LDS scratch of at least blockDim/64 entries).
template <int G, typename T>
__device__ __forceinline__ T gsum(T v, T* scratch) {
  if constexpr (G <= kWave) {
#pragma unroll
    for (int m = G / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, G);
    return v;
  } else {
#pragma unroll
    for (int m = kWave / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, kWave);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) scratch[w] = v;
    __syncthreads();
    T s = 0;
    const int nw = blockDim.x / kWave;
    for (int t = 0; t < nw; t++) s += scratch[t];
    __syncthreads();
    return s;
  }
}

template <int G, typename T>
__device__ __forceinline__ T gmin(T v) {
#pragma unroll
  for (int m = G / 2; m > 0; m >>= 1) { T o = __shfl_xor(v, m, G); v = o < v ? o : v; }
  return v;
}
template <int G, typename T>
__device__ __forceinline__ T gmax(T v) {
#pragma unroll
  for (int m = G / 2; m > 0; m >>= 1) { T o = __shfl_xor(v, m, G); v = o > v ? o : v; }
  return v;
}

// Exclusive prefix sum over a group (G <= 64).  Returns the exclusive value and
// writes the group total to *total.
template <int G, typename T>
__device__ __forceinline__ T gscan_excl(T v, int lane, T* total) {
  T inc = v;
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    T o = __shfl_up(inc, d, G);
    if (lane >= d) inc += o;
  }
  *total = __shfl(inc, G - 1, G);
  return inc - v;
}

// Block-wide exclusive scan (any blockDim multiple of 64, <= 1024).
template <typename T>
__device__ __forceinline__ T block_scan_excl(T v, T* scratch /* >= 17 */, T* total) {
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  T inc = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = __shfl_up(inc, d, kWave);
    if (lane >= d) inc += o;
  }
  __syncthreads();
  if (lane == kWave - 1) scratch[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int t = 0; t < nw; t++) { T x = scratch[t]; scratch[t] = run; run += x; }
    scratch[16] = run;
  }
  __syncthreads();
  T r = scratch[w] + inc - v;
  *total = scratch[16];
  __syncthreads();
  return r;
}

// Mask of this lane's group inside a 64-bit wave ballot.
template <int G>
__device__ __forceinline__ uint64_t group_bits(uint64_t ballot) {
  if constexpr (G >= kWave) {
    return ballot;
  } else {
    const int base = (threadIdx.x % kWave) / G * G;
    return (ballot >> base) & ((1ull << G) - 1ull);
  }
}

// Fibonacci hashing: top log2(T) bits of key * 2^32/phi.
__device__ __forceinline__ uint32_t hash_slot(int32_t key, int logT) {
  return ((uint32_t)key * 2654435769u) >> (32 - logT);
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Renumber them so that each XCD walks one contiguous eighth of the
// row list: neighbouring rows share B rows, and those then hit the same L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  constexpr int kXcd = 8;
  const int x = b % kXcd, k = b / kXcd;
  const int per = nb / kXcd, rem = nb % kXcd;
  return x * per + (x < rem ? x : rem) + k;
}

// Rows of a persistent window block.  The bin's row list is split into eight
// contiguous ranges, one per XCD (block b runs on XCD b % 8), sized by the
// XCD's block count; inside its range the blocks of one XCD take tiles of R
// rows in turn (group gi owns row gi of a tile).  At any moment the blocks
// of an XCD work on neighbouring rows, so the B rows they gather are shared
// in that XCD's L2.
struct RowWalk {
  int first, step, end;
};
__device__ __forceinline__ RowWalk row_walk(int nrows, int R, int gi) {
  constexpr int kXcd = 8;
  const int nb = gridDim.x, b = blockIdx.x;
  const int x = b % kXcd, k = b / kXcd;
  const int per = nb / kXcd, rem = nb % kXcd;
  const int nbx = per + (x < rem ? 1 : 0);
  const int before = x * per + (x < rem ? x : rem);
  const int x0 = (int)((int64_t)nrows * before / nb);
  const int x1 = (int)((int64_t)nrows * (before + nbx) / nb);
  return RowWalk{x0 + k * R + gi, nbx * R, x1};
}

__device__ __forceinline__ int64_t np2_i64(int64_t x) {
  if (x < 2) return 2;
  int64_t p = 2;
  while (p < x) p <<= 1;
  return p;
}

}  // namespace appt
