// kernels.hip -- gfx950 kernels of the SpGEMM pipeline.
//
// Pipeline (one call):  [transpose]  rowstats  prepass/bin  symbolic  scan
//                       numbin  numeric  [mirror]
//
// Value order: in the group bins (G <= 64 lanes per row) a row's A entries are
// visited in storage order, one at a time, and the G lanes split the B row;
// each output column is therefore touched by at most one lane per A entry and
// its fp64 sum is formed in exactly the order of the CPU Gustavson loop
// (acc = acc + a*b, product rounded first, via ds_add_f64).  Workgroup and
// heavy-row bins let the waves of a workgroup run ahead of each other, so
// their sums may be associated differently (bounded by the per-entry
// tolerance, not bit-equal).
#include <unordered_map>

#include "kernels.hpp"
#include "device_common.hpp"

#include <algorithm>
#include <type_traits>

namespace appt {

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------
// gate (optional) on a kernel: it runs only when *gate != 0 -- device-side
// control flow, no host synchronisation
__device__ __forceinline__ bool gated_off(const int32_t* gate) { return gate && *gate == 0; }

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v, const int32_t* gate) {
  if (gated_off(gate)) return;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s, const int32_t* gate) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v, gate);
}

__global__ void k_zero_u64x2(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0ull;
}

__global__ void k_fill3_i32(int32_t* a, int64_t na, int32_t* b, int64_t nb, int32_t* c, int64_t nc) {
  const int64_t tot = na + nb + nc;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    if (t < na) a[t] = 0;
    else if (t < na + nb) b[t - na] = 0;
    else c[t - na - nb] = 0;
  }
}
void set_dyn_lds(const void* fn, size_t bytes) {
  struct Key {
    const void* fn;
    int dev;
    bool operator==(const Key& o) const { return fn == o.fn && dev == o.dev; }
  };
  struct KeyHash {
    size_t operator()(const Key& k) const { return std::hash<const void*>()(k.fn) ^ (size_t)k.dev; }
  };
  thread_local std::unordered_map<Key, size_t, KeyHash> cap;
  int dev = 0;
  (void)hipGetDevice(&dev);
  auto it = cap.find(Key{fn, dev});
  if (it != cap.end() && it->second >= bytes) return;
  // a refused cap is not remembered: the launch that needs it then fails and
  // the stage's error check reports it
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) return;
  cap[Key{fn, dev}] = bytes;
}

void launch_fill3_i32(int32_t* a, int64_t na, int32_t* b, int64_t nb, int32_t* c, int64_t nc, hipStream_t s) {
  const int64_t tot = na + nb + nc;
  if (tot <= 0) return;
  int64_t blocks = (tot + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill3_i32, dim3((unsigned)blocks), dim3(256), 0, s, a, na, b, nb, c, nc);
}

__global__ void k_i64_to_i32(const int64_t* in, int64_t n, int32_t* out, const int32_t* gate) {
  if (gated_off(gate)) return;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (int32_t)in[t];
}
void launch_i64_to_i32(const int64_t* in, int64_t n, int32_t* out, hipStream_t s, const int32_t* gate) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_i64_to_i32, dim3((unsigned)blocks), dim3(256), 0, s, in, n, out, gate);
}
__global__ void k_copy_i64(const int64_t* in, int64_t n, int64_t* out) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    out[t] = in[t];
}
void launch_copy_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_copy_i64, dim3((unsigned)blocks), dim3(256), 0, s, in, n, out);
}

// ---------------------------------------------------------------------------
// B-row metadata {start, end, min column, max column}; 16 lanes per row.
// One 16-byte gather per A entry then serves the pre-pass and both product
// passes.
// ---------------------------------------------------------------------------
__global__ void k_rowstats(int32_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ ptr2,
                           const uint8_t* __restrict__ sel, int64_t off2, const int32_t* __restrict__ col,
                           int4* __restrict__ meta, int32_t* __restrict__ flag, int32_t min_len,
                           const int32_t* gate) {
  if (gated_off(gate)) return;
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  bool unsorted = false;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    int s, e;
    if (sel && sel[i]) {
      s = (int)(off2 + ptr2[i]);
      e = (int)(off2 + ptr2[i + 1]);
    } else {
      s = ptr[i];
      e = ptr[i + 1];
    }
    if (e - s <= min_len) continue;  // rows another kernel has described
    int mn = INT32_MAX, mx = -1;
    for (int p = s + lane; p < e; p += G) {
      const int c = col[p];
      mn = c < mn ? c : mn;
      mx = c > mx ? c : mx;
      if (p > s && col[p - 1] > c) unsorted = true;
    }
    mn = gmin<G>(mn);
    mx = gmax<G>(mx);
    if (lane == 0) meta[i] = make_int4(s, e, mn, mx);
  }
  // one plain store per wave that saw a descent (all writers store 1)
  if (__ballot(unsorted) != 0ull && threadIdx.x % kWave == 0) flag[0] = 1;
}
void launch_rowstats(int32_t n, const int32_t* ptr, const int32_t* ptr2, const uint8_t* sel, int64_t off2,
                     const int32_t* col, int4* meta, int32_t* flag, hipStream_t s, int32_t min_len,
                     const int32_t* gate) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_rowstats, dim3((unsigned)blocks), dim3(256), 0, s, n, ptr, ptr2, sel, off2, col, meta, flag,
                     min_len, gate);
}

// ---------------------------------------------------------------------------
// transpose B = A^T: column histogram, scan, atomic-cursor scatter.  The order
// inside a B row is not fixed; it does not affect any output value because a
// B row holds each column once (see the value-order note at the top).
// ---------------------------------------------------------------------------
__global__ void k_rows_strict(int32_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                              int32_t* __restrict__ flag, const int32_t* gate) {
  if (gated_off(gate)) return;
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb)
    for (int p = ptr[i] + lane; p + 1 < ptr[i + 1]; p += G) bad |= col[p] >= col[p + 1];
  if (__ballot(bad) != 0ull && threadIdx.x % kWave == 0) flag[0] = 1;
}
void launch_rows_strict(int32_t n, const int32_t* ptr, const int32_t* col, int32_t* flag, hipStream_t s,
                        const int32_t* gate) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_rows_strict, dim3((unsigned)blocks), dim3(256), 0, s, n, ptr, col, flag, gate);
}

__global__ void k_csr_check(int32_t n, int64_t nnz, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            int32_t* __restrict__ flag) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (int64_t)gridDim.x * blockDim.x;
  if (t0 == 0 && (ptr[0] != 0 || (int64_t)ptr[n] != nnz)) flag[0] = 1;
  bool bad = false;
  for (int64_t i = t0; i < n; i += st) bad |= ptr[i + 1] < ptr[i];
  if (__ballot(bad) != 0ull && threadIdx.x % kWave == 0) flag[1] = 1;
  bad = false;
  for (int64_t p = t0; p < nnz; p += st) bad |= (uint32_t)col[p] >= (uint32_t)n;
  if (__ballot(bad) != 0ull && threadIdx.x % kWave == 0) flag[2] = 1;
}
void launch_csr_check(int32_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, int32_t* flag, hipStream_t s) {
  const int64_t work = std::max<int64_t>((int64_t)n, nnz);
  int64_t blocks = (work + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_csr_check, dim3((unsigned)blocks), dim3(256), 0, s, n, nnz, ptr, col, flag);
}

__global__ void k_colcount(int64_t nnz, const int32_t* __restrict__ col, int32_t* __restrict__ cnt,
                           const int32_t* gate) {
  if (gated_off(gate)) return;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz;
       p += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[col[p]], 1);
}
void launch_colcount(int64_t nnz, const int32_t* col, int32_t* cnt, hipStream_t s, const int32_t* gate) {
  if (nnz <= 0) return;
  int64_t blocks = (nnz + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_colcount, dim3((unsigned)blocks), dim3(256), 0, s, nnz, col, cnt, gate);
}

// Symmetric-pattern transpose: when every row is strictly sorted and each
// stored (k, i) has a stored (i, k), B = A^T has A's pattern, so B row k is
// A row k with values a_ik looked up in row i: first at the mirrored index
// (entry d of row k is expected at index len_i - 1 - d of row i, exact for
// the interior rows of a symmetric stencil), else by binary search.  Any
// miss, duplicate or descent sets *asym and the scatter path below runs
// instead.
__global__ void k_transpose_gather(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                                   const double* __restrict__ av, int32_t* __restrict__ bc, double* __restrict__ bv,
                                   int32_t* __restrict__ asym) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t k = (int64_t)blockIdx.x * gpb + threadIdx.x / G; k < n; k += (int64_t)gridDim.x * gpb) {
    // a row that starts after an asymmetry was found is skipped; the flag is
    // read together with the row bounds (a select, not a branch), so its
    // round trip does not delay them
    const int stop = __hip_atomic_load(asym, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int s = ap[k], e0 = ap[k + 1];
    const int e = stop != 0 ? s : e0;
    bool miss = false;
    // two entries per lane per step (p and p + G), their dependent loads
    // issued together; the guessed mirror entry's value is loaded with its
    // column, so a row of up to 2G entries takes four dependent rounds
    for (int p0 = s + lane; p0 < e; p0 += 2 * G) {
      if (p0 - s >= 16 * G && (p0 - s) % (16 * G) == lane &&
          __hip_atomic_load(asym, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        break;
      int pp[2], i[2], prev[2], lo[2], hi[2], cg[2];
      bool in[2], gok[2];
      double vg[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        pp[u] = p0 + u * G;
        in[u] = pp[u] < e;
        i[u] = in[u] ? ac[pp[u]] : 0;
        prev[u] = (in[u] && pp[u] > s) ? ac[pp[u] - 1] : INT32_MIN;
      }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        lo[u] = in[u] ? ap[i[u]] : 0;
        hi[u] = in[u] ? ap[i[u] + 1] : 0;
      }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        // entry d of row k is expected at index len_i - 1 - d of row i
        const int g = hi[u] - 1 - (pp[u] - s);
        gok[u] = in[u] && g >= lo[u] && g < hi[u];
        cg[u] = gok[u] ? ac[g] : -1;
        vg[u] = gok[u] ? av[g] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if (!in[u]) continue;
        if (pp[u] > s && prev[u] >= i[u]) miss = true;   // descent or duplicate
        if (gok[u] && cg[u] == (int)k) {
          bc[pp[u]] = i[u];
          bv[pp[u]] = vg[u];
          continue;
        }
        int l = lo[u], h = hi[u];
        while (l < h) {
          const int mid = (int)(((unsigned)l + (unsigned)h) >> 1);
          if (ac[mid] < (int)k) l = mid + 1;
          else h = mid;
        }
        if (l < hi[u] && ac[l] == (int)k) {
          bc[pp[u]] = i[u];
          bv[pp[u]] = av[l];
        } else {
          miss = true;
        }
      }
    }
    if (miss) *asym = 1;   // every writer stores 1
  }
}
// B's row pointers are A's when the gather succeeded (*asym == 0)
__global__ void k_rowptr_if_symmetric(int32_t n, const int32_t* __restrict__ ap, int32_t* __restrict__ tp,
                                      const int32_t* __restrict__ asym) {
  if (*asym != 0) return;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= n; k += (int64_t)gridDim.x * blockDim.x)
    tp[k] = ap[k];
}

// gate (optional): the scatter and the row sort run only when *gate != 0
__global__ void k_transpose_scatter(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                                    const double* __restrict__ av, int32_t* __restrict__ cursor,
                                    int32_t* __restrict__ bc, double* __restrict__ bv,
                                    const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    const int s = ap[i], e = ap[i + 1];
    for (int p = s + lane; p < e; p += G) {
      const int k = ac[p];
      const int pos = atomicAdd(&cursor[k], 1);
      bc[pos] = (int32_t)i;
      bv[pos] = av[p];
    }
  }
}
void launch_transpose_scatter(int32_t n, const int32_t* ap, const int32_t* ac, const double* av,
                              int32_t* cursor, int32_t* bc, double* bv, const int32_t* gate, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_transpose_scatter, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, ac, av, cursor,
                     bc, bv, gate);
}

// ---------------------------------------------------------------------------
// exclusive scan int32 -> int64 (reduce / scan partials / apply)
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 2; }

__global__ void k_scan_reduce(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ part,
                              const int32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ int64_t scratch[17];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t sum = 0;
#pragma unroll
  for (int t = 0; t < kScanItems; t++) {
    const int64_t idx = base + (int64_t)t * kScanThreads + threadIdx.x;
    if (idx < n) sum += in[idx];
  }
  int64_t total;
  block_scan_excl<int64_t>(sum, scratch, &total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ void k_scan_partials(int64_t* part, int64_t nb, const int32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ int64_t scratch[17];
  int64_t carry = 0;
  for (int64_t base = 0; base < nb; base += blockDim.x) {
    const int64_t idx = base + threadIdx.x;
    const int64_t v = idx < nb ? part[idx] : 0;
    int64_t total;
    const int64_t ex = block_scan_excl<int64_t>(v, scratch, &total);
    if (idx < nb) part[idx] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) part[nb] = carry;
}

__global__ void k_scan_apply(const int32_t* __restrict__ in, int64_t n, const int64_t* __restrict__ part,
                             int64_t nb, int64_t* __restrict__ out64, int32_t* __restrict__ out32,
                             const int32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ int64_t scratch[17];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int64_t sum = 0;
#pragma unroll
  for (int t = 0; t < kScanItems; t++) {
    const int64_t idx = base + t;
    v[t] = idx < n ? in[idx] : 0;
    sum += v[t];
  }
  int64_t total;
  int64_t run = block_scan_excl<int64_t>(sum, scratch, &total) + part[blockIdx.x];
#pragma unroll
  for (int t = 0; t < kScanItems; t++) {
    const int64_t idx = base + t;
    if (idx < n) {
      out64[idx] = run;
      if (out32) out32[idx] = (int32_t)run;
    }
    run += v[t];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out64[n] = part[nb];
    if (out32) out32[n] = (int32_t)part[nb];
  }
}

// Small inputs: one block walks the tiles in order with a running carry
// (one launch instead of three).
constexpr int64_t kScanSmallTiles = 8;
__global__ void __launch_bounds__(kScanThreads) k_scan_small(const int32_t* __restrict__ in, int64_t n,
                                                             int64_t* __restrict__ out64, int32_t* __restrict__ out32,
                                                             const int32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ int64_t scratch[17];
  int64_t carry = 0;
  for (int64_t tb = 0; tb < n; tb += kScanTile) {
    const int64_t base = tb + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int64_t sum = 0;
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
      const int64_t idx = base + t;
      v[t] = idx < n ? in[idx] : 0;
      sum += v[t];
    }
    int64_t total;
    int64_t run = block_scan_excl<int64_t>(sum, scratch, &total) + carry;
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
      const int64_t idx = base + t;
      if (idx < n) {
        out64[idx] = run;
        if (out32) out32[idx] = (int32_t)run;
      }
      run += v[t];
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    out64[n] = carry;
    if (out32) out32[n] = (int32_t)carry;
  }
}

void launch_scan(const int32_t* in, int64_t n, int64_t* out64, int32_t* out32, int64_t* tmp,
                 hipStream_t s, const int32_t* gate) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 0) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int32_t*)out64, (int64_t)2, 0, gate);
    if (out32) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, out32, (int64_t)1, 0, gate);
    return;
  }
  if (nb <= kScanSmallTiles) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kScanThreads), 0, s, in, n, out64, out32, gate);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, tmp, gate);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, tmp, nb, gate);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, tmp, nb, out64, out32,
                     gate);
}

// ---------------------------------------------------------------------------
// Radix transpose: a stable LSD counting sort of A's entries by column, at
// most 8 bits per pass.  Entries come in row-major order, so a stable sort
// leaves each column's entries (B's row) ordered by row id.  One block per
// tile of kRadixTile entries: a per-tile digit histogram (digit-major over
// tiles), one exclusive scan of all histograms, then each block sorts its
// tile in LDS -- every wave walks its quarter in order, ranking equal digits
// inside a 64-entry chunk with ballots (lanes of equal digit found bit by
// bit) and a per-digit running slot -- and copies each digit's run out to
// consecutive addresses.  No global atomics; B's row pointers are the
// boundaries of the sorted keys.
// ---------------------------------------------------------------------------
int64_t radix_tiles(int64_t nnz) { return (nnz + kRadixTile - 1) / kRadixTile; }
static int key_bits(int32_t n) { return n > 1 ? 32 - __builtin_clz((unsigned)(n - 1)) : 1; }
int radix_passes(int32_t n) { return (key_bits(n) + kRadixMaxBits - 1) / kRadixMaxBits; }
int radix_bits(int32_t n) {
  const int p = radix_passes(n);
  return (key_bits(n) + p - 1) / p;
}

__global__ void k_row_ids(int32_t n, const int32_t* __restrict__ ap, int32_t* __restrict__ rid,
                          const int32_t* gate) {
  if (gated_off(gate)) return;
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb)
    for (int p = ap[i] + lane; p < ap[i + 1]; p += G) rid[p] = (int32_t)i;
}

constexpr int kRadixBuckets = 1 << kRadixMaxBits;

__global__ void __launch_bounds__(256) k_radix_hist(int64_t nnz, const int32_t* __restrict__ keys, int shift, int bits,
                                                    int64_t ntiles, int32_t* __restrict__ hist, const int32_t* gate) {
  if (gated_off(gate)) return;
  __shared__ int32_t cnt[kRadixBuckets];
  const int64_t tile = blockIdx.x;
  const int nb = 1 << bits;
  const uint32_t dm = (uint32_t)nb - 1u;
  for (int d = threadIdx.x; d < nb; d += blockDim.x) cnt[d] = 0;
  __syncthreads();
  const int64_t base = tile * kRadixTile;
  for (int t = threadIdx.x; t < kRadixTile; t += blockDim.x) {
    const int64_t e = base + t;
    if (e < nnz) atomicAdd(&cnt[((uint32_t)keys[e] >> shift) & dm], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nb; d += blockDim.x) hist[(int64_t)d * ntiles + tile] = cnt[d];
}

// One block per tile: (A) digit counts per wave over its kRadixChunks chunks,
// (B) stable ranks (wave order, then chunk, then lane: the input order) place
// every entry at its digit-sorted slot of the tile in LDS, (C) the block
// copies the sorted tile out, each digit's run to consecutive addresses.
__global__ void __launch_bounds__(256) k_radix_scatter(int64_t nnz, int shift, int bits, int64_t ntiles,
                                                       const int64_t* __restrict__ hoff,
                                                       const int32_t* __restrict__ kin, const int32_t* __restrict__ rin,
                                                       const double* __restrict__ vin, int32_t* __restrict__ kout,
                                                       int32_t* __restrict__ rout, double* __restrict__ vout,
                                                       const int32_t* gate) {
  if (gated_off(gate)) return;
  constexpr int T = (int)kRadixTile, NB = kRadixBuckets;
  // staging slots are skewed by one per 32 (slot + slot / 32): digit runs of
  // clustered keys start a multiple of 32 slots apart and would otherwise
  // land in one LDS bank
  constexpr int TS = T + T / 32;
  __shared__ int32_t sk[TS];
  __shared__ int32_t sr[TS];
  __shared__ double sv[TS];
  __shared__ int32_t wcnt[4][NB];   // per-wave digit counts, then running slots
  __shared__ int32_t bstart[NB];    // first slot of each digit in the tile
  __shared__ int64_t gbase[NB];     // global position of each digit's run
  __shared__ int64_t red[17];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t tile = blockIdx.x;
  const int nb = 1 << bits;
  const uint32_t dm = (uint32_t)nb - 1u;
  const int64_t base = tile * kRadixTile;
  const int64_t wbase = base + (int64_t)w * kWave * kRadixChunks;
  for (int d = threadIdx.x; d < 4 * NB; d += blockDim.x) (&wcnt[0][0])[d] = 0;
  // every entry of the wave's chunks is loaded once, up front (all loads in
  // flight together), and kept in registers through the ranking
  int32_t kr[kRadixChunks], rr[kRadixChunks];
  double vr[kRadixChunks];
#pragma unroll
  for (int c = 0; c < kRadixChunks; c++) {
    const int64_t e = wbase + (int64_t)c * kWave + lane;
    const bool valid = e < nnz;
    kr[c] = valid ? kin[e] : 0;
    rr[c] = valid && rin ? rin[e] : 0;
    vr[c] = valid ? vin[e] : 0.0;
  }
  __syncthreads();
  // (A) per-wave counts
#pragma unroll
  for (int c = 0; c < kRadixChunks; c++) {
    const int64_t e = wbase + (int64_t)c * kWave + lane;
    if (e < nnz) atomicAdd(&wcnt[w][((uint32_t)kr[c] >> shift) & dm], 1);
  }
  __syncthreads();
  // digit starts in the tile, and each wave's first slot per digit
  for (int d0 = 0; d0 < NB; d0 += blockDim.x) {
    const int d = d0 + threadIdx.x;
    int tot = 0;
    if (d < nb) {
      for (int v = 0; v < 4; v++) {
        const int x = wcnt[v][d];
        wcnt[v][d] = tot;
        tot += x;
      }
      gbase[d] = hoff[(int64_t)d * ntiles + tile];
    }
    int64_t all;
    const int64_t st = block_scan_excl<int64_t>((int64_t)tot, red, &all);
    if (d < nb) {
      bstart[d] = (int32_t)st;
      for (int v = 0; v < 4; v++) wcnt[v][d] += (int32_t)st;
    }
  }
  __syncthreads();
  // (B) stable ranks inside the wave, placed in LDS
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int c = 0; c < kRadixChunks; c++) {
    const int64_t e = wbase + (int64_t)c * kWave + lane;
    const bool valid = e < nnz;
    const int32_t key = kr[c];
    const uint32_t digit = ((uint32_t)key >> shift) & dm;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; b++) {
      const bool bit = (digit >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const int rank = __popcll(peers & below);
    const int slot = wcnt[w][digit] + rank;
    gsync<kWave>();  // every lane has read its digit's slot before it moves
    if (valid && rank == 0) wcnt[w][digit] += __popcll(peers);
    if (valid) {
      const int ps = slot + (slot >> 5);
      sk[ps] = key;
      sr[ps] = rr[c];
      sv[ps] = vr[c];
    }
    gsync<kWave>();
  }
  __syncthreads();
  // (C) copy out in slot order
  const int valid_n = (int)((nnz - base) < (int64_t)T ? (nnz - base) : (int64_t)T);
  for (int q = threadIdx.x; q < valid_n; q += blockDim.x) {
    const int ps = q + (q >> 5);
    const int32_t key = sk[ps];
    const uint32_t digit = ((uint32_t)key >> shift) & dm;
    const int64_t pos = gbase[digit] + (q - bstart[digit]);
    kout[pos] = key;
    if (rout) rout[pos] = sr[ps];
    vout[pos] = sv[ps];
  }
}

// B's row pointers from the sorted keys: tp[k] = first position whose key is
// >= k (positions p in [0, nnz] fill the columns since the previous key)
__global__ void k_radix_rowptr(int32_t n, int64_t nnz, const int32_t* __restrict__ keys, int32_t* __restrict__ tp,
                               const int32_t* gate) {
  if (gated_off(gate)) return;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p <= nnz; p += (int64_t)gridDim.x * blockDim.x) {
    const int kp = p < nnz ? keys[p] : n;
    const int kq = p > 0 ? keys[p - 1] : -1;
    for (int k = kq + 1; k <= kp; k++) tp[k] = (int32_t)p;
  }
}

// The radix passes over `count` entries with keys in [0, n): input (k0, r0,
// v0), ping-pong between the x and y arrays, ending in the y arrays (r*
// null: no row payload).  The input may share storage only with the arrays
// the first pass does not write: the x arrays when the pass count is odd,
// the y arrays when it is even.
static void radix_passes_run(int32_t n, int64_t count, const int32_t* k0, const int32_t* r0, const double* v0,
                             int32_t* xk, int32_t* xr, double* xv, int32_t* yk, int32_t* yr, double* yv,
                             int32_t* hist, int64_t* hoff, int64_t* htmp, const int32_t* gate, hipStream_t s) {
  const int P = radix_passes(n), bits = radix_bits(n);
  const int64_t nt = radix_tiles(count);
  if (nt <= 0) return;
  const unsigned tb = (unsigned)nt;  // one block per tile
  const int32_t* sk = k0;
  const int32_t* sr = r0;
  const double* sv = v0;
  for (int p = 0; p < P; p++) {
    const bool to_y = (P - 1 - p) % 2 == 0;
    int32_t* dk = to_y ? yk : xk;
    int32_t* dr = to_y ? yr : xr;
    double* dv = to_y ? yv : xv;
    hipLaunchKernelGGL(k_radix_hist, dim3(tb), dim3(256), 0, s, count, sk, p * bits, bits, nt, hist, gate);
    launch_scan(hist, ((int64_t)1 << bits) * nt, hoff, nullptr, htmp, s, gate);
    hipLaunchKernelGGL(k_radix_scatter, dim3(tb), dim3(256), 0, s, count, p * bits, bits, nt, (const int64_t*)hoff,
                       sk, sr, sv, dk, dr, dv, gate);
    sk = dk;
    sr = dr;
    sv = dv;
  }
}

void launch_radix_pairs(int32_t n, int64_t count, const RadixScratch& rx, int32_t* rows2, double* vals2,
                        hipStream_t s) {
  if (n <= 0 || count <= 0) return;
  const bool odd = radix_passes(n) % 2 == 1;
  const int32_t* k0 = odd ? rx.keys : rx.keys2;
  const int32_t* r0 = odd ? rx.rows : rows2;
  const double* v0 = odd ? rx.vals : vals2;
  radix_passes_run(n, count, k0, r0, v0, rx.keys, rx.rows, rx.vals, rx.keys2, rows2, vals2, rx.hist, rx.hoff,
                   rx.htmp, nullptr, s);
}

void launch_transpose_gather(int32_t n, const int32_t* ap, const int32_t* ac, const double* av, int32_t* tp,
                             int32_t* bc, double* bv, int32_t* asym, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_transpose_gather, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, ac, av, bc, bv, asym);
  int64_t b0 = ((int64_t)n + 256) / 256;
  if (b0 > 4096) b0 = 4096;
  hipLaunchKernelGGL(k_rowptr_if_symmetric, dim3((unsigned)b0), dim3(256), 0, s, n, ap, tp, asym);
}

void launch_transpose_general(int32_t n, int64_t nnz, const int32_t* ap, const int32_t* ac, const double* av,
                              int32_t* tp, int32_t* bc, double* bv, int32_t* cnt, int64_t* off64, int64_t* tmp,
                              int32_t* slist, const RadixScratch* rx, int32_t* asym, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (rx) {
    // asymmetric, large: radix passes; the last one writes (keys2, bc, bv).
    // The row ids go to the rows buffer the first pass does not write.
    int32_t* rid = (radix_passes(n) % 2 == 1) ? rx->rows : bc;
    hipLaunchKernelGGL(k_row_ids, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, rid, asym);
    radix_passes_run(n, nnz, ac, rid, av, rx->keys, rx->rows, rx->vals, rx->keys2, bc, bv, rx->hist, rx->hoff,
                     rx->htmp, asym, s);
    int64_t b1 = (nnz + 256) / 256;
    if (b1 > 8192) b1 = 8192;
    hipLaunchKernelGGL(k_radix_rowptr, dim3((unsigned)b1), dim3(256), 0, s, n, nnz, (const int32_t*)rx->keys2, tp,
                       asym);
    return;
  }
  // asymmetric, small: column histogram, scan, atomic-cursor scatter, row sort
  launch_fill_i32(cnt, n, 0, s, asym);
  launch_colcount(nnz, ac, cnt, s, asym);
  launch_scan(cnt, n, off64, tp, tmp, s, asym);
  launch_i64_to_i32(off64, n, cnt, s, asym);
  launch_transpose_scatter(n, ap, ac, av, cnt, bc, bv, asym, s);
  launch_sort_rows(n, tp, bc, bv, slist, asym, s);
}

void launch_transpose(int32_t n, int64_t nnz, const int32_t* ap, const int32_t* ac, const double* av,
                      int32_t* tp, int32_t* bc, double* bv, int32_t* cnt, int64_t* off64, int64_t* tmp,
                      int32_t* slist, const RadixScratch* rx, int32_t* asym, hipStream_t s) {
  launch_transpose_gather(n, ap, ac, av, tp, bc, bv, asym, s);
  launch_transpose_general(n, nnz, ap, ac, av, tp, bc, bv, cnt, off64, tmp, slist, rx, asym, s);
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t span_words(int lo, int hi) {
  return hi < lo ? 0 : (int64_t)((hi - lo) >> 5) + 1;
}

constexpr int kTouchCap = 64;   // touched-word list length per row
// products bound of the narrow-span symbolic rows kept wave-level, as a
// multiple of the last class's bound
constexpr int kSymStretch = 4;
// rowwords[i] holds the row's distinct-word count; kNoWordList marks rows whose
// symbolic pass left no sorted word list behind
constexpr int32_t kNoWordList = 1 << 30, kWordsMask = kNoWordList - 1;

template <typename T, bool W64>
__device__ __forceinline__ T ld_idx(const T* base, uint32_t idx) {
  if constexpr (W64) return base[idx];
  else return *(const T*)((const char*)base + (size_t)(idx * (uint32_t)sizeof(T)));
}

__host__ __device__ __forceinline__ int64_t align16(int64_t x) { return (x + 15) & ~(int64_t)15; }

// The symbolic window's bitmap is stored skewed by one LDS bank per 32
// words (word w at w + w/32): words a multiple of 32 apart -- the plane
// offsets of a 3-D stencil, 288 words on a 96^3 grid -- then fall in
// different banks instead of serialising the group's atomic ORs.
__host__ __device__ __forceinline__ uint32_t bank_skew(uint32_t w) { return w + (w >> 5); }
__host__ __device__ __forceinline__ int64_t skewed_words(int64_t W) { return W + (W >> 5) + 1; }
// symbolic window LDS per group: meta (G int2) | bits (skewed W, then G dummy
// u32) | touched (TL u32) | ctl (4 int)
__host__ __device__ __forceinline__ int64_t sym_window_bytes(int G, int64_t W) {
  return align16(G * 8) + align16((skewed_words(W) + G) * 4) + align16(kTouchCap * 4) + 16;
}
// numeric window LDS per group: meta (G int2 + G f64) | bits (W + G u32) | pre (W u16) |
// vals (Cm + G f64) | cols (Cm i32)
// Longest word list a numeric group of G lanes takes (G >= 16: the full
// list; 8-lane groups hold at most 4 entries per lane in registers).
__host__ __device__ __forceinline__ constexpr int list_cap(int G) { return G >= 16 ? kWordListCap : 32; }
// numeric list-window LDS per group: meta (G int2 + G f64) | word index (W + G u8) |
// listed words (list_cap(G) x (mask, first position)) | vals (Cm + G f64)
__host__ __device__ __forceinline__ int64_t num_list_bytes(int G, int64_t W, int64_t Cm) {
  return align16(G * 16) + align16(W + G) + align16(list_cap(G) * 8) + align16((Cm + G) * 8);
}
__host__ __device__ __forceinline__ int64_t num_window_bytes(int G, int64_t W, int64_t Cm) {
  return align16(G * 16) + align16((W + G) * 4) + align16(W * 2) + align16((Cm + G) * 8) + align16(Cm * 4);
}

__host__ __device__ __forceinline__ int64_t np2_hd(int64_t x) {
  if (x <= 2) return 2;
  return (int64_t)1 << (64 - __builtin_clzll((unsigned long long)(x - 1)));
}
// Lanes per row from the mean B-row length f / nA (compared by multiplication:
// f <= g * nA  <=>  ceil(f / nA) <= g).
__device__ __forceinline__ int group_index(const BinParams& P, int64_t f, int64_t nA) {
  return f <= (int64_t)P.glen_max[0] * nA ? 0 : f <= (int64_t)P.glen_max[1] * nA ? 1
       : f <= (int64_t)P.glen_max[2] * nA ? 2 : 3;
}

// Register-sort bin of a row with f products over nA A entries: slots
// N = np2(max(f, nA, 16)) <= P.sort_max, or -1.
__device__ __forceinline__ int sort_bin_of(const BinParams& P, int64_t f, int64_t nA) {
  const int64_t m = f > nA ? f : nA;
  if (m > P.sort_max) return -1;
  int b = kBinSort0;
  while ((int64_t)(16 << (b - kBinSort0)) < m) b++;
  return b;
}

// Symbolic bin from products f, span words and mean B-row length L.  *stat
// receives the LDS sizing statistic: span words (dense) or key bound (hash).
// Rows whose span does not fit a dense window go to a register-sort bin
// when they have at most kSortMax products, else to the word hash.
constexpr int64_t kWgMaxProducts = 65536;      // workgroup symbolic rows: at most this many products,
constexpr int64_t kWgWideWords = 8192;         // and over spans of more than this many words
constexpr int64_t kWgWideMaxProducts = 4096;   // at most this many
__device__ __forceinline__ int sym_bin_of(const BinParams& P, int64_t f, int64_t words, int64_t nA, int64_t* stat) {
  *stat = 0;
  if (f <= 0 || words <= 0) return 0;
  const int gidx = group_index(P, f, nA);
  const int64_t keys = f < words ? f : words;
  const int64_t hash_bytes = np2_hd((keys * 4) / 3 + 1) * 8;
  const int sb = sort_bin_of(P, f, nA);
  // sort bins: the statistic is the A-row length (sizes the entry arrays)
  if (sb >= 0 && P.sort_ratio > 0 && words > (int64_t)P.sort_ratio * f) { *stat = nA; return sb; }
  for (int c = 0; c < P.ncls; c++)
    if (f <= P.cls_max[c]) {
      if (sym_window_bytes(8 << gidx, words) <= P.cls_bytes[c]) { *stat = words; return 1 + (c * 2 + 0) * 4 + gidx; }
      if (sb >= 0) { *stat = nA; return sb; }
      if (hash_bytes <= P.cls_bytes[c]) { *stat = keys; return 1 + (c * 2 + 1) * 4 + gidx; }
      break;
    }
  if (sb >= 0) { *stat = nA; return sb; }
  // long rows over a narrow span (banded operands): one 64-lane group of the
  // last class per row rather than a workgroup per row, and the row keeps its
  // word list (every touched word fits the list), which the numeric pass uses
  const int lc = P.ncls - 1;
  if (f <= kSymStretch * (int64_t)P.cls_max[lc] && words <= kTouchCap &&
      sym_window_bytes(64, words) <= P.cls_bytes[lc]) {
    *stat = words;
    return 1 + (lc * 2 + 0) * 4 + 3;
  }
  // rows of many products take the heavy kernel even when their span fits a
  // workgroup: its products are dealt over 1024 lanes, several in flight
  // each, where a workgroup row walks each short B row with one lane; for
  // wide spans the workgroup bitmap's clear adds to that (R-MAT 2^20 A*A
  // row-pointer plan 581 -> 101 ms, 2^16 A*A 7.80 -> 7.38 ms)
  if (f <= kWgMaxProducts && (words <= kWgWideWords || f <= kWgWideMaxProducts)) {
    if (words * 4 <= P.wg_bytes) { *stat = words; return kBinWgBitmap; }
    if (hash_bytes <= P.wg_bytes) { *stat = keys; return kBinWgHash; }
  }
  *stat = words;
  return kBinHeavy;
}

__host__ __device__ __forceinline__ int64_t whash_bytes_hd(int64_t w, int64_t c) {
  const int64_t T = np2_hd((w * 3) / 2 + 1);
  const int64_t v = c > T ? c : T;
  return ((T * 12 + v * 8 + c * 4) + 15) & ~(int64_t)15;
}

// Numeric bin from row nnz c, span words, distinct words w and L.
__device__ __forceinline__ int num_bin_of(const BinParams& P, int64_t c, int64_t words, int64_t w, int64_t f,
                                          int64_t nA, int64_t listed, int64_t* stat) {
  *stat = 0;
  if (c <= 0 || words <= 0) return 0;
  const int gidx = group_index(P, f, nA);
  const int64_t lim = (int64_t)P.dense_ratio * w;
  const bool dense_ok = words <= (lim > 64 ? lim : 64);
  const int sb = sort_bin_of(P, f, nA);
  // scattered: under 4 columns per 3 distinct words (clustered rows, such as
  // stencils, keep their windows however wide the span)
  if (sb >= 0 && P.sort_ratio > 0 && words > (int64_t)P.sort_ratio * f && 4 * w > 3 * c) { *stat = nA; return sb; }
  for (int cl = 0; cl < P.ncls; cl++)
    if (c <= P.cls_max[cl]) {
      // listed: the row's word-list length, or -1 without a list
      if (listed >= 0 && listed <= list_cap(8 << gidx) && num_list_bytes(8 << gidx, words, c) <= P.cls_bytes[cl]) {
        *stat = words;
        return kBinList0 + cl * 4 + gidx;
      }
      if (num_window_bytes(8 << gidx, words, c) <= P.cls_bytes[cl]) { *stat = words; return 1 + (cl * 2 + 0) * 4 + gidx; }
      if (sb >= 0) { *stat = nA; return sb; }
      if (whash_bytes_hd(w, c) <= P.cls_bytes[cl]) { *stat = w; return 1 + (cl * 2 + 1) * 4 + gidx; }
      break;
    }
  if (sb >= 0) { *stat = nA; return sb; }
  // workgroup bins: bound each term separately so bin maxima still fit
  if (dense_ok && words * 8 <= P.wg_bytes / 2 && c * 12 <= P.wg_bytes / 2) { *stat = words; return kBinWgBitmap; }
  if (whash_bytes_hd(c, c) <= P.wg_bytes) { *stat = w; return kBinWgHash; }
  *stat = words;
  return kBinHeavy;
}

// Block-local bin statistics flushed with one global atomic per bin.
struct BinStatLds {
  int32_t v[kNumBins * kStatFields];
};
__device__ __forceinline__ void stats_init(BinStatLds& st) {
  for (int t = threadIdx.x; t < kNumBins * kStatFields; t += blockDim.x) st.v[t] = 0;
  __syncthreads();
}
// Wave-aggregated update: every active lane passes its bin (or -1); one LDS
// atomic per distinct bin in the wave.
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) { const int32_t o = __shfl_xor(v, m, 64); v = o > v ? o : v; }
  return v;
}
__device__ __forceinline__ void stats_add_wave(BinStatLds& st, int bin, int64_t metric, int64_t words) {
  const int lane = threadIdx.x % kWave;
  const int32_t m32 = (int32_t)(metric > INT32_MAX ? INT32_MAX : metric);
  const int32_t w32 = (int32_t)(words > INT32_MAX ? INT32_MAX : words);
  uint64_t todo = __ballot(bin >= 0);
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const int lb = __shfl(bin, leader, kWave);
    const uint64_t same = __ballot(bin == lb);
    const int32_t mm = wave_max_i32(bin == lb ? m32 : 0);
    const int32_t ww = wave_max_i32(bin == lb ? w32 : 0);
    if (lane == leader) {
      atomicAdd(&st.v[lb * kStatFields + 0], (int32_t)__popcll(same));
      atomicMax(&st.v[lb * kStatFields + 1], mm);
      atomicMax(&st.v[lb * kStatFields + 2], ww);
    }
    todo &= ~same;
  }
}
// Per-block statistics go to a field-major partial array part[field * nb + b]
// (and the block's product count to prod[b]); k_stats_reduce combines them.
// No same-address global atomics: thousands of blocks updating one counter
// serialise in the L2.
__device__ __forceinline__ void stats_flush(BinStatLds& st, int32_t* part, unsigned long long* prod,
                                           unsigned long long my_prod) {
  __syncthreads();
  const int nb = gridDim.x;
  for (int t = threadIdx.x; t < kNumBins * kStatFields; t += blockDim.x) part[(size_t)t * nb + blockIdx.x] = st.v[t];
  if (prod && threadIdx.x == 0) prod[blockIdx.x] = my_prod;
}

__global__ void __launch_bounds__(256) k_stats_reduce(const int32_t* __restrict__ part,
                                                      const unsigned long long* __restrict__ prod, int nb,
                                                      int32_t* __restrict__ gstats,
                                                      unsigned long long* __restrict__ total) {
  __shared__ int64_t red[17];
  const int f = blockIdx.x;
  if (f < kNumBins * kStatFields) {
    const bool is_sum = (f % kStatFields) == 0;
    int64_t v = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      const int64_t x = part[(size_t)f * nb + b];
      v = is_sum ? v + x : (x > v ? x : v);
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      const int64_t o = __shfl_xor(v, m, 64);
      v = is_sum ? v + o : (o > v ? o : v);
    }
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t r = 0;
      for (int w = 0; w < (int)blockDim.x / kWave; w++) r = is_sum ? r + red[w] : (red[w] > r ? red[w] : r);
      gstats[f] = (int32_t)(r > INT32_MAX ? INT32_MAX : r);
    }
  } else if (prod) {
    unsigned long long v = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += prod[b];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = (int64_t)v;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long r = 0;
      for (int w = 0; w < (int)blockDim.x / kWave; w++) r += (unsigned long long)red[w];
      *total = r;
    }
  }
}

static void stats_reduce(const int32_t* part, const unsigned long long* prod, int nb, int32_t* gstats,
                         unsigned long long* total, hipStream_t s) {
  const int fields = kNumBins * kStatFields + (prod ? 1 : 0);
  hipLaunchKernelGGL(k_stats_reduce, dim3(fields), dim3(256), 0, s, part, prod, nb, gstats, total);
}

int64_t stats_part_elems() { return (int64_t)kNumBins * kStatFields * kStatsMaxBlocks; }

// Structural pre-pass: products f_i (k_flops of the reference), output span
// [lo, hi] from the B rows' min/max columns, symbolic bin, bin statistics.
constexpr int kPrepassLanes = 4, kPrepassLong = 64;
constexpr int kPrepassCheckU = 8;   // template-check entries per lane per step (32 per row)
__global__ void k_prepass(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                          const int4* __restrict__ bmeta, int upper, BinParams P,
                          int64_t* __restrict__ rowf, int32_t* __restrict__ rowlo, int32_t* __restrict__ rowhi,
                          uint8_t* __restrict__ bin, int32_t* __restrict__ part,
                          unsigned long long* __restrict__ prodpart, TplArgs T, int tpl_on,
                          int32_t* __restrict__ rowcnt, int32_t* __restrict__ rowwords) {
  // kPrepassLanes lanes per row (more rows in flight, short reductions);
  // rows of more than kPrepassLong entries are left to the whole wave
  constexpr int G = kPrepassLanes;
  __shared__ BinStatLds st;
  __shared__ unsigned long long red[16];
  stats_init(st);
  const int lane = threadIdx.x % G, wl = threadIdx.x % kWave;
  const int64_t gpb = blockDim.x / G;
  unsigned long long myprod = 0;
  (void)upper;
  // one row's bin from its reduced products and span; `mine`: the lane that
  // writes the row
  // tpl: the row's checked template slot, or -1
  auto finish = [&](int64_t i, int64_t f, int lo, int hi, int nA, bool mine, int tpl) {
    // bitmap windows start on a 32-column boundary, so a window word and a
    // word-hash key (j >> 5) cover the same columns and distinct-word counts
    // agree between accumulator modes
    if (lo <= hi) lo &= ~31;
    const int64_t words = span_words(lo, hi);
    int64_t stat;
    int b;
    if (tpl >= 0) {
      // template rows: the count is the template's slot count, no symbolic
      // pass; the statistics carry the slot count and the longest B row
      const int4 inf = T.info[tpl];
      b = kBinTemplate;
      stat = inf.w;
      if (mine) {
        rowcnt[i] = inf.z;
        rowwords[i] = kNoWordList;
      }
    } else {
      b = sym_bin_of(P, f, words, (int64_t)nA, &stat);
    }
    if (mine) {
      if (tpl_on) T.tplid[i] = tpl;
      rowf[i] = f;
      rowlo[i] = lo;
      rowhi[i] = hi;
      bin[i] = (uint8_t)b;
      myprod += (unsigned long long)f;
    }
    return std::pair<int, int64_t>(b, tpl >= 0 ? (int64_t)T.info[tpl].z : stat);
  };
  // every wave runs the same number of iterations so the wave-wide statistics
  // updates below are executed by converged waves
  const int64_t stride = (int64_t)gridDim.x * gpb;
  const int64_t i0 = (int64_t)blockIdx.x * gpb + threadIdx.x / G;
  const int64_t wave_first = (int64_t)blockIdx.x * gpb + (threadIdx.x / kWave) * (kWave / G);
  for (int64_t wi = wave_first; wi < n; wi += stride) {
    const int64_t i = i0 + (wi - wave_first);
    const bool live = i < n;
    int64_t f = 0;
    int lo = INT32_MAX, hi = -1, s = 0, e = 0, tpl = -1, acl = -1;
    if (live) {
      s = ap[i];
      e = ap[i + 1];
      if (tpl_on) {
        tpl = T.tplid[i];
        acl = T.aclass[i];
      }
    }
    const bool lng = live && e - s > kPrepassLong;
    if (lng || acl < 0) tpl = -1;
    // a template candidate (template rows have at most kTplMaxLen <= kPrepassLong
    // entries): valid template of this row's length, same A class as its
    // representative, and the same B class at every entry (classes set: a
    // class check may clear a class after the template pass read it).  The
    // checks' loads depend on the row's template and entries only, so they
    // go out together: three dependent rounds (row, template and entries,
    // B classes and the representative's class).  The products are the
    // template's, the span unused.
    int same = 1;
    int64_t tf = 0;
    if (tpl >= 0) {
      const int m = e - s;
      const int4 inf = T.info[tpl];
      const int rep = T.trep[tpl];
      tf = T.posoff[(size_t)tpl * (kTplMaxLen + 1) + m];
      const int32_t* tup = T.tuple + (size_t)tpl * kTplMaxLen;
      constexpr int U = kPrepassCheckU;
      for (int u0 = lane; u0 < m; u0 += U * G) {
        int k[U], tv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const bool in = u0 + u * G < m;
          k[u] = in ? ac[s + u0 + u * G] : -1;
          tv[u] = in ? tup[u0 + u * G] : -1;
        }
        int bcl[U];
#pragma unroll
        for (int u = 0; u < U; u++) bcl[u] = k[u] >= 0 ? T.bclass[k[u]] : -1;
#pragma unroll
        for (int u = 0; u < U; u++)
          if (k[u] >= 0) same &= (bcl[u] >= 0 && bcl[u] == tv[u]) ? 1 : 0;
      }
      same &= (inf.x != 0 && inf.y == m && T.aclass[rep] == acl) ? 1 : 0;
    }
    same = gmin<G>(same);
    if (!same) tpl = -1;
    if (tpl >= 0) {
      if (lane == 0) f = tf;
    } else if (live && !lng) {
      // 4 entries per lane per step: their column and B-row loads in flight together
      for (int p0 = s + lane; p0 < e; p0 += 4 * G) {
        int k[4];
#pragma unroll
        for (int u = 0; u < 4; u++) k[u] = p0 + u * G < e ? ac[p0 + u * G] : -1;
        int4 bm[4];
#pragma unroll
        for (int u = 0; u < 4; u++) bm[u] = k[u] >= 0 ? bmeta[k[u]] : make_int4(0, 0, INT32_MAX, -1);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int len = bm[u].y - bm[u].x;
          f += len;
          if (len > 0) {
            lo = bm[u].z < lo ? bm[u].z : lo;
            hi = bm[u].w > hi ? bm[u].w : hi;
          }
        }
      }
    }
    f = gsum<G, int64_t>(f, nullptr);
    lo = gmin<G>(lo);
    hi = gmax<G>(hi);
    const bool leader = live && !lng && lane == 0;
    const auto bs = finish(i, f, lo, hi, e - s, leader, tpl);
    // template rows: metric = products (their slot-table size), words = longest B row
    stats_add_wave(st, leader ? bs.first : -1, f, tpl >= 0 ? T.info[tpl].w : bs.second);
    // this wave's long rows, one at a time with all 64 lanes
    uint64_t longs = __ballot(lng && lane == 0);
    while (longs) {
      const int src = __builtin_ctzll(longs);
      longs &= longs - 1ull;
      const int64_t li = __shfl((long long)i, src, kWave);
      const int ls = __shfl(s, src, kWave), le = __shfl(e, src, kWave);
      int64_t lf = 0;
      int llo = INT32_MAX, lhi = -1;
      for (int p = ls + wl; p < le; p += kWave) {
        const int4 bm = bmeta[ac[p]];
        const int len = bm.y - bm.x;
        lf += len;
        if (len > 0) {
          llo = bm.z < llo ? bm.z : llo;
          lhi = bm.w > lhi ? bm.w : lhi;
        }
      }
      lf = gsum<kWave, int64_t>(lf, nullptr);
      llo = gmin<kWave>(llo);
      lhi = gmax<kWave>(lhi);
      const auto lb = finish(li, lf, llo, lhi, le - ls, wl == 0, -1);
      stats_add_wave(st, wl == 0 ? lb.first : -1, lf, lb.second);
    }
  }
  // block reduction of the product count
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) myprod += __shfl_xor(myprod, m, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = myprod;
  __syncthreads();
  unsigned long long t = 0;
  for (int w = 0; w < (int)blockDim.x / 64; w++) t += red[w];
  stats_flush(st, part, prodpart, t);
}
void launch_prepass(int32_t n, const int32_t* ap, const int32_t* ac, const int4* bmeta, int upper,
                    const BinParams& P, int64_t* rowf, int32_t* rowlo, int32_t* rowhi, uint8_t* bin,
                    int32_t* stats, unsigned long long* total_products, int32_t* part, const TplArgs* T,
                    int32_t* rowcnt, int32_t* rowwords, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * kPrepassLanes + 255) / 256;
  if (blocks > kStatsMaxBlocks) blocks = kStatsMaxBlocks;
  unsigned long long* prodpart = reinterpret_cast<unsigned long long*>(part + stats_part_elems());
  const TplArgs none{};
  hipLaunchKernelGGL(k_prepass, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, ac, bmeta, upper, P, rowf, rowlo,
                     rowhi, bin, part, prodpart, T ? *T : none, T ? 1 : 0, rowcnt, rowwords);
  stats_reduce(part, prodpart, (int)blocks, stats, total_products, s);
}

// Numeric bin of each row from its exact output count.
__global__ void k_numbin(int32_t n, const int32_t* __restrict__ ap, const int64_t* __restrict__ rowf,
                         const int32_t* __restrict__ rowcnt, const int32_t* __restrict__ rowwords,
                         const int32_t* __restrict__ rowlo, const int32_t* __restrict__ rowhi, BinParams P,
                         uint8_t* __restrict__ bin, int32_t* __restrict__ part, const int32_t* __restrict__ tplid,
                         const int4* __restrict__ tinfo) {
  __shared__ BinStatLds st;
  stats_init(st);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t wave_first = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x / kWave) * kWave;
  for (int64_t wi = wave_first; wi < n; wi += stride) {
    const int64_t i = wi + threadIdx.x % kWave;
    int b = -1;
    int64_t c = 0, stat = 0;
    if (i < n) {
      c = rowcnt[i];
      const int64_t words = span_words(rowlo[i], rowhi[i]);
      const int64_t nA = ap[i + 1] - ap[i];
      const int rw = rowwords[i];
      const int64_t listed = (rw & kNoWordList) == 0 && (rw & kWordsMask) <= kWordListCap ? (rw & kWordsMask) : -1;
      const int t = tplid ? tplid[i] : -1;
      if (t >= 0) {
        b = kBinTemplate;
        stat = tinfo[t].w;  // widest A or B row of the template: lanes per row
      } else {
        b = num_bin_of(P, c, words, rw & kWordsMask, rowf[i], nA, listed, &stat);
      }
      bin[i] = (uint8_t)b;
    }
    stats_add_wave(st, b, c, stat);
  }
  stats_flush(st, part, nullptr, 0ull);
}
void launch_numbin(int32_t n, const int32_t* ap, const int64_t* rowf, const int32_t* rowcnt,
                   const int32_t* rowwords, const int32_t* rowlo, const int32_t* rowhi, const BinParams& P,
                   uint8_t* bin, int32_t* stats, int32_t* part, const TplArgs* T, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n + 255) / 256;
  if (blocks > kStatsMaxBlocks) blocks = kStatsMaxBlocks;
  hipLaunchKernelGGL(k_numbin, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, rowf, rowcnt, rowwords, rowlo,
                     rowhi, P, bin, part, T ? T->tplid : nullptr, T ? T->info : nullptr);
  stats_reduce(part, nullptr, (int)blocks, stats, nullptr, s);
}

// Partition rows by bin, keeping ascending row order inside each block tile
// (neighbouring rows share B rows, so order matters for cache reuse, never
// for results).  A tile is 1024 rows: wave w owns rows [w*256, w*256+256) of
// it in four chunks of 64; ranks come from wave ballots per distinct bin; one
// global atomic per (tile, bin) reserves the tile's range.
__global__ void __launch_bounds__(256) k_bin_scatter(int32_t n, const uint8_t* __restrict__ bin,
                                                     const int32_t* __restrict__ stats, int32_t* __restrict__ taken,
                                                     int32_t* __restrict__ rows, const int64_t* __restrict__ total_src,
                                                     int64_t* __restrict__ total_dst) {
  constexpr int kItems = 4, kWaves = 4;
  __shared__ int32_t cnt[kWaves][kNumBins];
  __shared__ int32_t woff[kWaves][kNumBins];
  __shared__ int32_t start[kNumBins];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  // first slot of each bin in the row list: exclusive scan of the bin
  // counts over bins 1.. (bin 0 is not listed), in every block
  for (int b = threadIdx.x; b < kNumBins; b += blockDim.x) start[b] = b > 0 ? stats[b * kStatFields] : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    if (blockIdx.x == 0 && total_src) *total_dst = *total_src;
    int32_t run = 0;
    for (int b = 0; b < kNumBins; b++) {
      const int32_t c = start[b];
      start[b] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int64_t tile = (int64_t)blockIdx.x * 1024; tile < n; tile += (int64_t)gridDim.x * 1024) {
    for (int t = lane; t < kNumBins; t += kWave) cnt[w][t] = 0;
    __syncthreads();
    int myb[kItems], myr[kItems];
#pragma unroll
    for (int t = 0; t < kItems; t++) {
      const int64_t i = tile + (int64_t)w * 256 + t * kWave + lane;
      const int b = i < n ? (int)bin[i] : 0;
      myb[t] = b;
      myr[t] = 0;
      uint64_t todo = __ballot(b > 0);
      while (todo) {
        const int leader = __builtin_ctzll(todo);
        const int lb = __shfl(b, leader, kWave);
        const uint64_t same = __ballot(b == lb) & todo;
        if (b == lb && b > 0) myr[t] = cnt[w][lb] + __popcll(same & ((1ull << lane) - 1ull));
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane == leader) cnt[w][lb] += __popcll(same);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        todo &= ~same;
      }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kNumBins; b += blockDim.x) {
      int tot = 0;
      for (int v = 0; v < kWaves; v++) tot += cnt[v][b];
      int base = (b > 0 && tot) ? start[b] + atomicAdd(&taken[b], tot) : 0;
      for (int v = 0; v < kWaves; v++) { woff[v][b] = base; base += cnt[v][b]; }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kItems; t++) {
      const int64_t i = tile + (int64_t)w * 256 + t * kWave + lane;
      if (myb[t] > 0) rows[woff[w][myb[t]] + myr[t]] = (int32_t)i;
    }
    __syncthreads();
  }
}
void launch_bin_scatter(int32_t n, const uint8_t* bin, const int32_t* stats, int32_t* taken, int32_t* rows,
                        const int64_t* total_src, int64_t* total_dst, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n + 1023) / 1024;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)blocks), dim3(256), 0, s, n, bin, stats, taken, rows, total_src,
                     total_dst);
}

// ---------------------------------------------------------------------------
// product enumeration for one row, shared by symbolic and numeric kernels.
//
// G <= 64: A entries are fetched G at a time (one per lane) and broadcast with
// lane shuffles; the group walks them in order, lanes split each B row.
// G  > 64: the whole workgroup owns the row; A-entry metadata is staged in
// LDS (stage arrays of G entries).
// fn(j, q, a) is called once per product with j = B column, q = B position,
// a = A value.
// ---------------------------------------------------------------------------
struct Stage {
  int32_t bs[256];
  int32_t be[256];
  double a[256];
  int32_t longs[256];
  int32_t nlong;
};

// fn(j, prod) is called once per product with j = B column and prod = a*b
// (rounded product; 0 when VALS is false).
//
// G <= 64: the row's A entries are fetched G at a time (one per lane) and
// broadcast with lane shuffles.  U consecutive entries are handled per step:
// their first G B-row entries are loaded together (U independent gathers in
// flight per lane), then applied strictly in entry order, each entry's
// remainder (B rows longer than G) before the next entry, so every output
// column still receives its contributions in A-row storage order.
// G > 64: the whole workgroup owns the row; A-entry metadata is staged in LDS.
template <int G, bool VALS, int U = 8, typename Fn>
__device__ __forceinline__ void for_each_product(int i, int lane, const RowArgs& A, Stage* stage, Fn&& fn) {
  const int s = A.ap[i], e = A.ap[i + 1];
  for (int base = s; base < e; base += G) {
    const int p = base + lane;
    int bs = 0, be = 0;
    double a = 0.0;
    if (p < e) {
      const int k = A.ac[p];
      const int4 bm = A.bmeta[k];
      bs = bm.x;
      be = bm.y;
      if (VALS) a = A.av[p];
    }
    const int m = (e - base) < G ? (e - base) : G;
    if constexpr (G <= kWave) {
      for (int t0 = 0; t0 < m; t0 += U) {
        int jj[U], qb[U], qe[U];
        double pv[U], aa[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int t = t0 + u;
          const int tt = t < m ? t : 0;
          qb[u] = __shfl(bs, tt, G) + lane;
          qe[u] = t < m ? __shfl(be, tt, G) : 0;
          aa[u] = VALS ? __shfl(a, tt, G) : 0.0;
          const bool ok = qb[u] < qe[u];
          jj[u] = ok ? A.bc[qb[u]] : -1;
          pv[u] = (VALS && ok) ? aa[u] * A.bv[qb[u]] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (jj[u] >= 0) fn(jj[u], pv[u]);
          for (int q = qb[u] + G; q < qe[u]; q += G) fn(A.bc[q], VALS ? aa[u] * A.bv[q] : 0.0);
        }
      }
    } else {
      // whole workgroup per row; the order of contributions is not fixed
      // here (atomics), so products are balanced over the lanes: B rows of
      // at most kShortRow entries are walked one lane per A entry, longer
      // ones with all lanes over the B row
      constexpr int kShortRow = 32;
      __syncthreads();
      stage->bs[lane] = bs;
      stage->be[lane] = be;
      if (VALS) stage->a[lane] = a;
      if (lane == 0) stage->nlong = 0;
      __syncthreads();
      if (be - bs > kShortRow) stage->longs[atomicAdd(&stage->nlong, 1)] = lane;
      else
        for (int q = bs; q < be; q++) fn(A.bc[q], VALS ? a * A.bv[q] : 0.0);
      __syncthreads();
      const int nl = stage->nlong;
      for (int t = 0; t < nl; t++) {
        const int src = stage->longs[t];
        const int tbs = stage->bs[src], tbe = stage->be[src];
        const double ta = VALS ? stage->a[src] : 0.0;
        for (int q = tbs + lane; q < tbe; q += G) fn(A.bc[q], VALS ? ta * A.bv[q] : 0.0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Accumulator modes of a row (per bin):
//   dense   a bitmap over the row's whole output span [lo, hi] (1 bit per
//           column); output positions are prefix popcounts of the words.
//   whash   "word hash": an LDS open-addressing table keyed by the 32-column
//           word index j>>5 with a 32-bit column mask per key.  Rows whose
//           span is wide but whose columns cluster (stencils) need a handful
//           of keys; scattered rows degrade to one key per column.  The
//           numeric pass sorts only the keys, then uses prefix popcounts.
// ---------------------------------------------------------------------------

// Find or claim the slot of word key wk.  Returns the slot; *claimed is set
// when this call inserted the key.
__device__ __forceinline__ uint32_t whash_insert(int32_t* keys, int logT, uint32_t mask, int32_t wk, bool* claimed) {
  uint32_t h = hash_slot(wk, logT);
  while (true) {
    const int cur = __atomic_load_n(&keys[h], __ATOMIC_RELAXED);
    if (cur == wk) { *claimed = false; return h; }
    if (cur == -1) {
      const int old = atomicCAS(&keys[h], -1, wk);
      if (old == -1) { *claimed = true; return h; }
      if (old == wk) { *claimed = false; return h; }
    }
    h = (h + 1u) & mask;
  }
}
// Slot of a key known to be present.
__device__ __forceinline__ uint32_t whash_find(const int32_t* keys, int logT, uint32_t mask, int32_t wk) {
  uint32_t h = hash_slot(wk, logT);
  while (keys[h] != wk) h = (h + 1u) & mask;
  return h;
}

// ---------------------------------------------------------------------------
// symbolic: distinct output columns (rowcnt) and distinct 32-column words
// (rowwords) per row
// ---------------------------------------------------------------------------
template <int G, bool DENSE>
__global__ void __launch_bounds__(256) k_symbolic_lds(RowArgs A, int32_t row_words, int32_t* __restrict__ rowcnt) {
  extern __shared__ uint32_t lds[];
  __shared__ Stage stage_buf[G > kWave ? 1 : 0 + 1];
  __shared__ int32_t red[16];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = xcd_block(blockIdx.x, gridDim.x) * R + gi;
  if (G <= kWave && r >= A.nrows) return;
  const int i = A.rows[r];
  uint32_t* tab = lds + (size_t)gi * row_words;
  int cnt = 0, nw = 0;
  const int lo = A.rowlo[i], hi = A.rowhi[i];
  if constexpr (DENSE) {
    const int words = (int)span_words(lo, hi);
    for (int t = lane; t < words; t += G) tab[t] = 0u;
    gsync<G>();
    for_each_product<G, false>(i, lane, A, stage_buf, [&](int j, double) {
      const uint32_t b = (uint32_t)(j - lo);
      const uint32_t bit = 1u << (b & 31u);
      const uint32_t old = atomicOr(&tab[b >> 5], bit);
      cnt += (old & bit) ? 0 : 1;
      nw += old == 0u ? 1 : 0;
    });
    if (G > kWave && A.wgbits) {
      // keep the row's bitmap for the numeric workgroup pass (streaming
      // stores: they would otherwise evict the B rows this pass gathers)
      gsync<G>();
      uint32_t* dst = A.wgbits + (size_t)r * A.wgstride;
      for (int t = lane; t < words; t += G) __builtin_nontemporal_store(tab[t], &dst[t]);
      if (lane == 0) A.wgslot[i] = r;
    }
  } else {
    const int64_t keys_bound = A.rowf[i] < span_words(lo, hi) ? A.rowf[i] : span_words(lo, hi);
    const int64_t T = np2_i64((keys_bound * 4) / 3 + 1);
    const int logT = __builtin_ctzll((unsigned long long)T);
    const uint32_t mask = (uint32_t)T - 1u;
    int32_t* keys = (int32_t*)tab;
    uint32_t* masks = tab + T;
    for (int t = lane; t < (int)T; t += G) { keys[t] = -1; masks[t] = 0u; }
    gsync<G>();
    for_each_product<G, false>(i, lane, A, stage_buf, [&](int j, double) {
      bool claimed;
      const uint32_t h = whash_insert(keys, logT, mask, j >> 5, &claimed);
      nw += claimed ? 1 : 0;
      const uint32_t bit = 1u << (j & 31);
      const uint32_t old = atomicOr(&masks[h], bit);
      cnt += (old & bit) ? 0 : 1;
    });
  }
  cnt = gsum<G, int32_t>(cnt, red);
  nw = gsum<G, int32_t>(nw, red);
  if (lane == 0) { rowcnt[i] = cnt; A.rowwords[i] = nw | kNoWordList; }
}

// ---------------------------------------------------------------------------
// numeric
// ---------------------------------------------------------------------------

// In-LDS bitonic sort by key (all comparators ascending; positions >= len act
// as +infinity and are never touched).
template <int G, typename V>
__device__ __forceinline__ void lds_sort_pairs(int32_t* keys, V* vals, int len, int lane) {
  int N = 1;
  while (N < len) N <<= 1;
  for (int k = 2; k <= N; k <<= 1) {
    const int half = k >> 1;
    for (int t = lane; t < (N >> 1); t += G) {
      const int blk = t / half, off = t % half;
      const int i1 = blk * k + off, i2 = blk * k + k - 1 - off;
      if (i2 < len) {
        const int32_t k1 = keys[i1], k2 = keys[i2];
        if (k1 > k2) {
          keys[i1] = k2; keys[i2] = k1;
          const V v1 = vals[i1]; vals[i1] = vals[i2]; vals[i2] = v1;
        }
      }
    }
    gsync<G>();
    for (int jj = k >> 2; jj > 0; jj >>= 1) {
      for (int t = lane; t < (N >> 1); t += G) {
        const int i1 = ((t & ~(jj - 1)) << 1) | (t & (jj - 1));
        const int i2 = i1 + jj;
        if (i2 < len) {
          const int32_t k1 = keys[i1], k2 = keys[i2];
          if (k1 > k2) {
            keys[i1] = k2; keys[i2] = k1;
            const V v1 = vals[i1]; vals[i1] = vals[i2]; vals[i2] = v1;
          }
        }
      }
      gsync<G>();
    }
  }
}

// ---------------------------------------------------------------------------
// Sorted transpose: the atomic-cursor scatter leaves each B row in arbitrary
// order; every row is then sorted by its column (the source row index), the
// values moving with their keys.  Rows of at most 64 entries: a group of 16,
// 32 or 64 lanes, in registers; up to kSortLds entries: one block in LDS; longer rows: one
// block, in place in global memory.
// ---------------------------------------------------------------------------
constexpr int kSortLds = 4096;

// Rows of (W/2, W] entries (2..16 for W = 16), one group of W lanes each.
template <int W>
__global__ void k_sort_rows_wave(int32_t n, const int32_t* __restrict__ bp, int32_t* __restrict__ bc,
                                 double* __restrict__ bv, const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  constexpr int kMin = W == 16 ? 2 : W / 2 + 1;
  const int lane = threadIdx.x % W;
  const int64_t ng = (int64_t)gridDim.x * blockDim.x / W;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / W; r < n; r += ng) {
    const int s = bp[r], len = bp[r + 1] - s;
    if (len < kMin || len > W) continue;  // uniform across the group
    int key = lane < len ? bc[s + lane] : INT32_MAX;
    double v = lane < len ? bv[s + lane] : 0.0;
#pragma unroll
    for (int k = 2; k <= W; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int pk = __shfl_xor(key, j, W);
        const double pv = __shfl_xor(v, j, W);
        const bool up = (lane & k) == 0, lower = (lane & j) == 0;
        // lower lane of an ascending pair keeps the minimum, and so on
        if (lower == up ? pk < key : pk > key) {
          key = pk;
          v = pv;
        }
      }
    }
    if (lane < len) {
      bc[s + lane] = key;
      bv[s + lane] = v;
    }
  }
}

// Rows longer than a wave, listed for balanced sorting: lengths in
// (kWave, kSortLds] from the front of `list`, longer ones from the back;
// count[0] / count[1] (zero beforehand) receive the two list lengths.
__global__ void k_select_rows(int32_t n, const int32_t* __restrict__ bp, int32_t* __restrict__ list,
                              int32_t* __restrict__ count, const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int len = bp[r + 1] - bp[r];
    if (len > kSortLds) list[n - 1 - atomicAdd(&count[1], 1)] = (int32_t)r;
    else if (len > kWave) list[atomicAdd(&count[0], 1)] = (int32_t)r;
  }
}

__global__ void __launch_bounds__(256) k_sort_rows_lds(int32_t n, const int32_t* __restrict__ bp,
                                                       int32_t* __restrict__ bc, double* __restrict__ bv,
                                                       const int32_t* __restrict__ list,
                                                       const int32_t* __restrict__ count,
                                                       const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  constexpr int T = 256;
  __shared__ int32_t keys[kSortLds];
  __shared__ double vals[kSortLds];
  const int m = count[0];
  for (int q = blockIdx.x; q < m; q += gridDim.x) {
    const int r = list[q];
    const int s = bp[r], len = bp[r + 1] - s;
    for (int t = threadIdx.x; t < len; t += T) {
      keys[t] = bc[s + t];
      vals[t] = bv[s + t];
    }
    __syncthreads();
    lds_sort_pairs<T, double>(keys, vals, len, threadIdx.x);
    for (int t = threadIdx.x; t < len; t += T) {
      bc[s + t] = keys[t];
      bv[s + t] = vals[t];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(1024) k_sort_rows_long(int32_t n, const int32_t* __restrict__ bp,
                                                         int32_t* __restrict__ bc, double* __restrict__ bv,
                                                         const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ count,
                                                         const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int m = count[1];
  for (int q = blockIdx.x; q < m; q += gridDim.x) {
    const int r = list[n - 1 - q];
    const int s = bp[r], len = bp[r + 1] - s;
    lds_sort_pairs<1024, double>(bc + s, bv + s, len, threadIdx.x);
    __syncthreads();
  }
}

int64_t sort_rows_scratch_elems(int32_t n) { return (int64_t)n + 2; }

void launch_sort_rows(int32_t n, const int32_t* bp, int32_t* bc, double* bv, int32_t* scratch, const int32_t* gate,
                      hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_sort_rows_wave<16>, dim3((unsigned)blocks), dim3(256), 0, s, n, bp, bc, bv, gate);
  hipLaunchKernelGGL(k_sort_rows_wave<32>, dim3((unsigned)blocks), dim3(256), 0, s, n, bp, bc, bv, gate);
  hipLaunchKernelGGL(k_sort_rows_wave<64>, dim3((unsigned)blocks), dim3(256), 0, s, n, bp, bc, bv, gate);
  int32_t* count = scratch + n;
  launch_fill_i32(count, 2, 0, s);
  int64_t b1 = ((int64_t)n + 255) / 256;
  if (b1 > 4096) b1 = 4096;
  hipLaunchKernelGGL(k_select_rows, dim3((unsigned)b1), dim3(256), 0, s, n, bp, scratch, count, gate);
  hipLaunchKernelGGL(k_sort_rows_lds, dim3(2048), dim3(256), 0, s, n, bp, bc, bv, scratch, count, gate);
  hipLaunchKernelGGL(k_sort_rows_long, dim3(256), dim3(1024), 0, s, n, bp, bc, bv, scratch, count, gate);
}

// Group-wide exclusive scan (G <= 64 shuffles, else block scan).
template <int G>
__device__ __forceinline__ int group_scan_excl(int v, int lane, int64_t* red64) {
  if constexpr (G <= kWave) {
    int tot;
    return gscan_excl<G, int>(v, lane, &tot);
  } else {
    int64_t tot;
    return (int)block_scan_excl<int64_t>(v, red64, &tot);
  }
}

// Dense numeric.  LDS per row: info[words] (uint2: bits, prefix),
// vals[c_max] (f64), cols[c_max] (i32).  Output positions come from prefix
// popcounts, so the row is produced already sorted.
template <int G>
__device__ __forceinline__ void numeric_dense_row(int i, int lane, const RowArgs& A, uint2* info, double* vals,
                                                  int32_t* cols, const int64_t* rowoff, int32_t* ccol,
                                                  double* cval, Stage* stage, int64_t* red64) {
  const int lo = A.rowlo[i], hi = A.rowhi[i];
  const int words = (int)span_words(lo, hi);
  const int64_t off = rowoff[i];
  const int c = (int)(rowoff[i + 1] - off);
  // the row's structure: the bitmap its symbolic pass kept, else a pass
  // over the products
  const int32_t slot = A.wgslot ? A.wgslot[i] : -1;
  if (slot >= 0) {
    const uint32_t* src = A.wgbits + (size_t)slot * A.wgstride;
    for (int t = lane; t < words; t += G) info[t] = make_uint2(__builtin_nontemporal_load(&src[t]), 0u);
    for (int t = lane; t < c; t += G) vals[t] = 0.0;
  } else {
    for (int t = lane; t < words; t += G) info[t] = make_uint2(0u, 0u);
    for (int t = lane; t < c; t += G) vals[t] = 0.0;
    gsync<G>();
    uint32_t* bits = (uint32_t*)info;
    for_each_product<G, false>(i, lane, A, stage, [&](int j, double) {
      const uint32_t b = (uint32_t)(j - lo);
      atomicOr(&bits[(b >> 5) * 2], 1u << (b & 31u));
    });
  }
  gsync<G>();
  const int chunk = (words + G - 1) / G;
  const int w0 = lane * chunk;
  const int w1 = (w0 + chunk) < words ? (w0 + chunk) : words;
  int local = 0;
  for (int w = w0; w < w1; w++) local += __popc(info[w].x);
  int run = group_scan_excl<G>(local, lane, red64);
  for (int w = w0; w < w1; w++) {
    uint32_t x = info[w].x;
    info[w].y = (uint32_t)run;
    while (x) {
      const int bpos = __builtin_ctz(x);
      cols[run++] = lo + w * 32 + bpos;
      x &= x - 1u;
    }
  }
  gsync<G>();
  for_each_product<G, true>(i, lane, A, stage, [&](int j, double prod) {
    if (A.upper && j < i) return;
    const uint32_t b = (uint32_t)(j - lo);
    const uint2 wi = info[b >> 5];
    const uint32_t pos = wi.y + __popc(wi.x & ((1u << (b & 31u)) - 1u));
    atomicAdd(&vals[pos], prod);
  });
  gsync<G>();
  for (int t = lane; t < c; t += G) {
    ccol[off + t] = cols[t];
    cval[off + t] = vals[t];
  }
}

// Word-hash numeric.  LDS per row (Tm = table slots of the bin maximum):
//   hkey[Tm] i32 | hinfo[Tm] uint2 (mask, prefix) | vals[max(c_max, Tm)] f64
//   (first holding the key sort buffer) | cols[c_max] i32
template <int G>
__device__ __forceinline__ void numeric_whash_row(int i, int lane, const RowArgs& A, int32_t* hkey, uint2* hinfo,
                                                  double* vals, int32_t* cols, const int64_t* rowoff,
                                                  int32_t* ccol, double* cval, Stage* stage, int64_t* red64) {
  const int64_t off = rowoff[i];
  const int c = (int)(rowoff[i + 1] - off);
  const int w = A.rowwords[i] & kWordsMask;
  const int64_t T = np2_i64((int64_t)(w * 3) / 2 + 1);
  const int logT = __builtin_ctzll((unsigned long long)T);
  const uint32_t mask = (uint32_t)T - 1u;
  for (int t = lane; t < (int)T; t += G) { hkey[t] = -1; hinfo[t] = make_uint2(0u, 0u); }
  gsync<G>();
  uint32_t* hbits = (uint32_t*)hinfo;
  for_each_product<G, false>(i, lane, A, stage, [&](int j, double) {
    bool claimed;
    const uint32_t h = whash_insert(hkey, logT, mask, j >> 5, &claimed);
    atomicOr(&hbits[h * 2], 1u << (j & 31));
  });
  gsync<G>();
  // gather the occupied slots: skey = word key, sslot = slot
  int32_t* skey = (int32_t*)vals;
  int32_t* sslot = skey + T;
  int count = 0;
  for (int base = 0; base < (int)T; base += G) {
    const int t = base + lane;
    const int32_t key = t < (int)T ? hkey[t] : -1;
    const bool occ = key != -1;
    int rank, tot;
    if constexpr (G <= kWave) {
      const uint64_t gb = group_bits<G>(__ballot(occ));
      rank = __popcll(gb & ((1ull << lane) - 1ull));
      tot = __popcll(gb);
    } else {
      int64_t t64;
      rank = (int)block_scan_excl<int64_t>(occ ? 1 : 0, red64, &t64);
      tot = (int)t64;
    }
    if (occ) { skey[count + rank] = key; sslot[count + rank] = t; }
    count += tot;
  }
  gsync<G>();
  lds_sort_pairs<G, int32_t>(skey, sslot, w, lane);
  // prefix popcounts over the sorted keys; emit the sorted column list
  const int chunk = (w + G - 1) / G;
  const int k0 = lane * chunk;
  const int k1 = (k0 + chunk) < w ? (k0 + chunk) : w;
  int local = 0;
  for (int k = k0; k < k1; k++) local += __popc(hinfo[sslot[k]].x);
  int run = group_scan_excl<G>(local, lane, red64);
  for (int k = k0; k < k1; k++) {
    const int sl = sslot[k];
    const int32_t wk = skey[k];
    uint32_t x = hinfo[sl].x;
    hinfo[sl].y = (uint32_t)run;
    while (x) {
      const int bpos = __builtin_ctz(x);
      cols[run++] = wk * 32 + bpos;
      x &= x - 1u;
    }
  }
  gsync<G>();
  for (int t = lane; t < c; t += G) vals[t] = 0.0;
  gsync<G>();
  for_each_product<G, true>(i, lane, A, stage, [&](int j, double prod) {
    if (A.upper && j < i) return;
    const uint32_t h = whash_find(hkey, logT, mask, j >> 5);
    const uint2 wi = hinfo[h];
    const uint32_t pos = wi.y + __popc(wi.x & ((1u << (j & 31)) - 1u));
    atomicAdd(&vals[pos], prod);
  });
  gsync<G>();
  for (int t = lane; t < c; t += G) {
    ccol[off + t] = cols[t];
    cval[off + t] = vals[t];
  }
}

__host__ __device__ __forceinline__ int64_t whash_slots_num(int64_t w) {
  int64_t T = 2;
  while (T < (w * 3) / 2 + 1) T <<= 1;
  return T;
}
__host__ __device__ __forceinline__ int64_t dense_row_bytes(int64_t words_max, int64_t c_max) {
  return ((words_max * 8 + c_max * 12) + 15) & ~(int64_t)15;
}
__host__ __device__ __forceinline__ int64_t whash_row_bytes(int64_t w_max, int64_t c_max) {
  const int64_t T = whash_slots_num(w_max);
  const int64_t v = c_max > T ? c_max : T;
  return ((T * 12 + v * 8 + c_max * 4) + 15) & ~(int64_t)15;
}

template <int G, bool DENSE>
__global__ void __launch_bounds__(256) k_numeric_lds(RowArgs A, int32_t words_max, int32_t c_max,
                                                     const int64_t* __restrict__ rowoff, int32_t* __restrict__ ccol,
                                                     double* __restrict__ cval) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ Stage stage_buf[G > kWave ? 1 : 0 + 1];
  __shared__ int64_t red64[17];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = xcd_block(blockIdx.x, gridDim.x) * R + gi;
  if (G <= kWave && r >= A.nrows) return;
  const int i = A.rows[r];
  if constexpr (DENSE) {
    unsigned char* base = lds_raw + (size_t)gi * dense_row_bytes(words_max, c_max);
    uint2* info = (uint2*)base;
    double* vals = (double*)(base + (size_t)words_max * 8);
    int32_t* cols = (int32_t*)(base + (size_t)words_max * 8 + (size_t)c_max * 8);
    numeric_dense_row<G>(i, lane, A, info, vals, cols, rowoff, ccol, cval, stage_buf, red64);
  } else {
    const int64_t Tm = whash_slots_num(words_max);
    const int64_t vm = c_max > Tm ? c_max : Tm;
    unsigned char* base = lds_raw + (size_t)gi * whash_row_bytes(words_max, c_max);
    uint2* hinfo = (uint2*)base;
    double* vals = (double*)(base + Tm * 8);
    int32_t* hkey = (int32_t*)(base + Tm * 8 + vm * 8);
    int32_t* cols = hkey + Tm;
    numeric_whash_row<G>(i, lane, A, hkey, hinfo, vals, cols, rowoff, ccol, cval, stage_buf, red64);
  }
}

// ---------------------------------------------------------------------------
// Register-sort bins: rows of at most N = G*E products whose output columns
// are too scattered for a window (random sparsity).  One group of G lanes
// per row:
//   1. the row's A entries go to LDS with the exclusive scan of their B-row
//      lengths (epos), so product t (in A-row storage order, then B-row
//      order) belongs to the last entry r with epos[r] <= t;
//   2. lane l forms products t = e*G + l (coalesced B gathers), keyed
//      (column << S | t) with S = max(9, log2 N) -- or the column alone in
//      the symbolic pass -- and the numeric pass parks each product value in
//      LDS at index t;
//   3. a bitonic network sorts the G*E keys held E per lane in registers
//      (lane-local compare-exchanges, then xor shuffles across lanes);
//   4. distinct columns are the run heads of the sorted keys; a numeric head
//      sums its run in t order starting from 0.0, which is exactly the
//      A-row-order accumulation of the other bins, and writes the output
//      entry at its rank among the heads.
// ---------------------------------------------------------------------------
template <typename K>
__device__ __forceinline__ K shfl_xor_key(K v, int m, int width) {
  if constexpr (sizeof(K) == 8) {
    const uint32_t lo = __shfl_xor((uint32_t)v, m, width), hi = __shfl_xor((uint32_t)(v >> 32), m, width);
    return ((K)hi << 32) | lo;
  } else {
    return (K)__shfl_xor((uint32_t)v, m, width);
  }
}

// Ascending bitonic sort of the G*E keys of a group; lane l holds positions
// l*E .. l*E + E - 1 on return.
template <int G, int E, typename K>
__device__ __forceinline__ void reg_bitonic(K (&k)[E], int lane) {
#pragma unroll
  for (int size = 2; size <= G * E; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      if (j >= E) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int p = lane * E + e;
          const K o = shfl_xor_key<K>(k[e], j / E, G);
          // the lower position of an ascending pair keeps the minimum
          const bool keep_min = ((p & size) == 0) == ((p & j) == 0);
          k[e] = keep_min ? (o < k[e] ? o : k[e]) : (o > k[e] ? o : k[e]);
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; e++) {
          if ((e & j) == 0) {
            const int p = lane * E + e;
            const K a = k[e], b = k[e | j];
            const bool asc = (p & size) == 0;
            const K mn = a < b ? a : b, mx = a < b ? b : a;
            k[e] = asc ? mn : mx;
            k[e | j] = asc ? mx : mn;
          }
        }
      }
    }
  }
}

// LDS of one row of a sort bin with N slots, A rows of at most `na`
// entries and keys of `kb` bytes: the entry area epos (na + 1) i32 | eb na
// i32, which the sorted keys (N x kb) overwrite, then (numeric) the product
// values val N f64.
__host__ __device__ __forceinline__ int64_t sort_entry_bytes(int64_t na) {
  return align16((na + 1) * 4) + align16(na * 4);
}
__host__ __device__ __forceinline__ int64_t sort_row_bytes(int N, int64_t na, int kb, bool vals) {
  const int64_t e = sort_entry_bytes(na), k = align16((int64_t)N * kb);
  return (e > k ? e : k) + (vals ? (int64_t)N * 8 : 0);
}

// Step 1: A entries of row i into LDS; returns the row's product count.
template <int G>
__device__ __forceinline__ int sort_stage_entries(int i, int lane, const RowArgs& A, int32_t* epos, int32_t* eb) {
  const int s = A.ap[i], nA = A.ap[i + 1] - s;
  int run = 0;
  for (int base = 0; base < nA; base += G) {
    const int r = base + lane;
    int len = 0, bs = 0;
    if (r < nA) {
      const int4 bm = A.bmeta[A.ac[s + r]];
      bs = bm.x;
      len = bm.y - bm.x;
    }
    int tot;
    const int ex = gscan_excl<G, int>(len, lane, &tot);
    if (r < nA) {
      epos[r] = run + ex;
      eb[r] = bs - (run + ex);
    }
    run += tot;
  }
  if (lane == 0) epos[nA] = run;
  gsync<G>();
  return run;
}

// Entry of product t: the last r in [0, nA) with epos[r] <= t.
__device__ __forceinline__ int sort_entry_of(const int32_t* epos, int nA, int t) {
  int lo = 0, hi = nA;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (epos[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int G, int E>
__global__ void __launch_bounds__(256) k_sym_sort(RowArgs A, int32_t na_max, int32_t* __restrict__ rowcnt) {
  constexpr int N = G * E;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = xcd_block(blockIdx.x, gridDim.x) * R + gi;
  if (r >= A.nrows) return;
  const int i = A.rows[r];
  unsigned char* base = lds_raw + (size_t)gi * sort_row_bytes(N, na_max, 4, false);
  int32_t* epos = (int32_t*)base;
  int32_t* eb = (int32_t*)(base + align16((int64_t)(na_max + 1) * 4));
  const int nA = A.ap[i + 1] - A.ap[i];
  const int f = sort_stage_entries<G>(i, lane, A, epos, eb);
  uint32_t key[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int t = e * G + lane;
    key[e] = 0xFFFFFFFFu;
    if (t < f) key[e] = (uint32_t)A.bc[eb[sort_entry_of(epos, nA, t)] + t];
  }
  reg_bitonic<G, E, uint32_t>(key, lane);
  gsync<G>();  // every lane is done with epos / eb before the keys overwrite them
  uint32_t* skey = (uint32_t*)epos;
#pragma unroll
  for (int e = 0; e < E; e++) skey[lane * E + e] = key[e];
  gsync<G>();
  // run heads of columns and of 32-column words
  int cnt = 0, nw = 0;
  bool whead[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int p = lane * E + e;
    const uint32_t prev = p > 0 ? skey[p - 1] : 0xFFFFFFFFu;
    const bool real = key[e] != 0xFFFFFFFFu;
    cnt += (real && (p == 0 || key[e] != prev)) ? 1 : 0;
    whead[e] = real && (p == 0 || (key[e] >> 5) != (prev >> 5));
    nw += whead[e] ? 1 : 0;
  }
  cnt = gsum<G, int32_t>(cnt, nullptr);
  int words;
  int rank = gscan_excl<G, int>(nw, lane, &words);
  // the row's sorted (word, mask) list for the numeric list bins, as the
  // window kernels leave it: words relative to the 32-aligned span start
  if (words <= kWordListCap) {
    uint2* wl = A.wlist + (size_t)i * kWordListCap;
    const uint32_t w0 = (uint32_t)A.rowlo[i] >> 5;
#pragma unroll
    for (int e = 0; e < E; e++) {
      if (!whead[e]) continue;
      const uint32_t wk = key[e] >> 5;
      uint32_t mask = 0u;
      for (int p = lane * E + e; p < N && (skey[p] >> 5) == wk; p++) mask |= 1u << (skey[p] & 31u);
      wl[rank++] = make_uint2(wk - w0, mask);
    }
  }
  if (lane == 0) {
    rowcnt[i] = cnt;
    A.rowwords[i] = words <= kWordListCap ? words : (words | kNoWordList);
  }
}

// Bits of the product index in a numeric sort key.
__host__ __device__ __forceinline__ constexpr int sort_key_shift(int N) {
  return N <= 512 ? 9 : N == 1024 ? 10 : 11;
}

template <int G, int E, bool UPPER, typename K>
__global__ void __launch_bounds__(256) k_num_sort(RowArgs A, int32_t na_max, const int64_t* __restrict__ rowoff,
                                                  int32_t* __restrict__ ccol, double* __restrict__ cval) {
  constexpr int N = G * E;
  constexpr int SH = sort_key_shift(N);
  constexpr K kNone = ~(K)0, kIdx = ((K)1 << SH) - 1;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = xcd_block(blockIdx.x, gridDim.x) * R + gi;
  if (r >= A.nrows) return;
  const int i = A.rows[r];
  const int64_t ebytes = sort_entry_bytes(na_max), kbytes = align16((int64_t)N * sizeof(K));
  unsigned char* base = lds_raw + (size_t)gi * sort_row_bytes(N, na_max, (int)sizeof(K), true);
  int32_t* epos = (int32_t*)base;
  int32_t* eb = (int32_t*)(base + align16((int64_t)(na_max + 1) * 4));
  double* val = (double*)(base + (ebytes > kbytes ? ebytes : kbytes));
  K* skey = (K*)base;
  const int s = A.ap[i], nA = A.ap[i + 1] - s;
  const int f = sort_stage_entries<G>(i, lane, A, epos, eb);
  K key[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int t = e * G + lane;
    key[e] = kNone;
    if (t < f) {
      const int en = sort_entry_of(epos, nA, t);
      const int q = eb[en] + t;
      const int j = A.bc[q];
      // AAtS: values only for j >= i (the mirror fills the rest)
      val[t] = (!UPPER || j >= i) ? A.av[s + en] * A.bv[q] : 0.0;
      key[e] = ((K)j << SH) | (K)t;
    }
  }
  reg_bitonic<G, E, K>(key, lane);
  gsync<G>();  // every lane is done with epos / eb before the keys overwrite them
#pragma unroll
  for (int e = 0; e < E; e++) skey[lane * E + e] = key[e];
  gsync<G>();
  bool head[E];
  int cnt = 0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int p = lane * E + e;
    head[e] = key[e] != kNone && (p == 0 || (skey[p - 1] >> SH) != (key[e] >> SH));
    cnt += head[e] ? 1 : 0;
  }
  int tot;
  int64_t out = rowoff[i] + gscan_excl<G, int>(cnt, lane, &tot);
#pragma unroll
  for (int e = 0; e < E; e++) {
    if (!head[e]) continue;
    const int col = (int)(key[e] >> SH);
    double acc = 0.0;
    if (!UPPER || col >= i) {
      acc = acc + val[(int)(key[e] & kIdx)];
      for (int p = lane * E + e + 1; p < f; p++) {
        const K kq = skey[p];
        if ((kq >> SH) != key[e] >> SH) break;
        acc = acc + val[(int)(kq & kIdx)];
      }
    }
    ccol[out] = col;
    cval[out] = acc;
    out++;
  }
}

// 32-bit keys hold column << S | t while columns stay below 2^(32 - S) - 1
// (the all-ones key marks an empty slot)
__host__ __device__ __forceinline__ bool sort_keys64(int32_t ncols, int N) {
  return (int64_t)ncols >= ((int64_t)1 << (32 - sort_key_shift(N))) - 1;
}

template <int G>
static void sort_launch_shape(const RowArgs& a, int64_t row_bytes, int* R, unsigned* blocks) {
  int r = 256 / G;
  while (r > 1 && (int64_t)r * row_bytes > a.lds_budget) r >>= 1;
  *R = r;
  *blocks = (unsigned)((a.nrows + r - 1) / r);
}
template <int G, int E>
static void sym_sort_launch(const RowArgs& a, int32_t na_max, int32_t* rowcnt, hipStream_t s) {
  const int64_t rb = sort_row_bytes(G * E, na_max, 4, false);
  int R; unsigned blocks;
  sort_launch_shape<G>(a, rb, &R, &blocks);
  const size_t lds = (size_t)R * rb;
  set_dyn_lds((const void*)k_sym_sort<G, E>, lds);
  hipLaunchKernelGGL((k_sym_sort<G, E>), dim3(blocks), dim3(R * G), lds, s, a, na_max, rowcnt);
}
template <int G, int E, bool UPPER, typename K>
static void num_sort_launch_k(const RowArgs& a, int32_t na_max, const int64_t* rowoff, int32_t* ccol, double* cval,
                              hipStream_t s) {
  const int64_t rb = sort_row_bytes(G * E, na_max, (int)sizeof(K), true);
  int R; unsigned blocks;
  sort_launch_shape<G>(a, rb, &R, &blocks);
  const size_t lds = (size_t)R * rb;
  set_dyn_lds((const void*)k_num_sort<G, E, UPPER, K>, lds);
  hipLaunchKernelGGL((k_num_sort<G, E, UPPER, K>), dim3(blocks), dim3(R * G), lds, s, a, na_max, rowoff, ccol,
                     cval);
}
template <int G, int E>
static void num_sort_launch(const RowArgs& a, int32_t na_max, const int64_t* rowoff, int32_t* ccol, double* cval,
                            hipStream_t s) {
  const bool k64 = sort_keys64(a.ncols, G * E);
  if (a.upper) {
    if (k64) num_sort_launch_k<G, E, true, uint64_t>(a, na_max, rowoff, ccol, cval, s);
    else num_sort_launch_k<G, E, true, uint32_t>(a, na_max, rowoff, ccol, cval, s);
  } else {
    if (k64) num_sort_launch_k<G, E, false, uint64_t>(a, na_max, rowoff, ccol, cval, s);
    else num_sort_launch_k<G, E, false, uint32_t>(a, na_max, rowoff, ccol, cval, s);
  }
}

// ---------------------------------------------------------------------------
// Heavy rows: outputs too wide for one workgroup's LDS.  One workgroup per
// row, rows claimed from a counter (heavy rows differ widely in cost).  The
// row's output span [lo, hi] is processed in column windows of `wc` columns
// held in LDS: a bitmap (symbolic) or a bitmap plus an f64 accumulator
// (numeric).  When every B row is sorted, a cursor per A entry (the first B
// position not yet consumed) carries over from one window to the next, so
// each product is loaded once; otherwise each window rescans the B rows and
// filters by column.  Inside a window the products of kHeavyT A entries are
// flattened (exclusive scan of their segment lengths) and dealt to the
// threads, kHeavyU products per thread in flight, so short and long B rows
// balance.
// ---------------------------------------------------------------------------
struct HeavyStage {
  int64_t off[kHeavyT];  // exclusive scan of the entries' segment lengths
  int32_t qs[kHeavyT];   // segment start (B position)
  double a[kHeavyT];     // a_ik
  int64_t red[17];
  int64_t tot;           // products of the last staged chunk
  int32_t row;
};

// A heavy row's entries kept in LDS across its windows (rows of at most
// kHeavyT entries over sorted B rows): cursor, B-row end and B-row maximum
// column per entry, and a_ik in HeavyStage::a, loaded once per row instead of
// an A-column, B-metadata and cursor load chain per window and pass.
struct HeavyRes {
  int32_t* cur;          // nullptr: not resident (cursors in global memory)
  int32_t* be;
  int32_t* bw;
};

// B ranges of a chunk of m A entries (from position cb) inside the window
// ending at wend: entry k's range [qs, qe) goes to hs.qs[k] and hs.off[k]
// (parked there until the caller's scan), a_ik to hs.a[k] when VALS.  With
// carry (B rows sorted), qs is the entry's cursor after the first window and
// qe the first B position at or past wend; keep moves the cursor there.
template <bool VALS>
__device__ __forceinline__ void heavy_stage_ranges(const RowArgs& A, HeavyStage& hs, int cb, int m, int64_t wend,
                                                   bool carry, bool first, bool keep, int32_t* cur, HeavyRes res) {
  const int tid = threadIdx.x;
  // each entry's B range inside the window, T lanes per entry (T = 1024 /
  // m rounded down to a power of two, at most 64): the window end is found
  // by a T-ary search, log_(T+1) dependent loads instead of log_2.  The
  // group's lane 0 stages qs, qe (parked in hs.off) and a_ik.
  int lg = 0;
  while (lg < 6 && m * (2 << lg) <= kHeavyT) lg++;
  const int T = 1 << lg;
  const int ent = tid >> lg, sub = tid & (T - 1);
  if (ent < m) {
    const int p = cb + ent;
    int qs, h;
    int64_t bmax;
    if (res.cur) {
      // resident: the group's lanes read the cursor before its lane 0 moves it
      qs = res.cur[ent];
      h = res.be[ent];
      bmax = res.bw[ent];
    } else {
      const int4 bm = A.bmeta[A.ac[p]];
      qs = (carry && !first) ? cur[p] : bm.x;
      h = bm.y;
      bmax = bm.w;
    }
    int l = qs;
    if (carry && l < h && bmax >= wend && T == 1) {
      // one lane per entry (chunks of more than kHeavyT / 2 entries, such as
      // the hub rows' ones): gallop from the cursor, since a window usually
      // ends a few positions past it, then bisect the last step (about
      // 2 log2 of the positions consumed instead of log2 of the row)
      int64_t step = 1;
      while (l < h) {
        const int p = (int64_t)(h - l) > step ? l + (int)step - 1 : h - 1;
        if ((int64_t)A.bc[p] < wend) {
          l = p + 1;
          step <<= 1;
        } else {
          h = p;
          break;
        }
      }
      while (l < h) {
        const int mid = l + ((h - l) >> 1);
        if ((int64_t)A.bc[mid] < wend) l = mid + 1;
        else h = mid;
      }
    } else if (carry && l < h && bmax >= wend) {
      // first position of the B row at or past the window end; l, h are
      // uniform over the group, whose lanes are active together
      const int base = (tid & (kWave - 1)) & ~(T - 1);
      const uint64_t gmask = T == kWave ? ~0ull : ((1ull << T) - 1ull);
      while (l < h) {
        const int len = h - l;
        const bool small = len <= T;
        const bool valid = !small || sub < len;
        const int pos = small ? (valid ? l + sub : l) : l + (int)((int64_t)len * (sub + 1) / (T + 1));
        const bool below = valid && (int64_t)A.bc[pos] < wend;
        const int c = __popcll((__ballot(below) >> base) & gmask);
        if (small) {
          l += c;
          h = l;
        } else {
          const int nl = c > 0 ? l + (int)((int64_t)len * c / (T + 1)) + 1 : l;
          const int nh = c < T ? l + (int)((int64_t)len * (c + 1) / (T + 1)) : h;
          l = nl;
          h = nh;
        }
      }
    } else {
      l = h;
    }
    if (sub == 0) {
      hs.qs[ent] = qs;
      hs.off[ent] = l;
      if (carry && keep) {
        if (res.cur) res.cur[ent] = l;
        else cur[p] = l;
      }
      if (VALS && !res.cur) hs.a[ent] = A.av[p];
    }
  }
}

enum HeavyMode { kHeavyOr = 0, kHeavySlot = 1, kHeavyDense = 2 };
constexpr int kHeavyU = 4;  // products in flight per thread in heavy_products

// A resident row's structure pass over a window of at most kHoldSteps steps
// of the product loop (kHoldSteps * kHeavyU * kHeavyT products) also forms
// the products' values and keeps them, with their columns and owners, in the
// thread's registers: once the window's slots (and its cut) are known they
// are added from there, with no value pass.
constexpr int kHoldSteps = 2;                       // steps a window's held products may take
constexpr int kHoldN = kHoldSteps * kHeavyU;         // products held per thread
struct HeavyHold {
  int32_t j[kHoldN];    // column, -1: none
  double v[kHoldN];
  int32_t l[kHoldN];    // owner (A entry of the chunk)
  bool on;              // every product of the window is held
};

template <int MODE>
__device__ __forceinline__ void heavy_products(const RowArgs& A, HeavyStage& hs, int i, int s, int e, int64_t w0,
                                               int64_t wend, bool carry, bool first, bool keep, int32_t* cur,
                                               uint32_t* bits, const uint32_t* pre, double* vals,
                                               HeavyRes res = HeavyRes{nullptr, nullptr, nullptr},
                                               bool reuse = false, HeavyHold* hold = nullptr,
                                               int32_t* ccol_w = nullptr, int64_t clim = 0) {
  // carry: B rows sorted and cursors kept across windows (heavy_stage_ranges).
  // reuse (resident rows, one chunk): hs still holds this window's ranges and
  // their scan from the previous call, whose window ended at the same column.
  // ccol_w (slot mode): the window's first output column entry; every
  // product stores its column at its slot there (slots below clim), so the
  // write-out only copies values.
  const int tid = threadIdx.x;
  for (int cb = s; cb < e; cb += kHeavyT) {
    const int m = (e - cb) < kHeavyT ? (e - cb) : kHeavyT;
    int64_t tot;
    if (!reuse) {
      heavy_stage_ranges<MODE != kHeavyOr>(A, hs, cb, m, wend, carry, first, keep, cur, res);
      __syncthreads();
      const int qs = tid < m ? hs.qs[tid] : 0;
      const int qe = tid < m ? (int)hs.off[tid] : 0;
      const int64_t ex = block_scan_excl<int64_t>((int64_t)(qe - qs), hs.red, &tot);
      hs.off[tid] = ex;
      if (tid == 0) hs.tot = tot;
      __syncthreads();
    } else {
      tot = hs.tot;
      // cursors move past the window: start + segment length
      if (keep && tid < m)
        res.cur[tid] = hs.qs[tid] + (int)((tid + 1 < m ? hs.off[tid + 1] : tot) - hs.off[tid]);
    }
    if (hold) {
      hold->on = res.cur != nullptr && e - s <= kHeavyT && tot <= (int64_t)kHoldN * kHeavyT;
#pragma unroll
      for (int u = 0; u < kHoldN; u++) hold->j[u] = -1;
    }
    // kHeavyU products per thread per step (t, t + 1024, ...): their owner
    // searches, gathers and LDS updates are issued together.  The owner of
    // t is the last entry whose segment starts at or before t, found by a
    // fixed-depth search (l = 0 holds: hs.off[0] == 0).
    int depth = 0;
    while ((1 << depth) < m) depth++;
    // one step from t0; with VH the values are formed for the structure
    // pass too and the step's columns, values and owners are returned
    auto step = [&](int64_t t0, auto vh, int32_t* jo, double* vo, int32_t* lo) {
      constexpr bool VH = decltype(vh)::value;
      int64_t tt[kHeavyU];
      int l[kHeavyU];
#pragma unroll
      for (int u = 0; u < kHeavyU; u++) {
        tt[u] = t0 + (int64_t)u * kHeavyT;
        l[u] = 0;
      }
      for (int d = depth - 1; d >= 0; d--) {
#pragma unroll
        for (int u = 0; u < kHeavyU; u++) {
          const int cand = l[u] + (1 << d);
          if (cand < m && hs.off[cand] <= tt[u]) l[u] = cand;
        }
      }
      int q[kHeavyU], j[kHeavyU];
      double pv[kHeavyU];
#pragma unroll
      for (int u = 0; u < kHeavyU; u++) {
        // steps past tot read B position qs of their owner's (in-range) start
        const bool ok = tt[u] < tot;
        q[u] = hs.qs[l[u]] + (ok ? (int)(tt[u] - hs.off[l[u]]) : 0);
        j[u] = ok ? A.bc[q[u]] : -1;
        if (MODE != kHeavyOr || VH) pv[u] = ok ? hs.a[l[u]] * A.bv[q[u]] : 0.0;
      }
      if constexpr (VH) {
#pragma unroll
        for (int u = 0; u < kHeavyU; u++) {
          jo[u] = j[u];
          vo[u] = pv[u];
          lo[u] = l[u];
        }
      }
#pragma unroll
      for (int u = 0; u < kHeavyU; u++) {
        if (j[u] < w0 || j[u] >= wend) continue;  // past tot, or B rows rescanned
        const uint32_t b = (uint32_t)(j[u] - w0);
        if (MODE == kHeavyOr) {
          atomicOr(&bits[b >> 5], 1u << (b & 31u));
        } else if (MODE == kHeavyDense) {
          atomicOr(&bits[b >> 5], 1u << (b & 31u));
          if (!A.upper || j[u] >= i) atomicAdd(&vals[b], pv[u]);
        } else {
          const uint32_t w = b >> 5;
          const uint32_t pos = pre[w] + __popc(bits[w] & ((1u << (b & 31u)) - 1u));
          if (ccol_w != nullptr && (int64_t)pos < clim) ccol_w[pos] = j[u];
          if (!A.upper || j[u] >= i) atomicAdd(&vals[pos], pv[u]);
        }
      }
    };
    if (MODE == kHeavyOr && hold && hold->on) {
      // every step of the window, unrolled so the held products stay in
      // registers
#pragma unroll
      for (int it = 0; it < kHoldSteps; it++) {
        const int64_t t0 = tid + (int64_t)it * kHeavyU * kHeavyT;
        if (t0 < tot)
          step(t0, std::true_type{}, hold->j + it * kHeavyU, hold->v + it * kHeavyU, hold->l + it * kHeavyU);
      }
    } else {
      for (int64_t t0 = tid; t0 < tot; t0 += (int64_t)kHeavyU * kHeavyT)
        step(t0, std::false_type{}, nullptr, nullptr, nullptr);
    }
    __syncthreads();
  }
}

// Symbolic heavy rows: windows of `wc` columns, a bitmap each; the row's
// count and distinct words are summed over the windows.
__global__ void __launch_bounds__(kHeavyT) k_heavy_sym(RowArgs A, int32_t wc, int32_t* __restrict__ work,
                                                       const int32_t* __restrict__ bflag, int32_t* __restrict__ cur,
                                                       int32_t* __restrict__ rowcnt) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ HeavyStage hs;
  uint32_t* bits = (uint32_t*)lds_raw;
  const int tid = threadIdx.x;
  const bool sorted = bflag[0] == 0;
  while (true) {
    __syncthreads();
    if (tid == 0) hs.row = atomicAdd(work, 1);
    __syncthreads();
    const int r = hs.row;
    if (r >= A.nrows) break;
    const int i = A.rows[r];
    const int lo = A.rowlo[i], hi = A.rowhi[i];
    const int s = A.ap[i], e = A.ap[i + 1];
    const int npass = (int)(((int64_t)hi - lo + wc) / wc);
    const bool carry = sorted && npass > 1;
    int64_t cnt = 0, nzw = 0;
    for (int pass = 0; pass < npass; pass++) {
      const int64_t w0 = (int64_t)lo + (int64_t)pass * wc;
      const int64_t w1 = (w0 + wc) < ((int64_t)hi + 1) ? (w0 + wc) : ((int64_t)hi + 1);
      const int words = (int)((w1 - w0 + 31) >> 5);
      for (int t = tid; t < words; t += kHeavyT) bits[t] = 0u;
      heavy_products<kHeavyOr>(A, hs, -1, s, e, w0, w1, carry, pass == 0, pass + 1 < npass, cur, bits, nullptr,
                               nullptr);
      int64_t c = 0, z = 0;
      for (int t = tid; t < words; t += kHeavyT) {
        const uint32_t x = bits[t];
        c += __popc(x);
        z += x != 0u ? 1 : 0;
      }
      int64_t tc, tz;
      (void)block_scan_excl<int64_t>(c, hs.red, &tc);
      (void)block_scan_excl<int64_t>(z, hs.red, &tz);
      cnt += tc;
      nzw += tz;
      __syncthreads();
    }
    if (tid == 0) {
      rowcnt[i] = (int32_t)cnt;
      A.rowwords[i] = (int32_t)nzw | kNoWordList;
    }
  }
}
// Numeric heavy rows.  Two window shapes, chosen per row:
// * dense: a window of up to dcols columns (without a tuning cap the width
//   heavy_dense_cols gives for the LDS left by the row's resident entries),
//   one f64 slot per column; one product pass sets the bits and
//   accumulates, then the window's set columns are compacted into C.
// * sparse: a window of up to `wcols` columns (1 bit plus a 32-bit prefix
//   count per word) holding at most `vcap` outputs: a structure pass sets the
//   window's bits, prefix counts give each column its slot, the window is cut
//   at the last word whose entries still fit, and a value pass accumulates
//   the columns before the cut; the next window starts at the cut.
// Sparse windows cost a second product pass but far fewer windows on wide
// rows (power-law hubs over millions of columns, where dense windows of a
// few 10^4 columns each pay a pass over the row's A entries for cursor
// searches, scans and barriers).  Measured on R-MAT rows: rows spanning more
// than kHeavySparseSpan dense windows run faster sparse (scale 21: 1.87 s
// against 3.22 s), narrower ones dense (scale 16: 5.6 ms against 6.3 ms).
// With lbytes > 0 (the launch's dynamic LDS, no tuning cap) each sparse
// window takes its width and slots from the row's remaining outputs and
// columns (heavy_row_shape): a scattered row (some 10^4 outputs over 2^21
// columns) gets a few windows of several 10^5 columns instead of 16 of
// 131,072, and a dense row windows whose structure pass ends near the cut.
// A cut caps the next window at 3/2 of the columns it took.
// Cursors carry over between windows as in k_heavy_sym (in LDS for resident
// rows).  LDS: bits (window columns / 32 words; with a tuning cap the dense
// slots start after wcols / 32 words), then either prefix counts + value
// slots (sparse) or one slot per column (dense), and the resident entries
// at the end.
__global__ void __launch_bounds__(kHeavyT) k_heavy_num(RowArgs A, int32_t wcols, int32_t vcap, int32_t dcols,
                                                       int64_t lbytes,
                                                       int32_t* __restrict__ work, const int32_t* __restrict__ bflag,
                                                       int32_t* __restrict__ cur, const int64_t* __restrict__ rowoff,
                                                       int32_t* __restrict__ ccol, double* __restrict__ cval,
                                                       int32_t* __restrict__ fault) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ HeavyStage hs;
  __shared__ int32_t cut_word;
  uint32_t* bits = (uint32_t*)lds_raw;
  double* dvals = (double*)(lds_raw + (size_t)(wcols / 32) * 4);  // dense windows
  const int tid = threadIdx.x;
  const bool sorted = bflag[0] == 0;
  while (true) {
    __syncthreads();
    if (tid == 0) hs.row = atomicAdd(work, 1);
    __syncthreads();
    const int r = hs.row;
    if (r >= A.nrows) break;
    const int i = A.rows[r];
    const int lo = A.rowlo[i], hi = A.rowhi[i];
    const int s = A.ap[i], e = A.ap[i + 1];
    int64_t out = rowoff[i];
    const int64_t end = rowoff[i + 1];
    // window shape of this row (uniform across the block).  Rows of more
    // than one chunk (hubs) whose outputs cover at least 12 % of their span
    // take dense windows however wide the span: a dense window makes one
    // product pass over each chunk where a sparse one makes two, and a hub
    // row's chunk passes (window-end searches, scans, barriers) cost more
    // than its extra windows (R-MAT 2^21: 28 % 0.838 s, 12 % 0.824-0.837 s, 4 %
    // 0.841 s; for rows of one chunk sparse windows stay faster)
    const int mrow = e - s;
    const int64_t span = (int64_t)hi - lo + 1;
    const bool sparse =
        span > kHeavySparseSpan * dcols && !(lbytes > 0 && mrow > kHeavyT && 100 * (end - out) >= 12 * span);
    // rows of one chunk over sorted B rows keep their entries in LDS
    // (HeavyRes) behind the windows, whose shapes take the rest
    const bool resident = lbytes > 0 && sorted && mrow > 0 && mrow <= kHeavyT;
    const int64_t rbytes = resident ? (((int64_t)12 * mrow + 15) & ~(int64_t)15) : 0;
    HeavyRes res{nullptr, nullptr, nullptr};
    if (resident) {
      res.cur = (int32_t*)(lds_raw + (size_t)(lbytes - rbytes));
      res.be = res.cur + mrow;
      res.bw = res.be + mrow;
      if (tid < mrow) {
        const int4 bm = A.bmeta[A.ac[s + tid]];
        res.cur[tid] = bm.x;
        res.be[tid] = bm.y;
        res.bw[tid] = bm.w;
        hs.a[tid] = A.av[s + tid];
      }
      __syncthreads();
    }
    // dense windows without a tuning cap: a bitmap of the window's own
    // words, f64 slots, and the resident entries share the LDS (65 / 8 bytes
    // per column), about 12 % wider than dcols
    int32_t dcr = dcols;
    double* dv = dvals;
    if (!sparse && lbytes > 0) {
      dcr = heavy_dense_cols(lbytes - rbytes);
      dv = (double*)(lds_raw + (size_t)(dcr / 32) * 4);
    }
    int64_t w0 = lo, wlim = 0;
    bool first = true;
    while (w0 <= hi) {
      // sparse windows: sized per window from the row's remaining outputs
      // and columns (heavy_row_shape) unless a tuning cap fixed the shape
      int32_t wcr = wcols, vcr = vcap;
      if (sparse && lbytes > 0)
        heavy_row_shape(lbytes - rbytes, end - out, (int64_t)hi + 1 - w0, wlim, &wcr, &vcr);
      uint32_t* pre = bits + wcr / 32;
      double* vals = sparse ? (double*)(lds_raw + (size_t)(wcr / 32) * 8) : dv;
      const int vlim = sparse ? vcr : dcr;
      const int64_t wspan = sparse ? wcr : dcr;
      const int64_t w1 = (w0 + wspan) < ((int64_t)hi + 1) ? (w0 + wspan) : ((int64_t)hi + 1);
      const int words = (int)((w1 - w0 + 31) >> 5);
      for (int t = tid; t < words; t += kHeavyT) bits[t] = 0u;
      int cw = words;
      int64_t wcut = w1;
      if (sparse) {
        if (tid == 0) cut_word = words;
        // structure of [w0, w1); cursors are not advanced.  A resident
        // row's window of one product step keeps its values (HeavyHold)
        HeavyHold hold;
        hold.on = false;
        heavy_products<kHeavyOr>(A, hs, i, s, e, w0, w1, sorted, first, false, cur, bits, pre, vals, res, false,
                                 &hold);
        // slots: exclusive prefix of the words' popcounts; the cut is the
        // first word whose entries would not all fit in vcap.  Each thread
        // takes a run of consecutive words, so a window of any width costs
        // one block scan
        const int per = (words + kHeavyT - 1) / kHeavyT;
        const int wa = tid * per < words ? tid * per : words;
        const int wz = wa + per < words ? wa + per : words;
        int64_t loc = 0;
        for (int w = wa; w < wz; w++) loc += __popc(bits[w]);
        int64_t run;
        int64_t p0 = block_scan_excl<int64_t>(loc, hs.red, &run);
        for (int w = wa; w < wz; w++) {
          const int pc = __popc(bits[w]);
          pre[w] = (uint32_t)(p0 < (int64_t)0xffffffffu ? p0 : 0xffffffffu);
          if (p0 <= vcr && p0 + pc > vcr) cut_word = w;  // unique: counts are monotone
          p0 += pc;
        }
        __syncthreads();
        cw = cut_word;
        if (cw == 0) {
          // the first word's outputs exceed vcap: no progress is possible.
          // heavy_num_shape keeps vcap >= 32, so this is a guard only; the
          // block abandons the row and the host reports the fault
          if (tid == 0) *fault = 1;
          break;
        }
        if (cw < words) wcut = w0 + (int64_t)cw * 32;
        const int nvals = cw < words ? (int)pre[cw] : (int)(run < vcr ? run : vcr);
        for (int t = tid; t < nvals; t += kHeavyT) vals[t] = 0.0;
        __syncthreads();
        if (hold.on) {
          // held products: those before the cut go to their slots.  The
          // cursors move past the window (start + segment length) or, for a
          // cut window, past each entry's products before the cut, counted
          // into hs.off, whose scan is not needed any more: B rows are
          // sorted, so those products are a prefix of the entry's segment
          const bool cut = cw < words;
          int32_t* cnt = (int32_t*)hs.off;
          if (cut) {
            if (tid < mrow) cnt[tid] = 0;
            __syncthreads();
          }
#pragma unroll
          for (int u = 0; u < kHoldN; u++) {
            const int64_t jj = hold.j[u];
            if (jj >= w0 && jj < wcut) {
              const uint32_t b = (uint32_t)(jj - w0);
              const uint32_t w = b >> 5;
              const uint32_t pos = pre[w] + __popc(bits[w] & ((1u << (b & 31u)) - 1u));
              // the column at its slot (products of one column store the same value)
              if (out + (int64_t)pos < end) ccol[out + pos] = (int32_t)jj;
              if (!A.upper || jj >= i) atomicAdd(&vals[pos], hold.v[u]);
              if (cut) atomicAdd(&cnt[hold.l[u]], 1);
            }
          }
          if (!cut && tid < mrow)
            res.cur[tid] = hs.qs[tid] + (int)((tid + 1 < mrow ? hs.off[tid + 1] : hs.tot) - hs.off[tid]);
          __syncthreads();
          if (cut) {
            if (tid < mrow) res.cur[tid] = hs.qs[tid] + cnt[tid];
            __syncthreads();
          }
        } else {
          // values of [w0, wcut); cursors move to the cut.  A resident row's
          // uncut window reuses the structure pass's ranges and their scan
          heavy_products<kHeavySlot>(A, hs, i, s, e, w0, wcut, sorted, first, true, cur, bits, pre, vals, res,
                                     resident && cw == words, nullptr, ccol + out, end - out);
        }
      } else {
        for (int t = tid; t < (int)(w1 - w0); t += kHeavyT) vals[t] = 0.0;
        heavy_products<kHeavyDense>(A, hs, i, s, e, w0, w1, sorted, first, true, cur, bits, pre, vals, res);
      }
      // columns and values of the words before the cut, in order
      int64_t run = 0;
      if (sparse) {
        // the columns are at their slots already (stored by the pass that
        // added the values): the values are copied out in slot order
        run = cw < words ? (int64_t)pre[cw] : (int64_t)pre[words - 1] + __popc(bits[words - 1]);
        for (int64_t k = tid; k < run; k += kHeavyT) {
          const int64_t o = out + k;
          if (o < end && k < vlim) cval[o] = vals[k];
        }
      } else if (cw <= kHeavyT) {
        // dense window of at most kHeavyT words: word prefixes (parked in
        // hs.off, free after the product pass), then one thread per column,
        // so consecutive set columns are stored to consecutive positions
        const uint32_t x = tid < cw ? bits[tid] : 0u;
        const int64_t p0 = block_scan_excl<int64_t>((int64_t)__popc(x), hs.red, &run);
        hs.off[tid] = p0;
        __syncthreads();
        const int ncol = cw * 32;
        for (int c = tid; c < ncol; c += kHeavyT) {
          const uint32_t y = bits[c >> 5];
          if ((y >> (c & 31)) & 1u) {
            const int64_t o = out + hs.off[c >> 5] + __popc(y & ((1u << (c & 31)) - 1u));
            if (o < end && c < vlim) {
              ccol[o] = (int32_t)(w0 + c);
              cval[o] = vals[c];
            }
          }
        }
      } else {
        for (int wb = 0; wb < cw; wb += kHeavyT) {
          const int w = wb + tid;
          const uint32_t x = w < cw ? bits[w] : 0u;
          int64_t tot;
          const int64_t p0 = run + block_scan_excl<int64_t>((int64_t)__popc(x), hs.red, &tot);
          if (w < cw) {
            uint32_t y = x;
            int64_t o = out + p0;
            while (y) {
              const int bpos = __builtin_ctz(y);
              const int slot = w * 32 + bpos;   // dense: the column offset
              if (o < end && slot < vlim) {
                ccol[o] = (int32_t)(w0 + (int64_t)w * 32 + bpos);
                cval[o] = vals[slot];
              }
              o++;
              y &= y - 1u;
            }
          }
          run += tot;
        }
      }
      out += run;
      __syncthreads();
      if (sparse) wlim = heavy_next_wlim(cw < words, wcut - w0, wlim);
      w0 = wcut;
      first = false;
    }
  }
}

// ---------------------------------------------------------------------------
// Window kernels (dense mode of the wave-level bins).  Persistent: each block
// walks rows of the bin's row list (row_walk: XCD-local ranges, tiles dealt
// to the XCD's blocks in turn), one row per group of G lanes at a time.  Per group, LDS holds a
// bitmap over the row's output window, kept all-zero between rows by
// clearing only the words the row touched (recorded in a touched list), so
// the per-row cost does not grow with the window width.  The product loop is
// branch-light: A-entry metadata is staged in LDS and read as broadcasts, B
// gathers use 32-bit offsets and are issued for every lane, and lanes without
// a product update a private dummy slot instead of branching.
// ---------------------------------------------------------------------------

constexpr int kBatch = 8;     // A entries whose B gathers are issued together
constexpr int kWordCap = kWordListCap;  // words per row kept in the symbolic -> numeric word list

// Stage G A entries of a row chunk: (bs, be) of their B rows (zero-padded so
// a batch may read past the chunk end) and, for the value pass, a_ik.
// Returns true when some staged B row is longer than G.
template <int G, bool VALS>
__device__ __forceinline__ bool stage_entries(int base, int e, int lane, const RowArgs& A, int2* mbe, double* ma) {
  const int p = base + lane;
  int2 v = make_int2(0, 0);
  double a = 0.0;
  if (p < e) {
    const int4 bm = A.bmeta[A.ac[p]];
    v = make_int2(bm.x, bm.y);
    if (VALS) a = A.av[p];
  }
  mbe[lane] = v;
  if (VALS) ma[lane] = a;
  return group_bits<G>(__ballot(v.y - v.x > G)) != 0ull;
}

// Software pipeline over a group's rows.  A row's metadata is a chain of
// dependent loads: rows[r] -> (ap[i], rowlo[i], row head) -> ac[p] ->
// (bmeta[k], av[p]).  At the top of every row iteration, hop h of that chain
// is issued for row r + h*R (h = 1..4) from values requested one iteration
// earlier, so no load waits on another load issued in the same iteration.
struct RowState {
  int i, s, e, lo;
  int k;       // this lane's first-chunk A column (-1: none)
  int2 be;     // B-row start / end of that entry
  double a;    // a_ik of that entry (value pass only)
};

__device__ __forceinline__ void hop_scalars(RowState& st, const RowArgs& A) {
  st.s = A.ap[st.i];
  st.e = A.ap[st.i + 1];
  st.lo = A.rowlo[st.i];
}
__device__ __forceinline__ void hop_entry(RowState& st, int lane, const RowArgs& A) {
  st.k = (st.s + lane < st.e) ? A.ac[st.s + lane] : -1;
}
template <bool VALS>
__device__ __forceinline__ void hop_bmeta(RowState& st, int lane, const RowArgs& A) {
  st.be = make_int2(0, 0);
  st.a = 0.0;
  if (st.k >= 0) {
    const int4 bm = A.bmeta[st.k];
    st.be = make_int2(bm.x, bm.y);
    if (VALS) st.a = A.av[st.s + lane];
  }
}

// Structure pass for one chunk whose entries are staged in mbe (m entries,
// zero-padded to G): OR every product's bit into the window, recording
// first-touched words when TRACK.  All gathers of the chunk are issued
// before the LDS updates.
template <int G, bool UPPER, bool W64, bool TRACK>
__device__ __forceinline__ void or_chunk(int i, int lo, int lane, int m, bool lng, const RowArgs& A,
                                         const int2* mbe, uint32_t* bits, int W, uint32_t* touched, int* ctl) {
  constexpr int B = G < 32 ? G : 32;
  static_assert(G % B == 0, "chunk depth must divide the group size");
  auto apply = [&](int j, bool ok) {
    const bool use = ok;
    const uint32_t b = (uint32_t)(j - lo);
    const uint32_t wd = (b >> 5) < (uint32_t)W ? (b >> 5) : (uint32_t)W;  // in range by construction
    // the symbolic window (TRACK) stores its bitmap bank-skewed; the numeric
    // dense window does not
    const uint32_t slot = TRACK ? (use ? bank_skew(wd) : (uint32_t)(skewed_words(W) + lane))
                                : (use ? wd : (uint32_t)(W + lane));
    const uint32_t old = atomicOr(&bits[slot], 1u << (b & 31u));
    if (TRACK && __builtin_expect(use && old == 0u, 0)) {
      const int idx = atomicAdd(&ctl[0], 1);
      if (idx < kTouchCap) touched[idx] = b >> 5;
    }
  };
  for (int t0 = 0; t0 < m; t0 += B) {
    int jj[B];
    uint32_t okm = 0u;
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int2 be = mbe[t0 + u];
      const int q = be.x + lane;
      const bool ok = q < be.y;
      okm |= (ok ? 1u : 0u) << u;
      jj[u] = ld_idx<int32_t, W64>(A.bc, ok ? (uint32_t)q : 0u);
    }
    // the chunk's atomic ORs are issued back to back; first touches are
    // recorded after all of them returned
    uint32_t old[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const bool use = (okm >> u) & 1u;
      const uint32_t b = (uint32_t)(jj[u] - lo);
      const uint32_t wd = (b >> 5) < (uint32_t)W ? (b >> 5) : (uint32_t)W;  // in range by construction
      const uint32_t slot = TRACK ? (use ? bank_skew(wd) : (uint32_t)(skewed_words(W) + lane))
                                  : (use ? wd : (uint32_t)(W + lane));
      old[u] = atomicOr(&bits[slot], 1u << (b & 31u));
    }
    if (TRACK) {
#pragma unroll
      for (int u = 0; u < B; u++) {
        if (__builtin_expect(((okm >> u) & 1u) && old[u] == 0u, 0)) {
          const int idx = atomicAdd(&ctl[0], 1);
          if (idx < kTouchCap) touched[idx] = (uint32_t)(jj[u] - lo) >> 5;
        }
      }
    }
  }
  if (lng) {
    for (int t = 0; t < m; t++) {
      const int2 be = mbe[t];
      for (int q = be.x + G + lane; q - lane < be.y; q += G) {
        const bool ok = q < be.y;
        apply(ld_idx<int32_t, W64>(A.bc, ok ? (uint32_t)q : 0u), ok);
      }
    }
  }
}

// Value pass for one staged chunk (mbe, ma): a_ik * b_kj added at its final
// sorted position, in A-row storage order per output column.
template <int G, bool UPPER, bool W64, bool LIST>
__device__ __forceinline__ void value_chunk(int i, int lo, int lane, int m, bool lng, const RowArgs& A,
                                            const int2* mbe, const double* ma, const uint32_t* bits,
                                            const uint16_t* pre, const uint8_t* widx, const uint2* wtab,
                                            double* vals, int Cm, int W) {
  // gathers in flight per lane; the list variant trades depth for occupancy
  // (84 VGPRs, 5 waves per SIMD: 4 per chunk beat 8 at 4 waves per SIMD)
  constexpr int kDepth = LIST ? 4 : 16;
  constexpr int B = G < kDepth ? G : kDepth;
  // chunks tile the G staged entries exactly (entries past the row's count
  // are zero-padded, never read past the staging)
  static_assert(G % B == 0, "chunk depth must divide the group size");
  auto add = [&](int j, double prod, bool ok) {
    const bool use = ok && (!UPPER || j >= i);
    const uint32_t bb = use ? (uint32_t)(j - lo) : 0u;
    // valid operands keep every index below in range: the word inside the
    // window, the listed entry below list_cap(G), the position below the
    // row's count (appt_csr_check is the check for untrusted input)
    const uint32_t w = bb >> 5;
    uint32_t pos;
    if constexpr (LIST) {
      const uint2 te = wtab[widx[w]];  // word -> listed entry (mask, first position)
      pos = te.y + __popc(te.x & ((1u << (bb & 31u)) - 1u));
    } else {
      pos = (uint32_t)pre[w] + __popc(bits[w] & ((1u << (bb & 31u)) - 1u));
    }
    const uint32_t slot = use ? pos : (uint32_t)(Cm + lane);
    atomicAdd(&vals[slot], prod);
  };
  if (!lng) {
    for (int t0 = 0; t0 < m; t0 += B) {
      int jj[B];
      double pv[B];
      uint32_t okm = 0u;
#pragma unroll
      for (int u = 0; u < B; u++) {
        const int2 be = mbe[t0 + u];
        const int q = be.x + lane;
        const bool ok = q < be.y;
        okm |= (ok ? 1u : 0u) << u;
        const uint32_t qq = ok ? (uint32_t)q : 0u;
        jj[u] = ld_idx<int32_t, W64>(A.bc, qq);
        pv[u] = ma[t0 + u] * ld_idx<double, W64>(A.bv, qq);
      }
#pragma unroll
      for (int u = 0; u < B; u++) add(jj[u], pv[u], (okm >> u) & 1u);
    }
  } else {
    for (int t = 0; t < m; t++) {
      const int2 be = mbe[t];
      const double a = ma[t];
      for (int q = be.x + lane; q - lane < be.y; q += G) {
        const bool ok = q < be.y;
        const uint32_t qq = ok ? (uint32_t)q : 0u;
        add(ld_idx<int32_t, W64>(A.bc, qq), a * ld_idx<double, W64>(A.bv, qq), ok);
      }
    }
  }
}

// Symbolic window: row count, distinct words, and (when it fits) the row's
// sorted word list (word, mask) for the numeric kernel.
template <int G, bool UPPER, bool W64>
__global__ void __launch_bounds__(256) k_sym_window(RowArgs A, int32_t W, int32_t* __restrict__ rowcnt) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  unsigned char* base = lds_raw + (size_t)gi * sym_window_bytes(G, W);
  int2* mbe = (int2*)base;
  uint32_t* bits = (uint32_t*)(base + align16(G * 8));
  uint32_t* touched = (uint32_t*)((unsigned char*)bits + align16((skewed_words(W) + G) * 4));
  int* ctl = (int*)((unsigned char*)touched + align16(kTouchCap * 4));
  for (int t = lane; t < skewed_words(W) + G; t += G) bits[t] = 0u;
  // unused touched entries hold a sentinel above every word index, so the
  // rank loop reads them four at a time without bounds tests
  for (int t = lane; t < kTouchCap; t += G) touched[t] = ~0u;
  const RowWalk walk = row_walk(A.nrows, R, gi);
  const int S = walk.step, r1 = walk.end;
  int r = walk.first;
  // pipeline prologue: rows r .. r+3R filled to hop depths 4, 3, 2, 1
  RowState q0{}, q1{}, q2{}, q3{};
  q0.k = q1.k = q2.k = q3.k = -1;
  if (r < r1) { q0.i = A.rows[r]; hop_scalars(q0, A); hop_entry(q0, lane, A); hop_bmeta<false>(q0, lane, A); }
  if (r + S < r1) { q1.i = A.rows[r + S]; hop_scalars(q1, A); hop_entry(q1, lane, A); }
  if (r + 2 * S < r1) { q2.i = A.rows[r + 2 * S]; hop_scalars(q2, A); }
  if (r + 3 * S < r1) q3.i = A.rows[r + 3 * S];
  gsync<G>();
  for (; r < r1; r += S) {
    // hops for the rows ahead (each uses values requested one iteration ago)
    RowState q4{};
    q4.k = -1;
    if (r + 4 * S < r1) q4.i = A.rows[r + 4 * S];
    if (r + 3 * S < r1) hop_scalars(q3, A);
    if (r + 2 * S < r1) hop_entry(q2, lane, A);
    if (r + S < r1) hop_bmeta<false>(q1, lane, A);
    const int i = q0.i;
    const int lo = q0.lo;
    if (lane == 0) ctl[0] = 0;
    mbe[lane] = q0.be;
    bool lng = group_bits<G>(__ballot(q0.be.y - q0.be.x > G)) != 0ull;
    gsync<G>();
    int m = (q0.e - q0.s) < G ? (q0.e - q0.s) : G;
    or_chunk<G, UPPER, W64, true>(i, lo, lane, m, lng, A, mbe, bits, W, touched, ctl);
    gsync<G>();
    for (int b0 = q0.s + G; b0 < q0.e; b0 += G) {                  // long A rows
      const bool l2 = stage_entries<G, false>(b0, q0.e, lane, A, mbe, nullptr);
      gsync<G>();
      m = (q0.e - b0) < G ? (q0.e - b0) : G;
      or_chunk<G, UPPER, W64, true>(i, lo, lane, m, l2, A, mbe, bits, W, touched, ctl);
      gsync<G>();
    }
    const int nt = ctl[0];
    int c = 0;
    if (nt <= kTouchCap) {
      constexpr int PL = (kTouchCap + G - 1) / G;
      uint32_t myw[PL], mym[PL];
#pragma unroll
      for (int k = 0; k < PL; k++) {
        const int u = lane + k * G;
        myw[k] = u < nt ? touched[u] : 0u;
        mym[k] = u < nt ? bits[bank_skew(myw[k])] : 0u;
        c += __popc(mym[k]);
      }
      if (nt <= kWordCap) {
        uint2* wl = A.wlist + (size_t)i * kWordCap;
        // rank = touched words below this lane's; each 16-byte broadcast
        // read serves all of the lane's slots (entries past nt are ~0u)
        const uint4* t4 = (const uint4*)touched;
        int rank[PL];
#pragma unroll
        for (int k = 0; k < PL; k++) rank[k] = 0;
        for (int v = 0; v < nt; v += 4) {
          const uint4 t = t4[v >> 2];
#pragma unroll
          for (int k = 0; k < PL; k++)
            rank[k] += (int)(t.x < myw[k]) + (int)(t.y < myw[k]) + (int)(t.z < myw[k]) + (int)(t.w < myw[k]);
        }
#pragma unroll
        for (int k = 0; k < PL; k++)
          if (lane + k * G < nt) wl[rank[k]] = make_uint2(myw[k], mym[k]);
      }
      gsync<G>();
#pragma unroll
      for (int k = 0; k < PL; k++)
        if (lane + k * G < nt) {
          bits[bank_skew(myw[k])] = 0u;
          touched[lane + k * G] = ~0u;
        }
    } else {
      const int words = (int)span_words(lo, A.rowhi[i]);
      for (int w = lane; w < words; w += G) {
        const uint32_t sw = bank_skew((uint32_t)w);
        c += __popc(bits[sw]);
        bits[sw] = 0u;
      }
      for (int t = lane; t < kTouchCap; t += G) touched[t] = ~0u;
    }
    c = gsum<G, int32_t>(c, nullptr);
    if (lane == 0) { rowcnt[i] = c; A.rowwords[i] = nt; }
    q0 = q1; q1 = q2; q2 = q3; q3 = q4;
    gsync<G>();
  }
}

// Symbolic window for bins whose output spans are at most kNarrowWords
// words (banded operands: a B row's consecutive columns put most lanes of a
// product step in one or two words).  One shared bitmap would serialise
// those lanes' atomic ORs on one address; each group instead keeps kNarrowC
// copies of its row bitmap, lane l ORing into copy l % kNarrowC, so the
// lanes of one word hit consecutive banks.  The copies are merged once per
// row (copy order rotated per word, one bank per lane), and the row's word
// list is the merged nonzero words, in order.  One row per group.
constexpr int kNarrowWords = 8;
template <int G>
__global__ void __launch_bounds__(256) k_sym_narrow(RowArgs A, int32_t* __restrict__ rowcnt) {
  constexpr int C = G < 32 ? G : 32;
  __shared__ uint32_t copies[256 / G][kNarrowWords * C];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = xcd_block(blockIdx.x, gridDim.x) * R + gi;
  if (r >= A.nrows) return;
  const int i = A.rows[r];
  uint32_t* bm = copies[gi];
  for (int t = lane; t < kNarrowWords * C; t += G) bm[t] = 0u;
  gsync<G>();
  const int lo = A.rowlo[i];
  const int cp = lane % C;
  for_each_product<G, false>(i, lane, A, nullptr, [&](int j, double) {
    const uint32_t b = (uint32_t)(j - lo);
    atomicOr(&bm[(b >> 5) * C + cp], 1u << (b & 31u));
  });
  gsync<G>();
  const int words = (int)span_words(lo, A.rowhi[i]);   // <= kNarrowWords by the launch
  uint32_t m = 0u;
  if (lane < words)
    for (int c = 0; c < C; c++) m |= bm[lane * C + (c + lane) % C];
  const uint64_t nz = group_bits<G>(__ballot(m != 0u));
  const int rank = __popcll(nz & ((1ull << lane) - 1ull));
  if (m != 0u) A.wlist[(size_t)i * kWordCap + rank] = make_uint2((uint32_t)lane, m);
  const int cnt = gsum<G, int32_t>(__popc(m), nullptr);
  if (lane == 0) {
    rowcnt[i] = cnt;
    A.rowwords[i] = __popcll(nz);
  }
}

template <int G, bool UPPER, bool W64, bool LIST>
__global__ void __launch_bounds__(256) k_num_window(RowArgs A, int32_t W, int32_t Cm,
                                                    const int64_t* __restrict__ rowoff, int32_t* __restrict__ ccol,
                                                    double* __restrict__ cval) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  unsigned char* base = lds_raw + (size_t)gi * (LIST ? num_list_bytes(G, W, Cm) : num_window_bytes(G, W, Cm));
  int2* mbe = (int2*)base;
  double* ma = (double*)(base + G * 8);
  // dense layout: bitmap | prefix counts | vals | cols
  // list layout:  word index | listed words | vals (columns go straight to C)
  uint32_t* bits = nullptr;
  uint16_t* pre = nullptr;
  int32_t* cols = nullptr;
  uint8_t* widx = nullptr;
  uint2* wtab = nullptr;
  double* vals;
  if constexpr (LIST) {
    widx = base + align16(G * 16);
    wtab = (uint2*)(widx + align16(W + G));
    vals = (double*)((unsigned char*)wtab + align16(list_cap(G) * 8));
  } else {
    bits = (uint32_t*)(base + align16(G * 16));
    pre = (uint16_t*)((unsigned char*)bits + align16((W + G) * 4));
    vals = (double*)((unsigned char*)pre + align16(W * 2));
    cols = (int32_t*)((unsigned char*)vals + align16((Cm + G) * 8));
    for (int t = lane; t < W + G; t += G) bits[t] = 0u;
  }
  const RowWalk walk = row_walk(A.nrows, R, gi);
  const int S = walk.step, r1 = walk.end;
  constexpr int kCap = list_cap(G);
  constexpr int PL = (kCap + G - 1) / G;
  // per-row output head: offsets, word count, sorted word list
  struct Head {
    int64_t off;
    int c, rw;
    uint2 ent[PL];
  };
  auto load_head = [&](int i, Head& h) {  // scalars; the list one hop later
    h.off = rowoff[i];
    h.c = (int)(rowoff[i + 1] - h.off);
    h.rw = A.rowwords[i];
  };
  auto load_list = [&](int i, Head& h) {  // only the listed entries
    const int nt = (h.rw & kNoWordList) ? 0 : (h.rw & kWordsMask);
    const uint2* wl = A.wlist + (size_t)i * kWordCap;
#pragma unroll
    for (int k = 0; k < PL; k++) {
      const int u = lane + k * G;
      h.ent[k] = (u < nt && u < kCap) ? wl[u] : make_uint2(0u, 0u);
    }
  };
  int r = walk.first;
  // pipeline prologue: rows r .. r+3R filled to hop depths 4, 3, 2, 1
  RowState q0{}, q1{}, q2{}, q3{};
  Head h0{}, h1{}, h2{}, h3{};
  q0.k = q1.k = q2.k = q3.k = -1;
  if (r < r1) {
    q0.i = A.rows[r]; hop_scalars(q0, A); load_head(q0.i, h0); load_list(q0.i, h0);
    hop_entry(q0, lane, A); hop_bmeta<true>(q0, lane, A);
  }
  if (r + S < r1) {
    q1.i = A.rows[r + S]; hop_scalars(q1, A); load_head(q1.i, h1); load_list(q1.i, h1); hop_entry(q1, lane, A);
  }
  if (r + 2 * S < r1) { q2.i = A.rows[r + 2 * S]; hop_scalars(q2, A); load_head(q2.i, h2); }
  if (r + 3 * S < r1) q3.i = A.rows[r + 3 * S];
  gsync<G>();
  for (; r < r1; r += S) {
    // hops for the rows ahead (each uses values requested one iteration ago)
    RowState q4{};
    q4.k = -1;
    if (r + 4 * S < r1) q4.i = A.rows[r + 4 * S];
    if (r + 3 * S < r1) { hop_scalars(q3, A); load_head(q3.i, h3); }
    if (r + 2 * S < r1) { hop_entry(q2, lane, A); load_list(q2.i, h2); }
    if (r + S < r1) hop_bmeta<true>(q1, lane, A);
    const int i = q0.i;
    const int lo = q0.lo;
    const int64_t off = h0.off;
    const int c = h0.c;
    const int nt = h0.rw & kWordsMask;
    const bool have_list = (h0.rw & kNoWordList) == 0 && nt <= kCap;
    uint32_t myw[PL];
    for (int t = lane; t < c; t += G) vals[t] = 0.0;
    mbe[lane] = q0.be;
    ma[lane] = q0.a;
    const bool lng = group_bits<G>(__ballot(q0.be.y - q0.be.x > G)) != 0ull;
    if (have_list) {
      // the symbolic pass left this row's words sorted: place them in the
      // window with their prefix popcounts and emit the column list; the
      // entries are strided over lanes (u = lane + k*G), scanned in that order
      int run = 0;
#pragma unroll
      for (int k = 0; k < PL; k++) {
        const int u = lane + k * G;
        const uint2 e = u < nt ? h0.ent[k] : make_uint2(0u, 0u);
        int tot;
        const int ex = gscan_excl<G, int>(__popc(e.y), lane, &tot);
        int rr = run + ex;
        if (u < nt && e.x < (uint32_t)W) {   // in range by construction
          uint32_t x = e.y;
          if constexpr (LIST) {
            widx[e.x] = (uint8_t)u;
            wtab[u] = make_uint2(e.y, (uint32_t)rr);
            while (x && rr < c) {
              ccol[off + rr++] = lo + (int)e.x * 32 + __builtin_ctz(x);
              x &= x - 1u;
            }
          } else {
            bits[e.x] = e.y;
            pre[e.x] = (uint16_t)rr;
            while (x && rr < Cm) {
              cols[rr++] = lo + (int)e.x * 32 + __builtin_ctz(x);
              x &= x - 1u;
            }
          }
        }
        myw[k] = e.x;
        run += tot;
      }
    } else if constexpr (!LIST) {
      gsync<G>();
      const int m0 = (q0.e - q0.s) < G ? (q0.e - q0.s) : G;
      or_chunk<G, UPPER, W64, false>(i, lo, lane, m0, lng, A, mbe, bits, W, nullptr, nullptr);
      gsync<G>();
      for (int b0 = q0.s + G; b0 < q0.e; b0 += G) {
        const bool l2 = stage_entries<G, false>(b0, q0.e, lane, A, mbe, nullptr);
        gsync<G>();
        const int m = (q0.e - b0) < G ? (q0.e - b0) : G;
        or_chunk<G, UPPER, W64, false>(i, lo, lane, m, l2, A, mbe, bits, W, nullptr, nullptr);
        gsync<G>();
      }
      const int words = (int)span_words(lo, A.rowhi[i]);
      const int chw = (words + G - 1) / G;
      const int w0 = lane * chw;
      const int w1 = (w0 + chw) < words ? (w0 + chw) : words;
      int local = 0;
      for (int w = w0; w < w1; w++) local += __popc(bits[w]);
      int tot;
      int run = gscan_excl<G, int>(local, lane, &tot);
      for (int w = w0; w < w1; w++) {
        uint32_t x = bits[w];
        pre[w] = (uint16_t)run;
        while (x) {
          cols[run++] = lo + w * 32 + __builtin_ctz(x);
          x &= x - 1u;
        }
      }
      // restage the first chunk (the structure pass reused mbe)
      mbe[lane] = q0.be;
      ma[lane] = q0.a;
#pragma unroll
      for (int k = 0; k < PL; k++) myw[k] = 0u;
    }
    gsync<G>();
    {
      const int m0 = (q0.e - q0.s) < G ? (q0.e - q0.s) : G;
      value_chunk<G, UPPER, W64, LIST>(i, lo, lane, m0, lng, A, mbe, ma, bits, pre, widx, wtab, vals, Cm, W);
      gsync<G>();
      for (int b0 = q0.s + G; b0 < q0.e; b0 += G) {                // long A rows
        const bool l2 = stage_entries<G, true>(b0, q0.e, lane, A, mbe, ma);
        gsync<G>();
        const int m = (q0.e - b0) < G ? (q0.e - b0) : G;
        value_chunk<G, UPPER, W64, LIST>(i, lo, lane, m, l2, A, mbe, ma, bits, pre, widx, wtab, vals, Cm, W);
        gsync<G>();
      }
    }
    for (int t = lane; t < c; t += G) {
      if constexpr (!LIST) ccol[off + t] = cols[t];
      cval[off + t] = vals[t];
    }
    if constexpr (LIST) {
      // word index entries are rewritten by the next row's placement
    } else if (have_list) {
#pragma unroll
      for (int k = 0; k < PL; k++)
        if (lane + k * G < nt && myw[k] < (uint32_t)W) bits[myw[k]] = 0u;
    } else {
      const int words = (int)span_words(lo, A.rowhi[i]);
      for (int w = lane; w < words; w += G) bits[w] = 0u;
    }
    q0 = q1; q1 = q2; q2 = q3; q3 = q4;
    h0 = h1; h1 = h2; h2 = h3;
    gsync<G>();
  }
}

static int g_num_cus = 256;
void set_num_cus(int n) { g_num_cus = n > 0 ? n : 256; }
int num_cus() { return g_num_cus; }

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static bool bin_dense(int bin) {
  return bin == kBinWgBitmap || (bin > 0 && bin < kBinWgBitmap && bin_is_bitmap(bin));
}

int64_t symbolic_row_lds_bytes(int bin, int32_t metric_max, int32_t words_max) {
  if (bin == kBinHeavy || bin == 0 || bin == kBinTemplate) return 0;
  if (bin_is_sort(bin)) return sort_row_bytes(sort_bin_slots(bin), words_max, 4, false);
  (void)metric_max;
  if (bin == kBinWgBitmap) return (int64_t)words_max * 4;
  if (bin_dense(bin)) return sym_window_bytes(bin_group(bin), words_max);
  return sym_slots(words_max) * 8;
}
int64_t numeric_row_lds_bytes(int bin, int32_t c_max, int32_t words_max) {
  if (bin == kBinHeavy || bin == 0) return 0;
  if (bin == kBinTemplate) return tpl_group_bytes(c_max);
  if (bin_is_sort(bin)) return sort_row_bytes(sort_bin_slots(bin), words_max, 8, true);
  if (bin == kBinWgBitmap) return dense_row_bytes(words_max, c_max);
  if (bin_is_list(bin)) return num_list_bytes(bin_group(bin), words_max, c_max);
  if (bin_dense(bin)) return num_window_bytes(bin_group(bin), words_max, c_max);
  return whash_row_bytes(words_max, c_max);
}

// rows per block so that rows * bytes <= budget, at least 1 and at most 256/G
static int rows_per_block(int G, int64_t row_bytes, int64_t budget) {
  int R = 256 / G;
  if (R < 1) R = 1;
  if (row_bytes > 0) {
    while (R > 1 && (int64_t)R * row_bytes > budget) R >>= 1;
  }
  return R;
}


template <int G, bool DENSE>
static void sym_lds(const RowArgs& a, int32_t words_max, int32_t* rowcnt, hipStream_t s) {
  const int32_t row_words = DENSE ? words_max : (int32_t)(sym_slots(words_max) * 2);
  const int64_t rb = (int64_t)row_words * 4;
  const int R = G > kWave ? 1 : rows_per_block(G, rb, a.lds_budget);
  const size_t lds = (size_t)R * rb;
  const unsigned blocks = (unsigned)((a.nrows + R - 1) / R);
  set_dyn_lds((const void*)k_symbolic_lds<G, DENSE>, lds);
  hipLaunchKernelGGL((k_symbolic_lds<G, DENSE>), dim3(blocks), dim3(R * G), lds, s, a, row_words, rowcnt);
}

template <int G, bool DENSE>
static void num_lds(const RowArgs& a, int32_t words_max, int32_t c_max, const int64_t* rowoff, int32_t* ccol,
                    double* cval, hipStream_t s) {
  const int64_t rb = DENSE ? dense_row_bytes(words_max, c_max) : whash_row_bytes(words_max, c_max);
  const int R = G > kWave ? 1 : rows_per_block(G, rb, a.lds_budget);
  const size_t lds = (size_t)R * rb;
  const unsigned blocks = (unsigned)((a.nrows + R - 1) / R);
  set_dyn_lds((const void*)k_numeric_lds<G, DENSE>, lds);
  hipLaunchKernelGGL((k_numeric_lds<G, DENSE>), dim3(blocks), dim3(R * G), lds, s, a, words_max, c_max, rowoff, ccol,
                     cval);
}

static int64_t g_lds_per_cu = 160 * 1024;
void set_lds_per_cu(int64_t b) { g_lds_per_cu = b > 0 ? b : 65536; }
int64_t lds_per_cu() { return g_lds_per_cu; }

// Groups per block for a window kernel and the persistent grid.
static void window_shape(int G, int64_t group_bytes, int32_t nrows, int64_t lds_budget, int* R_out,
                         unsigned* blocks_out) {
  int64_t budget = g_lds_per_cu / 2;   // two blocks per CU
  if (budget > lds_budget * 2) budget = lds_budget * 2;
  int R = (int)(budget / (group_bytes > 0 ? group_bytes : 1));
  const int per_wave = kWave / G;
  if (R > 256 / G) R = 256 / G;
  if (R >= per_wave) R -= R % per_wave;
  if (R < 1) R = 1;
  const int64_t blocks_per_cu = std::max<int64_t>(1, g_lds_per_cu / ((int64_t)R * group_bytes));
  int64_t blocks = (nrows + R - 1) / R;
  const int64_t cap = (int64_t)g_num_cus * blocks_per_cu * 2;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  *R_out = R;
  *blocks_out = (unsigned)blocks;
}

constexpr int64_t kNarrowIdx = (int64_t)1 << 29;  // 32-bit byte offsets into B's f64 values

template <int G, bool UPPER, bool W64>
static void sym_window_launch(const RowArgs& a, int32_t W, int32_t* rowcnt, hipStream_t s) {
  const int64_t gb = sym_window_bytes(G, W);
  int R; unsigned blocks;
  window_shape(G, gb, a.nrows, a.lds_budget, &R, &blocks);
  const size_t lds = (size_t)R * gb;
  set_dyn_lds((const void*)k_sym_window<G, UPPER, W64>, lds);
  hipLaunchKernelGGL((k_sym_window<G, UPPER, W64>), dim3(blocks), dim3(R * G), lds, s, a, W, rowcnt);
}
template <int G, bool UPPER, bool W64, bool LIST>
static void num_window_launch(const RowArgs& a, int32_t W, int32_t Cm, const int64_t* rowoff, int32_t* ccol,
                              double* cval, hipStream_t s) {
  const int64_t gb = LIST ? num_list_bytes(G, W, Cm) : num_window_bytes(G, W, Cm);
  int R; unsigned blocks;
  window_shape(G, gb, a.nrows, a.lds_budget, &R, &blocks);
  const size_t lds = (size_t)R * gb;
  set_dyn_lds((const void*)k_num_window<G, UPPER, W64, LIST>, lds);
  hipLaunchKernelGGL((k_num_window<G, UPPER, W64, LIST>), dim3(blocks), dim3(R * G), lds, s, a, W, Cm,
                     rowoff, ccol, cval);
}
template <int G>
static void sym_window_any(const RowArgs& a, int32_t W, int32_t* rowcnt, hipStream_t s) {
  if (W <= kNarrowWords) {
    const int R = 256 / G;
    const unsigned blocks = (unsigned)((a.nrows + R - 1) / R);
    hipLaunchKernelGGL(k_sym_narrow<G>, dim3(blocks), dim3(256), 0, s, a, rowcnt);
    return;
  }
  const bool wide = a.bnnz >= kNarrowIdx;
  if (a.upper) { if (wide) sym_window_launch<G, true, true>(a, W, rowcnt, s); else sym_window_launch<G, true, false>(a, W, rowcnt, s); }
  else { if (wide) sym_window_launch<G, false, true>(a, W, rowcnt, s); else sym_window_launch<G, false, false>(a, W, rowcnt, s); }
}
template <int G, bool LIST>
static void num_window_any(const RowArgs& a, int32_t W, int32_t Cm, const int64_t* rowoff, int32_t* ccol,
                           double* cval, hipStream_t s) {
  const bool wide = a.bnnz >= kNarrowIdx;
  if (a.upper) {
    if (wide) num_window_launch<G, true, true, LIST>(a, W, Cm, rowoff, ccol, cval, s);
    else num_window_launch<G, true, false, LIST>(a, W, Cm, rowoff, ccol, cval, s);
  } else {
    if (wide) num_window_launch<G, false, true, LIST>(a, W, Cm, rowoff, ccol, cval, s);
    else num_window_launch<G, false, false, LIST>(a, W, Cm, rowoff, ccol, cval, s);
  }
}

static void heavy_launch(bool vals, const RowArgs& a, const HeavyArgs& h, int32_t* work, int32_t* rowcnt,
                         const int64_t* rowoff, int32_t* ccol, double* cval, hipStream_t s) {
  const size_t lds =
      (size_t)(vals ? heavy_lds_bytes(true, h.wc_num, h.vcap, h.dcols) : heavy_lds_bytes(false, h.wc_sym, 0, 0));
  // one block per CU (kHeavyT threads and most of the CU's LDS each)
  const unsigned blocks = (unsigned)(a.nrows < g_num_cus ? a.nrows : g_num_cus);
  launch_fill_i32(work, 1, 0, s);
  if (vals) {
    set_dyn_lds((const void*)k_heavy_num, lds);
    hipLaunchKernelGGL(k_heavy_num, dim3(blocks), dim3(kHeavyT), lds, s, a, h.wc_num, h.vcap, h.dcols,
                       h.adapt ? (int64_t)lds : (int64_t)0, work,
                       h.bflag, h.cur, rowoff, ccol, cval, h.fault);
  } else {
    set_dyn_lds((const void*)k_heavy_sym, lds);
    hipLaunchKernelGGL(k_heavy_sym, dim3(blocks), dim3(kHeavyT), lds, s, a, h.wc_sym, work, h.bflag, h.cur, rowcnt);
  }
}

// Heavy rows of a row panel taking their symbolic results from a row-pointer
// plan over all rows: panel row i is plan row r0 + i.  The plan's distinct-
// word count is kept; such rows never have a word list.
__global__ void k_take_plan(const int32_t* __restrict__ rows, int32_t nrows, int32_t r0,
                            const int32_t* __restrict__ pcnt, const int32_t* __restrict__ pwords,
                            int32_t* __restrict__ rowcnt, int32_t* __restrict__ rowwords) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x) {
    const int i = rows[r];
    rowcnt[i] = pcnt[r0 + i];
    rowwords[i] = pwords[r0 + i] | kNoWordList;
  }
}
void launch_take_plan(const RowArgs& a, int32_t r0, const int32_t* pcnt, const int32_t* pwords, int32_t* rowcnt,
                      hipStream_t s) {
  if (a.nrows <= 0) return;
  int64_t blocks = ((int64_t)a.nrows + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_take_plan, dim3((unsigned)blocks), dim3(256), 0, s, a.rows, a.nrows, r0, pcnt, pwords, rowcnt,
                     a.rowwords);
}

void launch_symbolic_bin(int bin, const RowArgs& a, int32_t words_max, int32_t* rowcnt, const HeavyArgs& h,
                         hipStream_t s) {
  if (a.nrows <= 0 || bin <= 0) return;
  if (bin == kBinHeavy) {
    heavy_launch(false, a, h, h.work, rowcnt, nullptr, nullptr, nullptr, s);
    return;
  }
  if (bin_is_sort(bin)) {
    switch (sort_bin_slots(bin)) {
      case 16: sym_sort_launch<16, 1>(a, words_max, rowcnt, s); break;
      case 32: sym_sort_launch<32, 1>(a, words_max, rowcnt, s); break;
      case 64: sym_sort_launch<64, 1>(a, words_max, rowcnt, s); break;
      case 128: sym_sort_launch<64, 2>(a, words_max, rowcnt, s); break;
      case 256: sym_sort_launch<64, 4>(a, words_max, rowcnt, s); break;
      case 512: sym_sort_launch<64, 8>(a, words_max, rowcnt, s); break;
      case 1024: sym_sort_launch<64, 16>(a, words_max, rowcnt, s); break;
      default: sym_sort_launch<64, 32>(a, words_max, rowcnt, s); break;
    }
    return;
  }
  if (bin == kBinWgBitmap) { sym_lds<256, true>(a, words_max, rowcnt, s); return; }
  if (bin == kBinWgHash) { sym_lds<256, false>(a, words_max, rowcnt, s); return; }
  const int G = bin_group(bin);
  const bool bm = bin_is_bitmap(bin);
  switch (G * 2 + (bm ? 1 : 0)) {
    case 17: sym_window_any<8>(a, words_max, rowcnt, s); break;
    case 16: sym_lds<8, false>(a, words_max, rowcnt, s); break;
    case 33: sym_window_any<16>(a, words_max, rowcnt, s); break;
    case 32: sym_lds<16, false>(a, words_max, rowcnt, s); break;
    case 65: sym_window_any<32>(a, words_max, rowcnt, s); break;
    case 64: sym_lds<32, false>(a, words_max, rowcnt, s); break;
    case 129: sym_window_any<64>(a, words_max, rowcnt, s); break;
    case 128: sym_lds<64, false>(a, words_max, rowcnt, s); break;
    default: break;
  }
}

void launch_numeric_bin(int bin, const RowArgs& a, int32_t words_max, int32_t c_max, const int64_t* rowoff,
                        int32_t* ccol, double* cval, const HeavyArgs& h, hipStream_t s) {
  if (a.nrows <= 0 || bin <= 0) return;
  if (bin == kBinHeavy) {
    heavy_launch(true, a, h, h.work + 1, nullptr, rowoff, ccol, cval, s);
    return;
  }
  if (bin_is_sort(bin)) {
    switch (sort_bin_slots(bin)) {
      case 16: num_sort_launch<16, 1>(a, words_max, rowoff, ccol, cval, s); break;
      case 32: num_sort_launch<32, 1>(a, words_max, rowoff, ccol, cval, s); break;
      case 64: num_sort_launch<64, 1>(a, words_max, rowoff, ccol, cval, s); break;
      case 128: num_sort_launch<64, 2>(a, words_max, rowoff, ccol, cval, s); break;
      case 256: num_sort_launch<64, 4>(a, words_max, rowoff, ccol, cval, s); break;
      case 512: num_sort_launch<64, 8>(a, words_max, rowoff, ccol, cval, s); break;
      case 1024: num_sort_launch<64, 16>(a, words_max, rowoff, ccol, cval, s); break;
      default: num_sort_launch<64, 32>(a, words_max, rowoff, ccol, cval, s); break;
    }
    return;
  }
  if (bin_is_list(bin)) {
    switch (bin_group(bin)) {
      case 8: num_window_any<8, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
      case 16: num_window_any<16, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
      case 32: num_window_any<32, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
      default: num_window_any<64, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    }
    return;
  }
  if (bin == kBinWgBitmap) { num_lds<256, true>(a, words_max, c_max, rowoff, ccol, cval, s); return; }
  if (bin == kBinWgHash) { num_lds<256, false>(a, words_max, c_max, rowoff, ccol, cval, s); return; }
  const int G = bin_group(bin);
  const bool bm = bin_is_bitmap(bin);
  switch (G * 2 + (bm ? 1 : 0)) {
    case 17: num_window_any<8, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 16: num_lds<8, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 33: num_window_any<16, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 32: num_lds<16, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 65: num_window_any<32, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 64: num_lds<32, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 129: num_window_any<64, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 128: num_lds<64, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// Digest of a CSR result: out[0] hashes the structure (each row pointer with
// its index, each column with its position), out[1] the value bits with their
// positions.  Each term is a splitmix64 mix of (position, content); terms are
// summed modulo 2^64, so the digest does not depend on the reduction order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
template <typename RP>
__global__ void __launch_bounds__(256) k_digest(int32_t n, int64_t nnz, const RP* __restrict__ rp,
                                                const int32_t* __restrict__ col, const double* __restrict__ val,
                                                unsigned long long* __restrict__ out) {
  unsigned long long hs = 0, hv = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= n; t += stride)
    hs += mix64(mix64((unsigned long long)t) ^ (unsigned long long)(int64_t)rp[t]);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnz; t += stride) {
    const unsigned long long pos = mix64((unsigned long long)t + 0x5bd1e995ull);
    hs += mix64(pos ^ (unsigned long long)(uint32_t)col[t]);
    hv += mix64(pos ^ (unsigned long long)__double_as_longlong(val[t]));
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    hs += __shfl_xor(hs, m, 64);
    hv += __shfl_xor(hv, m, 64);
  }
  if (threadIdx.x % kWave == 0) {
    atomicAdd(&out[0], hs);
    atomicAdd(&out[1], hv);
  }
}
void launch_digest(int32_t n, int64_t nnz, const void* rp, bool rp64, const int32_t* col, const double* val,
                   unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u64x2, dim3(1), dim3(64), 0, s, out);
  const int64_t work = std::max<int64_t>((int64_t)n + 1, nnz);
  int64_t blocks = (work + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (rp64)
    hipLaunchKernelGGL(k_digest<int64_t>, dim3((unsigned)blocks), dim3(256), 0, s, n, nnz,
                       static_cast<const int64_t*>(rp), col, val, out);
  else
    hipLaunchKernelGGL(k_digest<int32_t>, dim3((unsigned)blocks), dim3(256), 0, s, n, nnz,
                       static_cast<const int32_t*>(rp), col, val, out);
}

// ---------------------------------------------------------------------------
// Numeric semantics (scipy's eliminate_zeros): entries whose value is exactly
// zero (structural products that cancelled, and explicit zeros) are dropped.
// Count per row, scan, order-preserving compaction with 16-lane ballots.
// ---------------------------------------------------------------------------
template <typename RP>
__global__ void k_count_nonzero(int32_t n, const RP* __restrict__ rp, const double* __restrict__ val,
                                int32_t* __restrict__ cnt) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    int c = 0;
    for (int64_t p = (int64_t)rp[i] + lane; p < (int64_t)rp[i + 1]; p += G) c += val[p] != 0.0 ? 1 : 0;
    c = gsum<G, int32_t>(c, nullptr);
    if (lane == 0) cnt[i] = c;
  }
}
template <typename RP>
__global__ void k_compact_nonzero(int32_t n, const RP* __restrict__ rp, const int32_t* __restrict__ col,
                                  const double* __restrict__ val, const int64_t* __restrict__ off,
                                  int32_t* __restrict__ ocol, double* __restrict__ oval) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    const int64_t s = rp[i], e = rp[i + 1];
    int64_t o = off[i];
    for (int64_t b = s; b < e; b += G) {
      const int64_t p = b + lane;
      const bool keep = p < e && val[p] != 0.0;
      const uint64_t m = group_bits<G>(__ballot(keep));
      if (keep) {
        const int64_t q = o + __popcll(m & ((1ull << lane) - 1ull));
        ocol[q] = col[p];
        oval[q] = val[p];
      }
      o += __popcll(m);
    }
  }
}
void launch_count_nonzero(int32_t n, const void* rp, bool rp64, const double* val, int32_t* cnt, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (rp64)
    hipLaunchKernelGGL(k_count_nonzero<int64_t>, dim3((unsigned)blocks), dim3(256), 0, s, n, (const int64_t*)rp, val, cnt);
  else
    hipLaunchKernelGGL(k_count_nonzero<int32_t>, dim3((unsigned)blocks), dim3(256), 0, s, n, (const int32_t*)rp, val, cnt);
}
void launch_compact_nonzero(int32_t n, const void* rp, bool rp64, const int32_t* col, const double* val,
                            const int64_t* off, int32_t* ocol, double* oval, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (rp64)
    hipLaunchKernelGGL(k_compact_nonzero<int64_t>, dim3((unsigned)blocks), dim3(256), 0, s, n, (const int64_t*)rp, col,
                       val, off, ocol, oval);
  else
    hipLaunchKernelGGL(k_compact_nonzero<int32_t>, dim3((unsigned)blocks), dim3(256), 0, s, n, (const int32_t*)rp, col,
                       val, off, ocol, oval);
}

// ---------------------------------------------------------------------------
// AAtS mirror.  C's structure is symmetric and every row is already complete
// and sorted; the numeric pass formed values only for j >= i.  Each lower
// entry (c, r), r < c, takes the value of (r, c) from row r's upper part,
// found by binary search (columns are ascending).  No atomics, no sort.
// Short rows are dealt by rows; long rows (power-law hubs, ~10^4-10^5 lower
// entries) by entries.
// ---------------------------------------------------------------------------
constexpr int kMirrorChunk = 4096;  // C entries per block of the long-row kernel
constexpr int kMirrorLong = 1024;   // rows longer than this go to the entry-parallel kernel

// Value of (r, c) from row r (columns ascending), or 0 when absent.
__device__ __forceinline__ double mirror_value(const int64_t* off, const int32_t* ccol, const double* cval, int r,
                                               int c) {
  int64_t lo = off[r], hi = off[r + 1];
  const int64_t end = hi;
  while (lo < hi) {  // first position of row r with column >= c
    const int64_t mid = (lo + hi) >> 1;
    if (ccol[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && ccol[lo] == c) ? cval[lo] : 0.0;
}

// Rows of at most kMirrorLong entries: 16 lanes per row over its lower part.
__global__ void k_mirror_rows(int32_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ ccol,
                              double* __restrict__ cval) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t c = (int64_t)blockIdx.x * gpb + threadIdx.x / G; c < n; c += (int64_t)gridDim.x * gpb) {
    const int64_t s = off[c], e = off[c + 1];
    if (e - s > kMirrorLong) continue;
    for (int64_t b = s; b < e; b += G) {
      if (ccol[b] >= c) break;  // the lower part is a prefix of the row
      const int64_t p = b + lane;
      if (p < e) {
        const int r = ccol[p];
        if (r < c) cval[p] = mirror_value(off, ccol, cval, r, (int)c);
      }
    }
  }
}

// Longer rows, entry-parallel: a block owns kMirrorChunk consecutive entries
// of C, so one long row is spread over many blocks.
__global__ void __launch_bounds__(256) k_mirror_long(int32_t n, int64_t nnz, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ ccol, double* __restrict__ cval) {
  __shared__ int32_t rows[2];
  const int64_t e0 = (int64_t)blockIdx.x * kMirrorChunk;
  const int64_t e1 = (e0 + kMirrorChunk) < nnz ? (e0 + kMirrorChunk) : nnz;
  if (threadIdx.x < 2) {
    // last row whose first entry is at or before e0 (thread 0) / e1 - 1 (thread 1)
    const int64_t e = threadIdx.x == 0 ? e0 : e1 - 1;
    int l = 0, h = n - 1;
    while (l < h) {
      const int mid = (int)(((int64_t)l + h + 1) >> 1);
      if (off[mid] <= e) l = mid;
      else h = mid - 1;
    }
    rows[threadIdx.x] = l;
  }
  __syncthreads();
  const int r0 = rows[0], r1 = rows[1];
  // skip the block when none of its rows is long (checked in one pass
  // when it spans few rows)
  if (r1 - r0 < (int)blockDim.x) {
    const int c = r0 + (int)threadIdx.x;
    const bool lng = c <= r1 && off[c + 1] - off[c] > kMirrorLong;
    if (!__syncthreads_or(lng)) return;
  }
  int c = r0;
  for (int64_t p = e0 + threadIdx.x; p < e1; p += blockDim.x) {
    while (off[c + 1] <= p) c++;  // the row holding entry p
    if (off[c + 1] - off[c] <= kMirrorLong) continue;
    const int r = ccol[p];
    if (r >= c) continue;         // upper part: formed by the numeric pass
    cval[p] = mirror_value(off, ccol, cval, r, c);
  }
}

// Radix mirror.  Strictly-upper entries (j, i), i > j, taken in row-major
// order and stably sorted by column i, come out grouped by i with j
// ascending: exactly row i's lower part, which starts at off[i].
// lower[i] / upper[i]: row i's entries left / right of the diagonal.
__global__ void k_mirror_counts(int32_t n, const int64_t* __restrict__ off, const int32_t* __restrict__ ccol,
                                int32_t* __restrict__ lower, int32_t* __restrict__ upper) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = off[i], hi = off[i + 1];
    const int64_t s = lo, e = hi;
    while (lo < hi) {  // first column >= i
      const int64_t mid = (lo + hi) >> 1;
      if (ccol[mid] < (int32_t)i) lo = mid + 1;
      else hi = mid;
    }
    const int diag = (lo < e && ccol[lo] == (int32_t)i) ? 1 : 0;
    lower[i] = (int32_t)(lo - s);
    upper[i] = (int32_t)(e - lo - diag);
  }
}
// The strictly-upper entries, compacted in row-major order: entry-parallel
// over the compacted positions (a block per chunk of them), so that long
// rows spread over many blocks.  Position q belongs to the last row j with
// up_off[j] <= q; a block finds the rows of its chunk's ends, and each
// thread searches between them.
constexpr int64_t kMirrorExtractChunk = 4096;
__device__ __forceinline__ int32_t last_row_at_most(const int64_t* __restrict__ pre, int32_t lo, int32_t hi,
                                                    int64_t q) {
  // largest j in [lo, hi] with pre[j] <= q (pre[lo] <= q holds)
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo + 1) / 2;
    if (pre[mid] <= q) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__global__ void __launch_bounds__(256) k_mirror_extract(int32_t n, int64_t count, const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ ccol,
                                                        const double* __restrict__ cval,
                                                        const int32_t* __restrict__ upper,
                                                        const int64_t* __restrict__ up_off,
                                                        int32_t* __restrict__ keys, double* __restrict__ vals) {
  __shared__ int32_t ends[2];
  const int64_t c0 = (int64_t)blockIdx.x * kMirrorExtractChunk;
  const int64_t c1 = c0 + kMirrorExtractChunk < count ? c0 + kMirrorExtractChunk : count;
  if (threadIdx.x < 2) ends[threadIdx.x] = last_row_at_most(up_off, 0, n - 1, threadIdx.x == 0 ? c0 : c1 - 1);
  __syncthreads();
  const int32_t rlo = ends[0], rhi = ends[1];
  for (int64_t q = c0 + threadIdx.x; q < c1; q += blockDim.x) {
    const int32_t j = last_row_at_most(up_off, rlo, rhi, q);
    const int64_t src = off[j + 1] - upper[j] + (q - up_off[j]);
    keys[q] = ccol[src];
    vals[q] = cval[src];
  }
}
__global__ void k_mirror_place(int64_t count, const int32_t* __restrict__ keys, const double* __restrict__ vals,
                               const int64_t* __restrict__ lo_off, const int64_t* __restrict__ off,
                               double* __restrict__ cval) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = keys[q];
    cval[off[i] + (q - lo_off[i])] = vals[q];
  }
}
void launch_mirror_counts(int32_t n, const int64_t* off, const int32_t* ccol, int32_t* lower, int32_t* upper,
                          hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_mirror_counts, dim3((unsigned)blocks), dim3(256), 0, s, n, off, ccol, lower, upper);
}
void launch_mirror_radix(int32_t n, int64_t count, const int64_t* off, const int32_t* ccol, double* cval,
                         const int32_t* upper, const int64_t* up_off, const int64_t* lo_off, const RadixScratch& rx,
                         double* vals2, hipStream_t s) {
  if (n <= 0 || count <= 0) return;
  // extract into the arrays the first pass does not write
  const bool odd = radix_passes(n) % 2 == 1;
  int32_t* ek = odd ? rx.keys : rx.keys2;
  double* ev = odd ? rx.vals : vals2;
  const int64_t blocks = (count + kMirrorExtractChunk - 1) / kMirrorExtractChunk;
  hipLaunchKernelGGL(k_mirror_extract, dim3((unsigned)blocks), dim3(256), 0, s, n, count, off, ccol,
                     (const double*)cval, upper, up_off, ek, ev);
  radix_passes_run(n, count, ek, nullptr, ev, rx.keys, nullptr, rx.vals, rx.keys2, nullptr, vals2, rx.hist, rx.hoff,
                   rx.htmp, nullptr, s);
  int64_t b2 = (count + 255) / 256;
  if (b2 > 16384) b2 = 16384;
  hipLaunchKernelGGL(k_mirror_place, dim3((unsigned)b2), dim3(256), 0, s, count, (const int32_t*)rx.keys2,
                     (const double*)vals2, lo_off, off, cval);
}

void launch_mirror_pull(int32_t n, int64_t nnz, const int64_t* off, const int32_t* ccol, double* cval,
                        hipStream_t s) {
  if (n <= 0 || nnz <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_mirror_rows, dim3((unsigned)blocks), dim3(256), 0, s, n, off, ccol, cval);
  const int64_t b2 = (nnz + kMirrorChunk - 1) / kMirrorChunk;
  hipLaunchKernelGGL(k_mirror_long, dim3((unsigned)b2), dim3(256), 0, s, n, nnz, off, ccol, cval);
}

}  // namespace appt
