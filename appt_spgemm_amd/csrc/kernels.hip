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
#include "kernels.hpp"
#include "device_common.hpp"

namespace appt {

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

__global__ void k_i64_to_i32(const int64_t* in, int64_t n, int32_t* out) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (int32_t)in[t];
}
void launch_i64_to_i32(const int64_t* in, int64_t n, int32_t* out, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_i64_to_i32, dim3((unsigned)blocks), dim3(256), 0, s, in, n, out);
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
// per-row min / max column (span of a B row); 16 lanes per row
// ---------------------------------------------------------------------------
__global__ void k_rowstats(int32_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                           int32_t* __restrict__ rmin, int32_t* __restrict__ rmax) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    const int s = ptr[i], e = ptr[i + 1];
    int mn = INT32_MAX, mx = -1;
    for (int p = s + lane; p < e; p += G) {
      const int c = col[p];
      mn = c < mn ? c : mn;
      mx = c > mx ? c : mx;
    }
    mn = gmin<G>(mn);
    mx = gmax<G>(mx);
    if (lane == 0) { rmin[i] = mn; rmax[i] = mx; }
  }
}
void launch_rowstats(int32_t n, const int32_t* ptr, const int32_t* col, int32_t* rmin, int32_t* rmax,
                     hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_rowstats, dim3((unsigned)blocks), dim3(256), 0, s, n, ptr, col, rmin, rmax);
}

// ---------------------------------------------------------------------------
// transpose B = A^T: column histogram, scan, atomic-cursor scatter.  The order
// inside a B row is not fixed; it does not affect any output value because a
// B row holds each column once (see the value-order note at the top).
// ---------------------------------------------------------------------------
__global__ void k_colcount(int64_t nnz, const int32_t* __restrict__ col, int32_t* __restrict__ cnt) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz;
       p += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[col[p]], 1);
}
void launch_colcount(int64_t nnz, const int32_t* col, int32_t* cnt, hipStream_t s) {
  if (nnz <= 0) return;
  int64_t blocks = (nnz + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_colcount, dim3((unsigned)blocks), dim3(256), 0, s, nnz, col, cnt);
}

__global__ void k_transpose_scatter(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                                    const double* __restrict__ av, int32_t* __restrict__ cursor,
                                    int32_t* __restrict__ bc, double* __restrict__ bv) {
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
                              int32_t* cursor, int32_t* bc, double* bv, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_transpose_scatter, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, ac, av, cursor,
                     bc, bv);
}

// ---------------------------------------------------------------------------
// exclusive scan int32 -> int64 (reduce / scan partials / apply)
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 2; }

__global__ void k_scan_reduce(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ part) {
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

__global__ void k_scan_partials(int64_t* part, int64_t nb) {
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
                             int64_t nb, int64_t* __restrict__ out64, int32_t* __restrict__ out32) {
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

void launch_scan(const int32_t* in, int64_t n, int64_t* out64, int32_t* out32, int64_t* tmp,
                 hipStream_t s) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 0) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int32_t*)out64, (int64_t)2, 0);
    if (out32) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, out32, (int64_t)1, 0);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, tmp, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, tmp, nb, out64, out32);
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bin_of(const BinParams& P, int64_t metric, int64_t words, int64_t L) {
  if (metric <= 0 || words <= 0) return 0;
  const int gidx = L <= P.glen_max[0] ? 0 : L <= P.glen_max[1] ? 1 : L <= P.glen_max[2] ? 2 : 3;
  for (int c = 0; c < P.ncls; c++)
    if (metric <= P.cls_max[c]) {
      const int mode = words <= P.cls_bm_words[c] ? 0 : 1;
      return 1 + (c * 2 + mode) * 4 + gidx;
    }
  if (metric <= P.wg_max) return words <= P.wg_bm_words ? kBinWgBitmap : kBinWgHash;
  return words <= P.gl_bm_words ? kBinGlBitmap : kBinGlHash;
}

__device__ __forceinline__ int64_t span_words(int lo, int hi) {
  return hi < lo ? 0 : (int64_t)((hi - lo) >> 5) + 1;
}

// Block-local bin statistics flushed with one global atomic per bin.
struct BinStatLds {
  int32_t v[kNumBins * kStatFields];
};
__device__ __forceinline__ void stats_init(BinStatLds& st) {
  for (int t = threadIdx.x; t < kNumBins * kStatFields; t += blockDim.x) st.v[t] = 0;
  __syncthreads();
}
__device__ __forceinline__ void stats_add(BinStatLds& st, int bin, int64_t metric, int64_t words) {
  atomicAdd(&st.v[bin * kStatFields + 0], 1);
  atomicMax(&st.v[bin * kStatFields + 1], (int32_t)(metric > INT32_MAX ? INT32_MAX : metric));
  atomicMax(&st.v[bin * kStatFields + 2], (int32_t)(words > INT32_MAX ? INT32_MAX : words));
}
__device__ __forceinline__ void stats_flush(BinStatLds& st, int32_t* g) {
  __syncthreads();
  for (int b = threadIdx.x; b < kNumBins; b += blockDim.x) {
    const int c = st.v[b * kStatFields + 0];
    if (c) {
      atomicAdd(&g[b * kStatFields + 0], c);
      atomicMax(&g[b * kStatFields + 1], st.v[b * kStatFields + 1]);
      atomicMax(&g[b * kStatFields + 2], st.v[b * kStatFields + 2]);
    }
  }
}

// Structural pre-pass: products f_i (k_flops of the reference), output span
// [lo, hi] from the B rows' min/max columns, symbolic bin, bin statistics.
__global__ void k_prepass(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                          const int32_t* __restrict__ bp, const int32_t* __restrict__ bmin,
                          const int32_t* __restrict__ bmax, int upper, BinParams P,
                          int64_t* __restrict__ rowf, int32_t* __restrict__ rowlo, int32_t* __restrict__ rowhi,
                          uint8_t* __restrict__ bin, int32_t* __restrict__ gstats,
                          unsigned long long* __restrict__ total_products) {
  constexpr int G = 16;
  __shared__ BinStatLds st;
  __shared__ unsigned long long red[16];
  stats_init(st);
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  unsigned long long myprod = 0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    const int s = ap[i], e = ap[i + 1];
    int64_t f = 0;
    int lo = INT32_MAX, hi = -1;
    for (int p = s + lane; p < e; p += G) {
      const int k = ac[p];
      const int len = bp[k + 1] - bp[k];
      f += len;
      if (len > 0) {
        const int a = bmin[k], b = bmax[k];
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
      }
    }
    f = gsum<G, int64_t>(f, nullptr);
    lo = gmin<G>(lo);
    hi = gmax<G>(hi);
    if (lane == 0) {
      if (upper && lo < (int)i) lo = (int)i;
      const int64_t words = span_words(lo, hi);
      const int64_t nA = e - s;
      const int64_t L = nA > 0 ? (f + nA - 1) / nA : 0;
      const int b = bin_of(P, f, words, L);
      rowf[i] = f;
      rowlo[i] = lo;
      rowhi[i] = hi;
      bin[i] = (uint8_t)b;
      stats_add(st, b, f, words);
      myprod += (unsigned long long)f;
    }
  }
  // block reduction of the product count
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) myprod += __shfl_xor(myprod, m, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = myprod;
  stats_flush(st, gstats);
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)blockDim.x / 64; w++) t += red[w];
    if (t) atomicAdd(total_products, t);
  }
}
void launch_prepass(int32_t n, const int32_t* ap, const int32_t* ac, const int32_t* bp, const int32_t* bmin,
                    const int32_t* bmax, int upper, const BinParams& P, int64_t* rowf, int32_t* rowlo,
                    int32_t* rowhi, uint8_t* bin, int32_t* stats, unsigned long long* total_products,
                    hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_prepass, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, ac, bp, bmin, bmax, upper, P,
                     rowf, rowlo, rowhi, bin, stats, total_products);
}

// Numeric bin of each row from its exact output count.
__global__ void k_numbin(int32_t n, const int32_t* __restrict__ ap, const int64_t* __restrict__ rowf,
                         const int32_t* __restrict__ rowcnt, const int32_t* __restrict__ rowlo,
                         const int32_t* __restrict__ rowhi, BinParams P, uint8_t* __restrict__ bin,
                         int32_t* __restrict__ gstats) {
  __shared__ BinStatLds st;
  stats_init(st);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = rowcnt[i];
    const int64_t words = span_words(rowlo[i], rowhi[i]);
    const int64_t nA = ap[i + 1] - ap[i];
    const int64_t f = rowf[i];
    const int64_t L = nA > 0 ? (f + nA - 1) / nA : 0;
    const int b = bin_of(P, c, words, L);
    bin[i] = (uint8_t)b;
    stats_add(st, b, c, words);
  }
  stats_flush(st, gstats);
}
void launch_numbin(int32_t n, const int32_t* ap, const int64_t* rowf, const int32_t* rowcnt,
                   const int32_t* rowlo, const int32_t* rowhi, const BinParams& P, uint8_t* bin,
                   int32_t* stats, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_numbin, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, rowf, rowcnt, rowlo, rowhi, P,
                     bin, stats);
}

// Partition rows by bin: one global atomic per (block, bin); rows keep their
// relative order within a block tile only approximately (LDS atomics), which
// affects scheduling, never results.
__global__ void k_bin_scatter(int32_t n, const uint8_t* __restrict__ bin, int32_t* __restrict__ cursor,
                              int32_t* __restrict__ rows) {
  constexpr int kItems = 4;
  __shared__ int32_t cnt[kNumBins];
  __shared__ int32_t base[kNumBins];
  for (int64_t tile = (int64_t)blockIdx.x * blockDim.x * kItems; tile < n;
       tile += (int64_t)gridDim.x * blockDim.x * kItems) {
    for (int t = threadIdx.x; t < kNumBins; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    int myb[kItems], myr[kItems];
#pragma unroll
    for (int t = 0; t < kItems; t++) {
      const int64_t i = tile + (int64_t)t * blockDim.x + threadIdx.x;
      myb[t] = i < n ? (int)bin[i] : -1;
      myr[t] = (myb[t] > 0) ? atomicAdd(&cnt[myb[t]], 1) : 0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kNumBins; t += blockDim.x)
      base[t] = (t > 0 && cnt[t]) ? atomicAdd(&cursor[t], cnt[t]) : 0;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kItems; t++) {
      const int64_t i = tile + (int64_t)t * blockDim.x + threadIdx.x;
      if (myb[t] > 0) rows[base[myb[t]] + myr[t]] = (int32_t)i;
    }
    __syncthreads();
  }
}
void launch_bin_scatter(int32_t n, const uint8_t* bin, int32_t* cursor, int32_t* rows, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n + 1023) / 1024;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)blocks), dim3(256), 0, s, n, bin, cursor, rows);
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
};

template <int G, bool VALS, typename Fn>
__device__ __forceinline__ void for_each_product(int i, int lane, const RowArgs& A, Stage* stage, Fn&& fn) {
  const int s = A.ap[i], e = A.ap[i + 1];
  for (int base = s; base < e; base += G) {
    const int p = base + lane;
    int bs = 0, be = 0;
    double a = 0.0;
    if (p < e) {
      const int k = A.ac[p];
      bs = A.bp[k];
      be = A.bp[k + 1];
      if (VALS) a = A.av[p];
    }
    const int m = (e - base) < G ? (e - base) : G;
    if constexpr (G <= kWave) {
      for (int t = 0; t < m; t++) {
        const int tbs = __shfl(bs, t, G), tbe = __shfl(be, t, G);
        const double ta = VALS ? __shfl(a, t, G) : 0.0;
        for (int q = tbs + lane; q < tbe; q += G) fn(A.bc[q], q, ta);
      }
    } else {
      __syncthreads();
      stage->bs[lane] = bs;
      stage->be[lane] = be;
      if (VALS) stage->a[lane] = a;
      __syncthreads();
      for (int t = 0; t < m; t++) {
        const int tbs = stage->bs[t], tbe = stage->be[t];
        const double ta = VALS ? stage->a[t] : 0.0;
        for (int q = tbs + lane; q < tbe; q += G) fn(A.bc[q], q, ta);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// symbolic: distinct output columns per row
// ---------------------------------------------------------------------------
template <int G, bool BITMAP>
__global__ void __launch_bounds__(1024) k_symbolic_lds(RowArgs A, int32_t row_words, int32_t* __restrict__ rowcnt) {
  extern __shared__ uint32_t lds[];
  __shared__ Stage stage_buf[G > kWave ? 1 : 0 + 1];
  __shared__ int32_t red[16];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = blockIdx.x * R + gi;
  if (G <= kWave && r >= A.nrows) return;
  const int i = A.rows[r];
  uint32_t* tab = lds + (size_t)gi * row_words;
  int cnt = 0;
  if constexpr (BITMAP) {
    const int lo = A.rowlo[i], hi = A.rowhi[i];
    const int words = (int)span_words(lo, hi);
    for (int t = lane; t < words; t += G) tab[t] = 0u;
    gsync<G>();
    for_each_product<G, false>(i, lane, A, stage_buf, [&](int j, int, double) {
      if (A.upper && j < i) return;
      const uint32_t b = (uint32_t)(j - lo);
      const uint32_t bit = 1u << (b & 31u);
      const uint32_t old = atomicOr(&tab[b >> 5], bit);
      cnt += (old & bit) ? 0 : 1;
    });
  } else {
    const int64_t T = np2_i64((A.rowf[i] * 4) / 3 + 1);
    const int logT = __builtin_ctzll((unsigned long long)T);
    const uint32_t mask = (uint32_t)T - 1u;
    int32_t* keys = (int32_t*)tab;
    for (int t = lane; t < (int)T; t += G) keys[t] = -1;
    gsync<G>();
    for_each_product<G, false>(i, lane, A, stage_buf, [&](int j, int, double) {
      if (A.upper && j < i) return;
      uint32_t h = hash_slot(j, logT);
      while (true) {
        const int cur = __atomic_load_n(&keys[h], __ATOMIC_RELAXED);
        if (cur == j) break;
        if (cur == -1) {
          const int old = atomicCAS(&keys[h], -1, j);
          if (old == -1) { cnt++; break; }
          if (old == j) break;
        }
        h = (h + 1u) & mask;
      }
    });
  }
  cnt = gsum<G, int32_t>(cnt, red);
  if (lane == 0) rowcnt[i] = cnt;
}

// Heavy rows: one workgroup per row, bitmap window in a global scratch slot.
__global__ void __launch_bounds__(256) k_symbolic_global(RowArgs A, int32_t slot_words, uint32_t* __restrict__ gscratch,
                                                         int32_t* __restrict__ rowcnt) {
  constexpr int G = 256;
  __shared__ Stage stage_buf[1];
  __shared__ int32_t red[16];
  uint32_t* tab = gscratch + (size_t)blockIdx.x * slot_words;
  const int lane = threadIdx.x;
  for (int r = blockIdx.x; r < A.nrows; r += gridDim.x) {
    const int i = A.rows[r];
    const int lo = A.rowlo[i], hi = A.rowhi[i];
    const int words = (int)span_words(lo, hi);
    __syncthreads();
    for (int t = lane; t < words; t += G) tab[t] = 0u;
    __threadfence();
    __syncthreads();
    int cnt = 0;
    for_each_product<G, false>(i, lane, A, stage_buf, [&](int j, int, double) {
      if (A.upper && j < i) return;
      const uint32_t b = (uint32_t)(j - lo);
      const uint32_t bit = 1u << (b & 31u);
      const uint32_t old = atomicOr(&tab[b >> 5], bit);
      cnt += (old & bit) ? 0 : 1;
    });
    cnt = gsum<G, int32_t>(cnt, red);
    if (lane == 0) rowcnt[i] = cnt;
  }
}

// ---------------------------------------------------------------------------
// numeric
// ---------------------------------------------------------------------------

// In-LDS bitonic sort (all comparators ascending; positions >= len act as
// +infinity and are never touched).
template <int G>
__device__ __forceinline__ void lds_sort_pairs(int32_t* keys, double* vals, int len, int lane) {
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
          const double v1 = vals[i1]; vals[i1] = vals[i2]; vals[i2] = v1;
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
            const double v1 = vals[i1]; vals[i1] = vals[i2]; vals[i2] = v1;
          }
        }
      }
      gsync<G>();
    }
  }
}

// Bitmap-window numeric.  LDS per row: info[words] (uint2: bits, prefix),
// vals[c_max] (f64), cols[c_max] (i32).  Output positions come from prefix
// popcounts, so the row is produced already sorted.
template <int G>
__device__ __forceinline__ void numeric_bitmap_row(int i, int lane, const RowArgs& A, uint2* info, double* vals,
                                                   int32_t* cols, const int64_t* rowoff, int32_t* ccol,
                                                   double* cval, Stage* stage, int64_t* red64) {
  const int lo = A.rowlo[i], hi = A.rowhi[i];
  const int words = (int)span_words(lo, hi);
  const int64_t off = rowoff[i];
  const int c = (int)(rowoff[i + 1] - off);
  for (int t = lane; t < words; t += G) info[t] = make_uint2(0u, 0u);
  for (int t = lane; t < c; t += G) vals[t] = 0.0;
  gsync<G>();
  uint32_t* bits = (uint32_t*)info;
  for_each_product<G, false>(i, lane, A, stage, [&](int j, int, double) {
    if (A.upper && j < i) return;
    const uint32_t b = (uint32_t)(j - lo);
    atomicOr(&bits[(b >> 5) * 2], 1u << (b & 31u));
  });
  gsync<G>();
  // prefix popcount over words: contiguous chunk per lane
  const int chunk = (words + G - 1) / G;
  const int w0 = lane * chunk;
  const int w1 = (w0 + chunk) < words ? (w0 + chunk) : words;
  int local = 0;
  for (int w = w0; w < w1; w++) local += __popc(info[w].x);
  int run;
  if constexpr (G <= kWave) {
    int tot;
    run = gscan_excl<G, int>(local, lane, &tot);
  } else {
    int64_t tot;
    run = (int)block_scan_excl<int64_t>(local, red64, &tot);
  }
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
  for_each_product<G, true>(i, lane, A, stage, [&](int j, int q, double a) {
    if (A.upper && j < i) return;
    const uint32_t b = (uint32_t)(j - lo);
    const uint2 wi = info[b >> 5];
    const uint32_t pos = wi.y + __popc(wi.x & ((1u << (b & 31u)) - 1u));
    atomicAdd(&vals[pos], a * A.bv[q]);
  });
  gsync<G>();
  for (int t = lane; t < c; t += G) {
    ccol[off + t] = cols[t];
    cval[off + t] = vals[t];
  }
}

// Hash numeric: keys[T] (i32), vals[T] (f64), T = np2(3c/2+1); compact in
// place, bitonic sort, write.
template <int G>
__device__ __forceinline__ void numeric_hash_row(int i, int lane, const RowArgs& A, int32_t* keys, double* vals,
                                                 const int64_t* rowoff, int32_t* ccol, double* cval, Stage* stage,
                                                 int64_t* red64) {
  const int64_t off = rowoff[i];
  const int c = (int)(rowoff[i + 1] - off);
  const int64_t T = np2_i64((int64_t)(c * 3) / 2 + 1);
  const int logT = __builtin_ctzll((unsigned long long)T);
  const uint32_t mask = (uint32_t)T - 1u;
  for (int t = lane; t < (int)T; t += G) { keys[t] = -1; vals[t] = 0.0; }
  gsync<G>();
  for_each_product<G, true>(i, lane, A, stage, [&](int j, int q, double a) {
    if (A.upper && j < i) return;
    uint32_t h = hash_slot(j, logT);
    while (true) {
      const int cur = __atomic_load_n(&keys[h], __ATOMIC_RELAXED);
      if (cur == j) break;
      if (cur == -1) {
        const int old = atomicCAS(&keys[h], -1, j);
        if (old == -1 || old == j) break;
      }
      h = (h + 1u) & mask;
    }
    atomicAdd(&vals[h], a * A.bv[q]);
  });
  gsync<G>();
  // in-place compaction (destination index never exceeds source index)
  int count = 0;
  for (int base = 0; base < (int)T; base += G) {
    const int t = base + lane;
    const int32_t key = t < (int)T ? keys[t] : -1;
    const double v = t < (int)T ? vals[t] : 0.0;
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
    gsync<G>();
    if (occ) { keys[count + rank] = key; vals[count + rank] = v; }
    count += tot;
    gsync<G>();
  }
  lds_sort_pairs<G>(keys, vals, c, lane);
  for (int t = lane; t < c; t += G) {
    ccol[off + t] = keys[t];
    cval[off + t] = vals[t];
  }
}

template <int G, bool BITMAP>
__global__ void __launch_bounds__(1024) k_numeric_lds(RowArgs A, int32_t words_max, int32_t c_max,
                                                      const int64_t* __restrict__ rowoff, int32_t* __restrict__ ccol,
                                                      double* __restrict__ cval) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ Stage stage_buf[G > kWave ? 1 : 0 + 1];
  __shared__ int64_t red64[17];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = blockIdx.x * R + gi;
  if (G <= kWave && r >= A.nrows) return;
  const int i = A.rows[r];
  if constexpr (BITMAP) {
    const size_t row_bytes = (size_t)words_max * 8 + (size_t)c_max * 12;
    const size_t rb = (row_bytes + 15) & ~(size_t)15;
    unsigned char* base = lds_raw + (size_t)gi * rb;
    uint2* info = (uint2*)base;
    double* vals = (double*)(base + (size_t)words_max * 8);
    int32_t* cols = (int32_t*)(base + (size_t)words_max * 8 + (size_t)c_max * 8);
    numeric_bitmap_row<G>(i, lane, A, info, vals, cols, rowoff, ccol, cval, stage_buf, red64);
  } else {
    const int64_t T = np2_i64((int64_t)(c_max * 3) / 2 + 1);
    const size_t rb = ((size_t)T * 12 + 15) & ~(size_t)15;
    unsigned char* base = lds_raw + (size_t)gi * rb;
    double* vals = (double*)base;
    int32_t* keys = (int32_t*)(base + (size_t)T * 8);
    numeric_hash_row<G>(i, lane, A, keys, vals, rowoff, ccol, cval, stage_buf, red64);
  }
}

// Heavy rows: global bitmap slot (bits + prefix as uint2 per word); values are
// accumulated straight into C at their final sorted position.
__global__ void __launch_bounds__(256) k_numeric_global(RowArgs A, int32_t slot_words, uint32_t* __restrict__ gscratch,
                                                        const int64_t* __restrict__ rowoff, int32_t* __restrict__ ccol,
                                                        double* __restrict__ cval) {
  constexpr int G = 256;
  __shared__ Stage stage_buf[1];
  __shared__ int64_t red64[17];
  uint2* info = (uint2*)(gscratch + (size_t)blockIdx.x * slot_words * 2);
  uint32_t* bits = (uint32_t*)info;
  const int lane = threadIdx.x;
  for (int r = blockIdx.x; r < A.nrows; r += gridDim.x) {
    const int i = A.rows[r];
    const int lo = A.rowlo[i], hi = A.rowhi[i];
    const int words = (int)span_words(lo, hi);
    const int64_t off = rowoff[i];
    const int c = (int)(rowoff[i + 1] - off);
    __syncthreads();
    for (int t = lane; t < words; t += G) info[t] = make_uint2(0u, 0u);
    for (int t = lane; t < c; t += G) cval[off + t] = 0.0;
    __threadfence();
    __syncthreads();
    for_each_product<G, false>(i, lane, A, stage_buf, [&](int j, int, double) {
      if (A.upper && j < i) return;
      const uint32_t b = (uint32_t)(j - lo);
      atomicOr(&bits[(b >> 5) * 2], 1u << (b & 31u));
    });
    __threadfence();
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int chunk = (words + G - 1) / G;
    const int w0 = lane * chunk;
    const int w1 = (w0 + chunk) < words ? (w0 + chunk) : words;
    int local = 0;
    for (int w = w0; w < w1; w++) local += __popc(__hip_atomic_load(&bits[w * 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    int64_t tot;
    int run = (int)block_scan_excl<int64_t>(local, red64, &tot);
    for (int w = w0; w < w1; w++) {
      uint32_t x = __hip_atomic_load(&bits[w * 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      info[w].y = (uint32_t)run;
      while (x) {
        const int bpos = __builtin_ctz(x);
        ccol[off + run++] = lo + w * 32 + bpos;
        x &= x - 1u;
      }
    }
    __threadfence();
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for_each_product<G, true>(i, lane, A, stage_buf, [&](int j, int q, double a) {
      if (A.upper && j < i) return;
      const uint32_t b = (uint32_t)(j - lo);
      const uint2 wi = info[b >> 5];
      const uint32_t pos = wi.y + __popc(wi.x & ((1u << (b & 31u)) - 1u));
      unsafeAtomicAdd(&cval[off + pos], a * A.bv[q]);
    });
  }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
int64_t symbolic_row_lds_bytes(int bin, int32_t metric_max, int32_t words_max) {
  if (bin == kBinGlBitmap || bin == kBinGlHash || bin == 0) return 0;
  const bool bm = (bin == kBinWgBitmap) || (bin < kBinWgBitmap && bin_is_bitmap(bin));
  return bm ? (int64_t)words_max * 4 : sym_slots(metric_max) * 4;
}
int64_t numeric_row_lds_bytes(int bin, int32_t c_max, int32_t words_max) {
  if (bin == kBinGlBitmap || bin == kBinGlHash || bin == 0) return 0;
  const bool bm = (bin == kBinWgBitmap) || (bin < kBinWgBitmap && bin_is_bitmap(bin));
  if (bm) return (((int64_t)words_max * 8 + (int64_t)c_max * 12) + 15) & ~(int64_t)15;
  return ((num_slots(c_max) * 12) + 15) & ~(int64_t)15;
}

// rows per block so that rows * bytes <= budget, at least 1 and at most 1024/G
static int rows_per_block(int G, int64_t row_bytes, int64_t budget) {
  int R = 256 / G;
  if (R < 1) R = 1;
  if (row_bytes > 0) {
    while (R > 1 && (int64_t)R * row_bytes > budget) R >>= 1;
  }
  return R;
}

static int64_t g_lds_budget = 64 * 1024;
void set_lds_budget(int64_t b) { g_lds_budget = b; }

template <int G, bool BM>
static void sym_lds(const RowArgs& a, int32_t row_words, int32_t* rowcnt, hipStream_t s) {
  const int64_t rb = (int64_t)row_words * 4;
  const int R = G > kWave ? 1 : rows_per_block(G, rb, g_lds_budget);
  const size_t lds = (size_t)R * rb;
  const unsigned blocks = (unsigned)((a.nrows + R - 1) / R);
  hipFuncSetAttribute((const void*)k_symbolic_lds<G, BM>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_symbolic_lds<G, BM>), dim3(blocks), dim3(R * G), lds, s, a, row_words, rowcnt);
}

template <int G, bool BM>
static void num_lds(const RowArgs& a, int32_t words_max, int32_t c_max, const int64_t* rowoff, int32_t* ccol,
                    double* cval, hipStream_t s) {
  int64_t rb;
  if (BM) rb = (((int64_t)words_max * 8 + (int64_t)c_max * 12) + 15) & ~(int64_t)15;
  else rb = ((num_slots(c_max) * 12) + 15) & ~(int64_t)15;
  const int R = G > kWave ? 1 : rows_per_block(G, rb, g_lds_budget);
  const size_t lds = (size_t)R * rb;
  const unsigned blocks = (unsigned)((a.nrows + R - 1) / R);
  hipFuncSetAttribute((const void*)k_numeric_lds<G, BM>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_numeric_lds<G, BM>), dim3(blocks), dim3(R * G), lds, s, a, words_max, c_max, rowoff, ccol,
                     cval);
}

void launch_symbolic_bin(int bin, const RowArgs& a, int32_t lds_row_words, int32_t* rowcnt, uint32_t* gscratch,
                         int32_t gslots, hipStream_t s) {
  if (a.nrows <= 0 || bin <= 0) return;
  if (bin == kBinGlBitmap) {
    const unsigned blocks = (unsigned)(a.nrows < gslots ? a.nrows : gslots);
    hipLaunchKernelGGL(k_symbolic_global, dim3(blocks), dim3(256), 0, s, a, lds_row_words, gscratch, rowcnt);
    return;
  }
  if (bin == kBinWgBitmap) { sym_lds<256, true>(a, lds_row_words, rowcnt, s); return; }
  if (bin == kBinWgHash) { sym_lds<256, false>(a, lds_row_words, rowcnt, s); return; }
  const int G = bin_group(bin);
  const bool bm = bin_is_bitmap(bin);
  switch (G * 2 + (bm ? 1 : 0)) {
    case 17: sym_lds<8, true>(a, lds_row_words, rowcnt, s); break;
    case 16: sym_lds<8, false>(a, lds_row_words, rowcnt, s); break;
    case 33: sym_lds<16, true>(a, lds_row_words, rowcnt, s); break;
    case 32: sym_lds<16, false>(a, lds_row_words, rowcnt, s); break;
    case 65: sym_lds<32, true>(a, lds_row_words, rowcnt, s); break;
    case 64: sym_lds<32, false>(a, lds_row_words, rowcnt, s); break;
    case 129: sym_lds<64, true>(a, lds_row_words, rowcnt, s); break;
    case 128: sym_lds<64, false>(a, lds_row_words, rowcnt, s); break;
    default: break;
  }
}

void launch_numeric_bin(int bin, const RowArgs& a, int32_t words_max, int32_t c_max, const int64_t* rowoff,
                        int32_t* ccol, double* cval, uint32_t* gscratch, int32_t gslots, hipStream_t s) {
  if (a.nrows <= 0 || bin <= 0) return;
  if (bin == kBinGlBitmap) {
    const unsigned blocks = (unsigned)(a.nrows < gslots ? a.nrows : gslots);
    hipLaunchKernelGGL(k_numeric_global, dim3(blocks), dim3(256), 0, s, a, words_max, gscratch, rowoff, ccol, cval);
    return;
  }
  if (bin == kBinWgBitmap) { num_lds<256, true>(a, words_max, c_max, rowoff, ccol, cval, s); return; }
  if (bin == kBinWgHash) { num_lds<256, false>(a, words_max, c_max, rowoff, ccol, cval, s); return; }
  const int G = bin_group(bin);
  const bool bm = bin_is_bitmap(bin);
  switch (G * 2 + (bm ? 1 : 0)) {
    case 17: num_lds<8, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 16: num_lds<8, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 33: num_lds<16, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 32: num_lds<16, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 65: num_lds<32, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 64: num_lds<32, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 129: num_lds<64, true>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    case 128: num_lds<64, false>(a, words_max, c_max, rowoff, ccol, cval, s); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// AAtS mirror: lower counts, placement, per-row sort of the lower segments
// ---------------------------------------------------------------------------
__global__ void k_lower_count(int32_t n, const int64_t* __restrict__ uoff, const int32_t* __restrict__ ucol,
                              int32_t* __restrict__ low) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t r = (int64_t)blockIdx.x * gpb + threadIdx.x / G; r < n; r += (int64_t)gridDim.x * gpb) {
    for (int64_t t = uoff[r] + lane; t < uoff[r + 1]; t += G) {
      const int c = ucol[t];
      if (c > r) atomicAdd(&low[c], 1);
    }
  }
}
void launch_lower_count(int32_t n, const int64_t* uoff, const int32_t* ucol, int32_t* low, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_lower_count, dim3((unsigned)blocks), dim3(256), 0, s, n, uoff, ucol, low);
}

__global__ void k_add_counts(int32_t n, const int32_t* __restrict__ a, const int64_t* __restrict__ boff,
                             int32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] + (int32_t)(boff[i + 1] - boff[i]);
}
void launch_add_counts(int32_t n, const int32_t* a, const int64_t* boff, int32_t* out, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_add_counts, dim3((unsigned)blocks), dim3(256), 0, s, n, a, boff, out);
}

__global__ void k_mirror_scatter(int32_t n, const int64_t* __restrict__ uoff, const int32_t* __restrict__ ucol,
                                 const double* __restrict__ uval, const int64_t* __restrict__ coff,
                                 const int32_t* __restrict__ low, int32_t* __restrict__ cursor,
                                 int32_t* __restrict__ ccol, double* __restrict__ cval) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t r = (int64_t)blockIdx.x * gpb + threadIdx.x / G; r < n; r += (int64_t)gridDim.x * gpb) {
    const int64_t u0 = uoff[r], u1 = uoff[r + 1];
    const int64_t dst = coff[r] + low[r];
    for (int64_t t = u0 + lane; t < u1; t += G) {
      const int c = ucol[t];
      const double v = uval[t];
      ccol[dst + (t - u0)] = c;
      cval[dst + (t - u0)] = v;
      if (c > r) {
        const int64_t pos = coff[c] + atomicAdd(&cursor[c], 1);
        ccol[pos] = (int32_t)r;
        cval[pos] = v;
      }
    }
  }
}
void launch_mirror_scatter(int32_t n, const int64_t* uoff, const int32_t* ucol, const double* uval,
                           const int64_t* coff, const int32_t* low, int32_t* cursor, int32_t* ccol, double* cval,
                           hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_mirror_scatter, dim3((unsigned)blocks), dim3(256), 0, s, n, uoff, ucol, uval, coff, low,
                     cursor, ccol, cval);
}

// Segment sort.  Short segments: one wave per segment, in LDS.  Up to 4096:
// one workgroup per segment in LDS.  Longer: one workgroup per segment,
// bitonic network directly in global memory.
constexpr int kSegWave = 256, kSegWg = 4096;

__global__ void __launch_bounds__(256) k_segsort_wave(int32_t n, const int64_t* __restrict__ off,
                                                      const int32_t* __restrict__ len, int32_t* __restrict__ col,
                                                      double* __restrict__ val) {
  __shared__ int32_t sk[4][kSegWave];
  __shared__ double sv[4][kSegWave];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < n; r += (int64_t)gridDim.x * 4) {
    const int L = len[r];
    if (L <= 1 || L > kSegWave) continue;
    const int64_t o = off[r];
    for (int t = lane; t < L; t += kWave) { sk[w][t] = col[o + t]; sv[w][t] = val[o + t]; }
    gsync<kWave>();
    lds_sort_pairs<kWave>(sk[w], sv[w], L, lane);
    for (int t = lane; t < L; t += kWave) { col[o + t] = sk[w][t]; val[o + t] = sv[w][t]; }
    gsync<kWave>();
  }
}

__global__ void __launch_bounds__(256) k_segsort_wg(int32_t n, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ len, int32_t* __restrict__ col,
                                                    double* __restrict__ val) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  double* sv = (double*)lds_raw;
  int32_t* sk = (int32_t*)(lds_raw + kSegWg * 8);
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const int L = len[r];
    if (L <= kSegWave || L > kSegWg) continue;
    const int64_t o = off[r];
    __syncthreads();
    for (int t = threadIdx.x; t < L; t += 256) { sk[t] = col[o + t]; sv[t] = val[o + t]; }
    __syncthreads();
    lds_sort_pairs<256>(sk, sv, L, threadIdx.x);
    for (int t = threadIdx.x; t < L; t += 256) { col[o + t] = sk[t]; val[o + t] = sv[t]; }
  }
}

__global__ void __launch_bounds__(256) k_segsort_global(int32_t n, const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ len, int32_t* __restrict__ col,
                                                        double* __restrict__ val) {
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const int L = len[r];
    if (L <= kSegWg) continue;
    const int64_t o = off[r];
    __syncthreads();
    lds_sort_pairs<256>(col + o, val + o, L, threadIdx.x);
  }
}

void launch_segment_sort(int32_t n, const int64_t* off, const int32_t* len, int32_t* col, double* val,
                         int32_t max_len, hipStream_t s) {
  if (n <= 0 || max_len <= 1) return;
  int64_t blocks = ((int64_t)n + 3) / 4;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_segsort_wave, dim3((unsigned)blocks), dim3(256), 0, s, n, off, len, col, val);
  if (max_len > kSegWave) {
    const size_t lds = (size_t)kSegWg * 12;
    hipFuncSetAttribute((const void*)k_segsort_wg, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int64_t b2 = n < 8192 ? n : 8192;
    hipLaunchKernelGGL(k_segsort_wg, dim3((unsigned)b2), dim3(256), lds, s, n, off, len, col, val);
  }
  if (max_len > kSegWg) {
    int64_t b3 = n < 1024 ? n : 1024;
    hipLaunchKernelGGL(k_segsort_global, dim3((unsigned)b3), dim3(256), 0, s, n, off, len, col, val);
  }
}

}  // namespace appt
