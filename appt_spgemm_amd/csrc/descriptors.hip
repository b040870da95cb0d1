// descriptors.hip -- structural descriptors on the GPU.
//
// Restates the host descriptor pass of bench/dispatch.cu:26-37 (deg_cv with
// ddof=0 over the stored row lengths, fill_aa = sum_p deg(col_p)/nnz,
// fill_aat = sum_k colcount_k^2/nnz), the psym estimate of
// src/fetch_and_prep.py:39-46 (the numeric pattern after merging duplicates),
// the decision rule pick_ours of bench/dispatch.cu:14-17, and adds bandwidth
// and products.  Reads A as it lies (no staging): per-entry products, order
// and bandwidth, per-row statistics, the column histogram, psym; one host
// synchronisation.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.hpp"
#include "engine.hpp"

namespace appt {

static unsigned long long* as_ull(void* p) { return static_cast<unsigned long long*>(p); }

// Accumulator words of one descriptor pass.
enum : int {
  kDescSumD2 = 0,   // sum of d^2, low 64 bits
  kDescSumD2Hi,     // carries out of the low word
  kDescMaxD,
  kDescBandwidth,
  kDescUnsorted,    // rows not strictly ascending
  kDescProducts,    // sum_p deg(col_p)
  kDescMaxRowProd,
  kDescColSq,       // sum_k colcount_k^2
  kDescInter,       // |P & P^T|
  kDescPattern,     // |P|
  kDescWords
};

// Entry-parallel kernels: a block takes kDescChunk consecutive entries of A
// (lane-strided, so loads coalesce); the rows of its chunk come from one
// search per chunk (k_desc_chunks) and are expanded into a row id per entry
// in LDS, so a power-law hub row is spread over many blocks instead of one
// lane group.
constexpr int kDescChunk = 2048;

__device__ __forceinline__ int32_t desc_row_of(const int32_t* __restrict__ ap, int32_t lo, int32_t hi, int64_t e) {
  while (lo < hi) {  // the last row r in [lo, hi] with ap[r] <= e
    const int mid = (int)(((int64_t)lo + hi + 1) >> 1);
    if ((int64_t)ap[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// crow[c] = the row of entry min(c * kDescChunk, nnz - 1), c = 0 .. nchunks:
// block c's entries lie in rows [crow[c], crow[c + 1]].  One thread per
// chunk, so no block starts with a search over all rows.
__global__ void k_desc_chunks(int32_t n, int64_t nnz, const int32_t* __restrict__ ap, int64_t nchunks,
                              int32_t* __restrict__ crow) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= nchunks; c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = c * kDescChunk < nnz ? c * kDescChunk : nnz - 1;
    crow[c] = desc_row_of(ap, 0, n - 1, e);
  }
}

// Row of every entry of the block's chunk [e0, e1), in LDS (rows r0 .. r1
// expanded by the block's threads, a row at a time each).
__device__ __forceinline__ void desc_rowids(const int32_t* __restrict__ ap, int32_t r0, int32_t r1, int64_t e0,
                                            int64_t e1, int32_t* rowid) {
  for (int32_t r = r0 + (int32_t)threadIdx.x; r <= r1; r += (int32_t)blockDim.x) {
    const int64_t a = (int64_t)ap[r] > e0 ? (int64_t)ap[r] : e0;
    const int64_t b = (int64_t)ap[r + 1] < e1 ? (int64_t)ap[r + 1] : e1;
    for (int64_t e = a; e < b; e++) rowid[e - e0] = r;
  }
  __syncthreads();
}

// Sum of v over the lanes of a wave that share key (equal keys on
// consecutive lanes); the segment's last lane gets the total and *last.
__device__ __forceinline__ unsigned long long seg_sum(unsigned long long v, int key, bool live, bool* last) {
  const int lane = threadIdx.x % kWave;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const unsigned long long o = __shfl_up(v, d, kWave);
    const int ko = __shfl_up(key, d, kWave);
    if (lane >= d && ko == key) v += o;
  }
  const int kn = __shfl_down(key, 1, kWave);
  const bool ln = __shfl_down(live ? 1 : 0, 1, kWave) != 0;
  *last = live && (lane == kWave - 1 || kn != key || !ln);
  return v;
}

// Zero the accumulators, the column histogram, the per-row products and the
// per-row order flags.
__global__ void k_desc_clear(int32_t n, unsigned long long* __restrict__ acc, int32_t* __restrict__ cnt,
                             unsigned long long* __restrict__ rowf, uint8_t* __restrict__ notasc) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    cnt[i] = 0;
    rowf[i] = 0ull;
    notasc[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < kDescWords) acc[threadIdx.x] = 0ull;
}

// Per entry: products deg(col) (summed per row into rowf, and over A),
// bandwidth, the column histogram, and each row's order (notasc[i] = 1 on a
// descent or a repeated column).
__global__ void __launch_bounds__(256) k_desc_entries(int32_t n, int64_t nnz, const int32_t* __restrict__ ap,
                                                      const int32_t* __restrict__ ac,
                                                      const int32_t* __restrict__ crow, int32_t* __restrict__ cnt,
                                                      unsigned long long* __restrict__ rowf,
                                                      uint8_t* __restrict__ notasc,
                                                      unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long red[2][4];
  __shared__ int32_t rowid[kDescChunk];
  const int64_t e0 = (int64_t)blockIdx.x * kDescChunk;
  const int64_t e1 = e0 + kDescChunk < nnz ? e0 + kDescChunk : nnz;
  desc_rowids(ap, crow[blockIdx.x], crow[blockIdx.x + 1], e0, e1, rowid);
  unsigned long long prod = 0, bw = 0;
  // a fixed trip count per wave keeps the segmented sums converged
  for (int64_t b = e0; b < e1; b += blockDim.x) {
    const int64_t e = b + threadIdx.x;
    const bool live = e < e1;
    unsigned long long f = 0;
    int32_t r = -1;
    if (live) {
      r = rowid[e - e0];
      const int c = ac[e];
      f = (unsigned long long)(ap[c + 1] - ap[c]);
      const long long d = (long long)c - r;
      const unsigned long long ad = (unsigned long long)(d < 0 ? -d : d);
      bw = ad > bw ? ad : bw;
      if (e + 1 < (int64_t)ap[r + 1] && ac[e + 1] <= c) notasc[r] = 1;
      atomicAdd(&cnt[c], 1);
      prod += f;
    }
    bool last;
    const unsigned long long seg = seg_sum(f, r, live, &last);
    if (last) atomicAdd(&rowf[r], seg);
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    prod += __shfl_xor(prod, m, 64);
    const unsigned long long o = __shfl_xor(bw, m, 64);
    bw = o > bw ? o : bw;
  }
  if (threadIdx.x % 64 == 0) {
    red[0][threadIdx.x / 64] = prod;
    red[1][threadIdx.x / 64] = bw;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    prod = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    bw = red[1][0];
    for (int w = 1; w < 4; w++) bw = red[1][w] > bw ? red[1][w] : bw;
    if (prod) atomicAdd(&acc[kDescProducts], prod);
    if (bw) atomicMax(&acc[kDescBandwidth], bw);
  }
}

// Merged value of position (r, c): the stored values of row r at column c,
// summed in storage order; *found says whether any is stored.  A strictly
// ascending row holds at most one, found by binary search.
// guess: a position to try first (stencil-like patterns put (r, c) at the
// mirror of (c, r)'s offset in its row).
__device__ __forceinline__ double merged_value(const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                                               const double* __restrict__ av, bool asc, int32_t r, int32_t c,
                                               int64_t guess, bool* found) {
  int lo = ap[r], hi = ap[r + 1];
  if (asc && guess >= lo && guess < hi && ac[guess] == c) {
    *found = true;
    return av[guess];
  }
  if (asc) {
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const int v = ac[mid];
      if (v == c) {
        *found = true;
        return av[mid];
      }
      if (v < c) lo = mid + 1;
      else hi = mid;
    }
    *found = false;
    return 0.0;
  }
  double v = 0.0;
  bool f = false;
  for (int q = lo; q < hi; q++)
    if (ac[q] == c) {
      v += av[q];
      f = true;
    }
  *found = f;
  return v;
}

// Per row (grid-stride): sum of d^2 (with its carry word), longest row, rows
// out of order, largest row products, sum of squared column counts.  Per
// entry (block chunks as above): psym -- each position (i, j) of P, the
// first stored occurrence of column j in row i kept when the row's merged
// value there is not zero, counts in |P|, and in |P & P^T| when (j, i) is in
// P too.  One atomic per accumulator and block.
__global__ void __launch_bounds__(256) k_desc_finish(int32_t n, int64_t nnz, const int32_t* __restrict__ ap,
                                                     const int32_t* __restrict__ ac, const double* __restrict__ av,
                                                     const int32_t* __restrict__ crow, const int32_t* __restrict__ cnt,
                                                     const unsigned long long* __restrict__ rowf,
                                                     const uint8_t* __restrict__ notasc,
                                                     unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long red[7][4];
  __shared__ int32_t rowid[kDescChunk];
  unsigned long long sd2 = 0, mx = 0, uns = 0, pmax = 0, csq = 0, hits = 0, pat = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long d = (unsigned long long)(ap[i + 1] - ap[i]);
    sd2 += d * d;
    mx = d > mx ? d : mx;
    uns += notasc[i];
    pmax = rowf[i] > pmax ? rowf[i] : pmax;
    const unsigned long long c = (unsigned long long)cnt[i];
    csq += c * c;
  }
  const int64_t e0 = (int64_t)blockIdx.x * kDescChunk;
  const int64_t e1 = e0 + kDescChunk < nnz ? e0 + kDescChunk : nnz;
  if (e0 < e1) {
    desc_rowids(ap, crow[blockIdx.x], crow[blockIdx.x + 1], e0, e1, rowid);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
      const int32_t r = rowid[e - e0];
      const int32_t j = ac[e];
      double v;
      if (notasc[r] == 0) {
        v = av[e];
      } else {
        const int s = ap[r], t = ap[r + 1];
        bool first = true;
        for (int q = s; q < (int)e; q++)
          if (ac[q] == j) {
            first = false;
            break;
          }
        if (!first) continue;
        v = 0.0;
        for (int q = (int)e; q < t; q++)
          if (ac[q] == j) v += av[q];
      }
      if (!(v != 0.0)) continue;  // NaN stays in the pattern, as (A != 0) keeps it
      pat++;
      bool found;
      const int64_t mirror = (int64_t)ap[j] + ((int64_t)ap[r + 1] - 1 - e);
      const double w = merged_value(ap, ac, av, notasc[j] == 0, j, r, mirror, &found);
      if (found && w != 0.0) hits++;
    }
  }
  unsigned long long v[7] = {sd2, mx, uns, pmax, csq, hits, pat};
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
#pragma unroll
    for (int k = 0; k < 7; k++) {
      const unsigned long long o = __shfl_xor(v[k], m, 64);
      v[k] = (k == 1 || k == 3) ? (o > v[k] ? o : v[k]) : v[k] + o;
    }
  }
  if (threadIdx.x % 64 == 0)
    for (int k = 0; k < 7; k++) red[k][threadIdx.x / 64] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; k++)
      for (int w = 1; w < 4; w++)
        v[k] = (k == 1 || k == 3) ? (red[k][w] > v[k] ? red[k][w] : v[k]) : v[k] + red[k][w];
    if (v[0]) {
      const unsigned long long old = atomicAdd(&acc[kDescSumD2], v[0]);
      if (old + v[0] < old) atomicAdd(&acc[kDescSumD2Hi], 1ull);
    }
    if (v[1]) atomicMax(&acc[kDescMaxD], v[1]);
    if (v[2]) atomicAdd(&acc[kDescUnsorted], v[2]);
    if (v[3]) atomicMax(&acc[kDescMaxRowProd], v[3]);
    if (v[4]) atomicAdd(&acc[kDescColSq], v[4]);
    if (v[5]) atomicAdd(&acc[kDescInter], v[5]);
    if (v[6]) atomicAdd(&acc[kDescPattern], v[6]);
  }
}

static bool pick_ours_rule(double dcv, double fill, int64_t nnz) {
  if (dcv <= 0.63) return fill <= 53.23;
  return nnz > 20164 && fill <= 47.17;
}

void Engine::descriptors(const appt_csr_view& A, appt_op op, appt_descriptors* out) {
  validate(A);
  APPT_HIP(hipSetDevice(device_));
  hipStream_t s = stream_;
  const int32_t n = A.n;
  const int64_t nnz = A.nnz;
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->nnz = nnz;
  if (n == 0) return;
  unsigned long long* acc = as_ull(ensure(ws_.digest, sizeof(unsigned long long) * kDescWords));
  int32_t* cnt = (int32_t*)ensure(ws_.cursor, sizeof(int32_t) * ((size_t)n + 1));
  uint8_t* notasc = (uint8_t*)ensure(ws_.symbin, (size_t)n);
  auto* rowf = as_ull(ensure(ws_.rowf, sizeof(unsigned long long) * (size_t)n));
  const int64_t rb = std::min<int64_t>(((int64_t)n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_desc_clear, dim3((unsigned)rb), dim3(256), 0, s, n, acc, cnt, rowf, notasc);
  // a block per kDescChunk entries, and enough threads for the row statistics
  const int64_t eb = (nnz + kDescChunk - 1) / kDescChunk;
  int32_t* crow = (int32_t*)ensure(ws_.cur, sizeof(int32_t) * (size_t)(eb + 1));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_desc_chunks, dim3((unsigned)std::min<int64_t>((eb + 256) / 256, 1024)), dim3(256), 0, s, n,
                       nnz, A.rowptr, eb, crow);
    hipLaunchKernelGGL(k_desc_entries, dim3((unsigned)eb), dim3(256), 0, s, n, nnz, A.rowptr, A.col, crow, cnt,
                       rowf, notasc, acc);
  }
  hipLaunchKernelGGL(k_desc_finish, dim3((unsigned)std::max(eb, rb)), dim3(256), 0, s, n, nnz, A.rowptr, A.col,
                     A.val, crow, cnt, rowf, notasc, acc);
  APPT_HIP(hipGetLastError());
  // one copy into the engine's pinned staging, one synchronisation
  static_assert(sizeof(HostStats::sym) >= sizeof(unsigned long long) * kDescWords, "staging too small");
  APPT_HIP(hipMemcpyAsync(hs_->sym, acc, sizeof(unsigned long long) * kDescWords, hipMemcpyDeviceToHost, s));
  wait_stream();
  unsigned long long h[kDescWords];
  std::memcpy(h, hs_->sym, sizeof(h));

  // variance exactly: n^2 var = n * sum d^2 - nnz^2 (the two-pass
  // sum (d - mean)^2 of bench/dispatch.cu:28-30, without its rounding)
  const unsigned __int128 sd2 = ((unsigned __int128)h[kDescSumD2Hi] << 64) | h[kDescSumD2];
  const __int128 num = (__int128)((unsigned __int128)n * sd2) - (__int128)nnz * (__int128)nnz;
  const long double var = (long double)num / ((long double)n * (long double)n);
  const double dmean = (double)nnz / n;
  out->deg_mean = dmean;
  out->deg_cv = (double)(sqrtl(var > 0 ? var : 0.0L) / (long double)(dmean > 1e-9 ? dmean : 1e-9));
  out->deg_max = (int32_t)h[kDescMaxD];
  out->bandwidth = (int32_t)h[kDescBandwidth];
  out->products_aa = (int64_t)h[kDescProducts];
  out->max_row_products_aa = (int32_t)std::min<unsigned long long>(h[kDescMaxRowProd], (unsigned long long)INT32_MAX);
  out->products_aat = (int64_t)h[kDescColSq];
  out->fill_aa = nnz > 0 ? (double)h[kDescProducts] / nnz : 0.0;
  out->fill_aat = nnz > 0 ? (double)h[kDescColSq] / nnz : 0.0;
  out->psym = (double)h[kDescInter] / (double)std::max<unsigned long long>(h[kDescPattern], 1);
  const double fill = op == APPT_OP_AA ? out->fill_aa : out->fill_aat;
  out->pick_ours = pick_ours_rule(out->deg_cv, fill, nnz) ? 1 : 0;
}

}  // namespace appt
