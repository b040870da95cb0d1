// descriptors.hip -- structural descriptors on the GPU.
//
// Restates the host descriptor pass of bench/dispatch.cu:26-37 (deg_cv with
// ddof=0 over the stored row lengths, fill_aa = sum_p deg(col_p)/nnz,
// fill_aat = sum_k colcount_k^2/nnz), the psym estimate of
// src/fetch_and_prep.py:39-46 (the numeric pattern after merging duplicates),
// the decision rule pick_ours of bench/dispatch.cu:14-17, and adds bandwidth
// and products.  Reads A once or twice with no staging, in four kernels and
// one host synchronisation.
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

// One G-lane group per row: degree, products, bandwidth, order; writes the
// row's strictly-ascending flag for the psym pass.
__global__ void __launch_bounds__(256) k_desc_rows(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                            uint8_t* __restrict__ ascending, unsigned long long* __restrict__ acc) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  unsigned long long sd2 = 0, mx = 0, bw = 0, uns = 0, prod = 0, pmax = 0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    const int s = ap[i], e = ap[i + 1];
    unsigned long long f = 0, b = 0, u = 0;
    for (int p = s + lane; p < e; p += G) {
      const int c = ac[p];
      f += (unsigned long long)(ap[c + 1] - ap[c]);
      const long long d = (long long)c - i;
      const unsigned long long ad = (unsigned long long)(d < 0 ? -d : d);
      b = ad > b ? ad : b;
      if (p + 1 < e && ac[p + 1] <= c) u = 1;
    }
    f = gsum<G, unsigned long long>(f, nullptr);
    b = gmax<G>(b);
    u = gmax<G>(u);
    if (lane == 0) {
      ascending[i] = u ? 0 : 1;
      const unsigned long long d = (unsigned long long)(e - s);
      sd2 += d * d;
      mx = d > mx ? d : mx;
      bw = b > bw ? b : bw;
      uns += u;
      prod += f;
      pmax = f > pmax ? f : pmax;
    }
  }
  // the block's groups, then one atomic per accumulator and block
  __shared__ unsigned long long red[6][16];
  const int g = threadIdx.x / G;
  if (lane == 0) {
    red[0][g] = sd2;
    red[1][g] = mx;
    red[2][g] = bw;
    red[3][g] = uns;
    red[4][g] = prod;
    red[5][g] = pmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int ng = blockDim.x / G;
    for (int t = 1; t < ng; t++) {
      sd2 += red[0][t];
      mx = red[1][t] > mx ? red[1][t] : mx;
      bw = red[2][t] > bw ? red[2][t] : bw;
      uns += red[3][t];
      prod += red[4][t];
      pmax = red[5][t] > pmax ? red[5][t] : pmax;
    }
    if (sd2) {
      const unsigned long long old = atomicAdd(&acc[kDescSumD2], sd2);
      if (old + sd2 < old) atomicAdd(&acc[kDescSumD2Hi], 1ull);
    }
    if (mx) atomicMax(&acc[kDescMaxD], mx);
    if (bw) atomicMax(&acc[kDescBandwidth], bw);
    if (uns) atomicAdd(&acc[kDescUnsorted], uns);
    if (prod) atomicAdd(&acc[kDescProducts], prod);
    if (pmax) atomicMax(&acc[kDescMaxRowProd], pmax);
  }
}

__global__ void __launch_bounds__(256) k_sumsq(int32_t n, const int32_t* __restrict__ cnt,
                                               unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[4];
  unsigned long long s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long c = (unsigned long long)cnt[i];
    s += c * c;
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0 && red[0] + red[1] + red[2] + red[3]) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

// Merged value of position (r, c): the stored values of row r at column c,
// summed in storage order; *found says whether any is stored.  A strictly
// ascending row holds at most one, found by binary search.
__device__ __forceinline__ double merged_value(const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                                               const double* __restrict__ av, bool asc, int32_t r, int32_t c,
                                               bool* found) {
  int lo = ap[r], hi = ap[r + 1];
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

// psym: each position (i, j) of P -- the first stored occurrence of column j
// in row i, kept when the row's merged value there is not zero -- counts in
// |P|, and in |P & P^T| when (j, i) is in P too.
__global__ void __launch_bounds__(256) k_psym(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                       const double* __restrict__ av, const uint8_t* __restrict__ ascending,
                       unsigned long long* __restrict__ acc) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  unsigned long long hits = 0, pat = 0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    const int s = ap[i], e = ap[i + 1];
    const bool asc = ascending[i] != 0;
    for (int p = s + lane; p < e; p += G) {
      const int32_t j = ac[p];
      double v;
      if (asc) {
        v = av[p];
      } else {
        bool first = true;
        for (int q = s; q < p; q++)
          if (ac[q] == j) {
            first = false;
            break;
          }
        if (!first) continue;
        v = 0.0;
        for (int q = p; q < e; q++)
          if (ac[q] == j) v += av[q];
      }
      if (!(v != 0.0)) continue;  // NaN stays in the pattern, as (A != 0) keeps it
      pat++;
      bool found;
      const double w = merged_value(ap, ac, av, ascending[j] != 0, j, (int32_t)i, &found);
      if (found && w != 0.0) hits++;
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    hits += __shfl_xor(hits, m, 64);
    pat += __shfl_xor(pat, m, 64);
  }
  __shared__ unsigned long long red[2][4];
  if (threadIdx.x % 64 == 0) {
    red[0][threadIdx.x / 64] = hits;
    red[1][threadIdx.x / 64] = pat;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    hits = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    pat = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    if (hits) atomicAdd(&acc[kDescInter], hits);
    if (pat) atomicAdd(&acc[kDescPattern], pat);
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
  uint8_t* asc = (uint8_t*)ensure(ws_.symbin, (size_t)n);
  launch_fill_i32((int32_t*)acc, 2 * kDescWords, 0, s);
  launch_fill_i32(cnt, n, 0, s);
  const int64_t blocks = std::min<int64_t>(((int64_t)n * 16 + 255) / 256, 8192);
  hipLaunchKernelGGL(k_desc_rows, dim3((unsigned)blocks), dim3(256), 0, s, n, A.rowptr, A.col, asc, acc);
  launch_colcount(nnz, A.col, cnt, s);
  const int64_t b2 = std::min<int64_t>(((int64_t)n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_sumsq, dim3((unsigned)b2), dim3(256), 0, s, n, cnt, acc + kDescColSq);
  hipLaunchKernelGGL(k_psym, dim3((unsigned)blocks), dim3(256), 0, s, n, A.rowptr, A.col, A.val, asc, acc);
  APPT_HIP(hipGetLastError());
  unsigned long long h[kDescWords];
  APPT_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
  APPT_HIP(hipStreamSynchronize(s));

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
