// gram.hip -- triangle-only Gram product C = A*A^T (op AAtS).
//
// The reference hashes only j >= i for AAtS in both phases and mirrors the
// upper triangle afterwards (src/spgemm_engine.cuh:82,104,121-127).  Here the
// other triangle's products are never formed at all: B = A^T holds column k
// of A with its rows ascending, so for an A entry p = (i, k) the products of
// row i with j >= i are exactly the suffix of B row k from row i (and those
// with j <= i the prefix up to row i).  Each A entry gets that suffix (or
// prefix) as a B-row descriptor of its own ({start, end, min, max}, the
// layout every bin kernel reads through ac[p]), and the unchanged symbolic
// and numeric bins compute T = the rows of one triangle of C, diagonal
// included, with about half of A*A^T's products.  Upper (the reference's
// choice) keeps the low-index rows whole; lower keeps the high-index ones,
// which unloads power-law matrices whose hubs sit at low indices.  The
// expand stage then builds the full symmetric C: row j = its row of T
// joined with column j of T off the diagonal (rows ascending), on the side
// where those columns fall.  The column-major order of the off-diagonal
// entries comes from a stable radix sort of them by column, so the placement
// is fixed and the mirrored values are bit-identical to their images.
#include "device_common.hpp"
#include "kernels.hpp"

namespace appt {

// One 16-lane group per A row: the suffix of B row ac[p] from its first row
// >= i (lower: the prefix up to its last row <= i), found by binary search
// (B rows ascending; A rows with repeated columns repeat i there, and the
// range keeps every repeat).
__global__ void k_gram_operand(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                               const int4* __restrict__ bmeta, const int32_t* __restrict__ bc, int lower,
                               int4* __restrict__ vmeta, int32_t* __restrict__ vac) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    const int s = ap[i], e = ap[i + 1];
    for (int p = s + lane; p < e; p += G) {
      const int4 m = bmeta[ac[p]];
      int lo = m.x, hi = m.y;
      // first position with row >= i (lower: > i)
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = bc[mid];
        if (v < (int32_t)i || (lower && v == (int32_t)i)) lo = mid + 1;
        else hi = mid;
      }
      if (lower) vmeta[p] = lo > m.x ? make_int4(m.x, lo, m.z, bc[lo - 1]) : make_int4(m.x, m.x, INT32_MAX, -1);
      else vmeta[p] = lo < m.y ? make_int4(lo, m.y, bc[lo], m.w) : make_int4(m.y, m.y, INT32_MAX, -1);
      vac[p] = p;
    }
  }
}

// sum over B rows of nnz(row)^2: the products of the full A*A^T (one
// atomic per block)
__global__ void __launch_bounds__(256) k_gram_products(int32_t nb, const int4* __restrict__ bmeta,
                                                       unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[4];
  unsigned long long s = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long d = (unsigned long long)(bmeta[k].y - bmeta[k].x);
    s += d * d;
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = red[0] + red[1] + red[2] + red[3];
    if (t) atomicAdd(out, t);
  }
}

void launch_gram_operand(int32_t n, const int32_t* ap, const int32_t* ac, int32_t nb, const int4* bmeta,
                         const int32_t* bc, int lower, int4* vmeta, int32_t* vac,
                         unsigned long long* full_products, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_gram_operand, dim3((unsigned)blocks), dim3(256), 0, s, n, ap, ac, bmeta, bc, lower, vmeta,
                     vac);
  int64_t b2 = ((int64_t)nb + 255) / 256;
  if (b2 > 1024) b2 = 1024;
  if (nb > 0) hipLaunchKernelGGL(k_gram_products, dim3((unsigned)b2), dim3(256), 0, s, nb, bmeta, full_products);
}

// ---------------------------------------------------------------------------
// expand
// ---------------------------------------------------------------------------
constexpr int kGramChunk = 2048;  // entries per block of the entry-parallel kernels

// Row of entry e0 under row offsets off[0..n] (the last row r with off[r] <= e0).
__device__ __forceinline__ int32_t row_of(const int64_t* __restrict__ off, int32_t n, int64_t e0) {
  int l = 0, h = n - 1;
  while (l < h) {
    const int mid = (int)(((int64_t)l + h + 1) >> 1);
    if (off[mid] <= e0) l = mid;
    else h = mid - 1;
  }
  return l;
}

// Entry-parallel walk over [0, total) of a CSR with offsets off: each block
// takes kGramChunk consecutive entries, finds its first row once, and each
// thread advances its own row as its entry index grows.  The walk stops at
// the last row; an entry at or past off[n] (offsets that disagree with
// `total`) is skipped, and the host check of off[n] reports it.
#define GRAM_ENTRY_LOOP(off, n, total, BODY)                                          \
  __shared__ int32_t row0_s;                                                          \
  const int64_t e0 = (int64_t)blockIdx.x * kGramChunk;                                \
  if (e0 >= (total)) return;                                                          \
  const int64_t e1 = e0 + kGramChunk < (total) ? e0 + kGramChunk : (total);           \
  if (threadIdx.x == 0) row0_s = row_of((off), (n), e0);                              \
  __syncthreads();                                                                    \
  int32_t r = row0_s;                                                                 \
  for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {                       \
    while (r < (n) - 1 && (off)[r + 1] <= e) r++;                                     \
    if ((off)[r + 1] <= e) continue;                                                  \
    BODY                                                                              \
  }

// mir[j] += 1 for every off-diagonal T entry (i, j): row j's mirrored entries
__global__ void __launch_bounds__(256) k_gram_mirrored(int32_t n, int64_t nu, const int64_t* __restrict__ uoff,
                                                       const int32_t* __restrict__ ucol, int32_t* __restrict__ mir) {
  GRAM_ENTRY_LOOP(uoff, n, nu, {
    const int32_t j = ucol[e];
    if (j != r) atomicAdd(&mir[j], 1);
  })
}

// strict[i] = T row i less its diagonal (every non-empty row of T has it:
// first in U's rows, last in L's); total[i] = mir[i] + T row i
__global__ void k_gram_counts(int32_t n, const int32_t* __restrict__ ucnt, const int32_t* __restrict__ lower,
                              int32_t* __restrict__ strict, int32_t* __restrict__ total) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = ucnt[i];
    strict[i] = u > 0 ? u - 1 : 0;
    total[i] = lower[i] + u;
  }
}

// off-diagonal T entries in row-major order: key = column, payload = (row, value)
__global__ void __launch_bounds__(256) k_gram_extract(int32_t n, int64_t nu, const int64_t* __restrict__ uoff,
                                                      const int32_t* __restrict__ ucol,
                                                      const double* __restrict__ uval,
                                                      const int64_t* __restrict__ sup_off, int lower,
                                                      int32_t* __restrict__ keys, int32_t* __restrict__ rows,
                                                      double* __restrict__ vals) {
  GRAM_ENTRY_LOOP(uoff, n, nu, {
    const int64_t t = e - uoff[r];
    const bool off_diag = lower ? e < uoff[r + 1] - 1 : t > 0;
    if (off_diag) {
      const int64_t d = sup_off[r] + (lower ? t : t - 1);
      keys[d] = ucol[e];
      rows[d] = r;
      vals[d] = uval[e];
    }
  })
}

// C row j, upper T: [column j of U above the diagonal: rows srow, values
// sval from lo_off[j]] ++ [U row j]; lower T: [L row j] ++ [column j of L
// below the diagonal]
__global__ void __launch_bounds__(256) k_gram_assemble(int32_t n, int64_t nc, const int64_t* __restrict__ coff,
                                                       int lower_t, const int32_t* __restrict__ lower,
                                                       const int64_t* __restrict__ lo_off,
                                                       const int32_t* __restrict__ srow,
                                                       const double* __restrict__ sval,
                                                       const int64_t* __restrict__ uoff,
                                                       const int32_t* __restrict__ ucol,
                                                       const double* __restrict__ uval, int32_t* __restrict__ ccol,
                                                       double* __restrict__ cval) {
  GRAM_ENTRY_LOOP(coff, n, nc, {
    const int64_t t = e - coff[r];
    // mirrored entries come first in an upper T's rows, last in a lower T's
    const int64_t own = uoff[r + 1] - uoff[r];
    const int64_t m = lower_t ? t - own : t;
    if (lower_t ? t >= own : t < (int64_t)lower[r]) {
      ccol[e] = srow[lo_off[r] + m];
      cval[e] = sval[lo_off[r] + m];
    } else {
      const int64_t q = uoff[r] + (lower_t ? t : t - lower[r]);
      ccol[e] = ucol[q];
      cval[e] = uval[q];
    }
  })
}

static unsigned chunk_blocks(int64_t total) { return (unsigned)((total + kGramChunk - 1) / kGramChunk); }

void launch_gram_mirrored(int32_t n, int64_t nu, const int64_t* uoff, const int32_t* ucol, int32_t* mir,
                          hipStream_t s) {
  if (n <= 0 || nu <= 0) return;
  hipLaunchKernelGGL(k_gram_mirrored, dim3(chunk_blocks(nu)), dim3(256), 0, s, n, nu, uoff, ucol, mir);
}

void launch_gram_counts(int32_t n, const int32_t* ucnt, const int32_t* lower, int32_t* strict, int32_t* total,
                        hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = ((int64_t)n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_gram_counts, dim3((unsigned)blocks), dim3(256), 0, s, n, ucnt, lower, strict, total);
}

void launch_gram_extract(int32_t n, int64_t nu, const int64_t* uoff, const int32_t* ucol, const double* uval,
                         const int64_t* sup_off, int lower, int32_t* keys, int32_t* rows, double* vals,
                         hipStream_t s) {
  if (n <= 0 || nu <= 0) return;
  hipLaunchKernelGGL(k_gram_extract, dim3(chunk_blocks(nu)), dim3(256), 0, s, n, nu, uoff, ucol, uval, sup_off,
                     lower, keys, rows, vals);
}

void launch_gram_assemble(int32_t n, int64_t nc, const int64_t* coff, int lower_t, const int32_t* mir,
                          const int64_t* mir_off, const int32_t* srow, const double* sval, const int64_t* uoff,
                          const int32_t* ucol, const double* uval, int32_t* ccol, double* cval, hipStream_t s) {
  if (n <= 0 || nc <= 0) return;
  hipLaunchKernelGGL(k_gram_assemble, dim3(chunk_blocks(nc)), dim3(256), 0, s, n, nc, coff, lower_t, mir, mir_off,
                     srow, sval, uoff, ucol, uval, ccol, cval);
}

}  // namespace appt
