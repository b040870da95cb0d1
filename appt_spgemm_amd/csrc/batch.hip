// batch.hip -- kernels of the batched path: many independent matrices run
// as one block-diagonal product.
//
// Packed layout (all device, global row numbering): matrix b owns rows
// [blk_row[b], blk_row[b+1]) and entries [blk_nnz[b], blk_nnz[b+1]) of one
// int32 rowptr with global offsets; its column indices stay local (0..n_b-1).
// The A side of the pipeline reads B-row ids acg[p] = col[p] + row base; the
// B side of an A*A block is the block's own rows (local columns), that of an
// A*A^T block its transpose, stored after the nnz(A) packed entries in the
// same col/val buffers, so output columns come out local as well.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace appt {

// rowbase[i] / rowblk[i] / rowop[i]: first row, index and op (0 A*A, 1 A*A^T)
// of the matrix that owns global row i.
__global__ void k_batch_rows(int32_t N, int32_t K, const int32_t* __restrict__ blk_row,
                             const uint8_t* __restrict__ blk_op, int32_t* __restrict__ rowbase,
                             int32_t* __restrict__ rowblk, uint8_t* __restrict__ rowop) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    // last block whose first row is <= i (empty blocks share their start
    // with the next block and are skipped by taking the last)
    int l = 0, h = K - 1;
    while (l < h) {
      const int mid = (l + h + 1) >> 1;
      if (blk_row[mid] <= i) l = mid;
      else h = mid - 1;
    }
    rowbase[i] = blk_row[l];
    rowblk[i] = l;
    rowop[i] = blk_op[l];
  }
}
void launch_batch_rows(int32_t N, int32_t K, const int32_t* blk_row, const uint8_t* blk_op, int32_t* rowbase,
                       int32_t* rowblk, uint8_t* rowop, hipStream_t s) {
  if (N <= 0) return;
  int64_t blocks = ((int64_t)N + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_batch_rows, dim3((unsigned)blocks), dim3(256), 0, s, N, K, blk_row, blk_op, rowbase, rowblk,
                     rowop);
}

// acg[p] = col[p] + rowbase(row of p); with tcnt, entries of A*A^T rows
// also count their transpose row (tcnt, zero beforehand); with rid, they
// record their row local to the matrix.
__global__ void k_batch_acg(int32_t N, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const int32_t* __restrict__ rowbase, const uint8_t* __restrict__ rowop,
                            int32_t* __restrict__ acg, int32_t* __restrict__ tcnt, int32_t* __restrict__ rid) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < N; i += (int64_t)gridDim.x * gpb) {
    const int base = rowbase[i];
    const bool t = rowop[i] != 0;
    for (int p = rp[i] + lane; p < rp[i + 1]; p += G) {
      const int k = col[p] + base;
      acg[p] = k;
      if (t && tcnt) atomicAdd(&tcnt[k], 1);
      if (t && rid) rid[p] = (int32_t)(i - base);
    }
  }
}
void launch_batch_acg(int32_t N, const int32_t* rp, const int32_t* col, const int32_t* rowbase,
                      const uint8_t* rowop, int32_t* acg, int32_t* tcnt, int32_t* rid, hipStream_t s) {
  if (N <= 0) return;
  int64_t blocks = ((int64_t)N * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_batch_acg, dim3((unsigned)blocks), dim3(256), 0, s, N, rp, col, rowbase, rowop, acg, tcnt,
                     rid);
}

// Transpose entries of A*A^T rows into the region after the nnz(A) packed
// entries: entry (i, k) lands in row k with local column i - rowbase(i).
__global__ void k_batch_tscatter(int32_t N, const int32_t* __restrict__ rp, const int32_t* __restrict__ acg,
                                 const int32_t* __restrict__ rowbase, const uint8_t* __restrict__ rowop,
                                 int32_t* __restrict__ cursor, int64_t nnz_a, int32_t* __restrict__ bcol,
                                 double* __restrict__ bval) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < N; i += (int64_t)gridDim.x * gpb) {
    if (!rowop[i]) continue;
    const int li = (int)i - rowbase[i];
    for (int p = rp[i] + lane; p < rp[i + 1]; p += G) {
      const int64_t q = nnz_a + atomicAdd(&cursor[acg[p]], 1);
      bcol[q] = li;
      bval[q] = bval[p];
    }
  }
}
void launch_batch_tscatter(int32_t N, const int32_t* rp, const int32_t* acg, const int32_t* rowbase,
                           const uint8_t* rowop, int32_t* cursor, int64_t nnz_a, int32_t* bcol, double* bval,
                           hipStream_t s) {
  if (N <= 0) return;
  int64_t blocks = ((int64_t)N * 16 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_batch_tscatter, dim3((unsigned)blocks), dim3(256), 0, s, N, rp, acg, rowbase, rowop, cursor,
                     nnz_a, bcol, bval);
}

// Transposes of the A*A^T matrices, one workgroup per matrix, for batches
// whose A*A^T matrices have at most kBatchTBlockN rows: column counts in LDS
// (LDS atomics, where the global-counter form serialises on the columns a
// band's neighbouring rows share), their scan gives the transposed row
// pointers (tp, relative to the transposed region; matrix b's region starts
// at toff[b], the A*A^T entries before it), and entries are placed through
// LDS cursors, walked in storage order kU at a time per thread with their
// local rows from rid.  A*A matrices only write their empty transposed
// rows.  Rows come out unordered within a transposed row (the per-row sort
// follows).
__global__ void __launch_bounds__(1024) k_batch_tblock(int32_t K, const int32_t* __restrict__ blk_row,
                                                       const int64_t* __restrict__ blk_nnz,
                                                       const uint8_t* __restrict__ blk_op,
                                                       const int32_t* __restrict__ toff,
                                                       const int32_t* __restrict__ rid,
                                                       const int32_t* __restrict__ col,
                                                       const double* __restrict__ val, int32_t* __restrict__ tp,
                                                       int32_t* __restrict__ bcol, double* __restrict__ bval) {
  constexpr int kU = 4;
  extern __shared__ int32_t cnt[];
  __shared__ int64_t red[17];
  const int tid = threadIdx.x, T = blockDim.x;
  for (int b = blockIdx.x; b < K; b += gridDim.x) {
    const int r0 = blk_row[b], r1 = blk_row[b + 1], nb = r1 - r0;
    const int base = toff[b];
    if (!blk_op[b]) {
      for (int t = tid; t < nb; t += T) tp[r0 + t] = base;
      continue;
    }
    for (int t = tid; t < nb; t += T) cnt[t] = 0;
    __syncthreads();
    const int64_t e0 = blk_nnz[b], e1 = blk_nnz[b + 1];
    for (int64_t p0 = e0 + tid; p0 < e1; p0 += (int64_t)T * kU) {
      int c[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) c[u] = p0 + (int64_t)u * T < e1 ? col[p0 + (int64_t)u * T] : -1;
#pragma unroll
      for (int u = 0; u < kU; u++)
        if (c[u] >= 0) atomicAdd(&cnt[c[u]], 1);
    }
    __syncthreads();
    // exclusive scan of the counts: a contiguous chunk per thread
    const int chunk = (nb + T - 1) / T;
    const int k0 = tid * chunk, k1 = (k0 + chunk) < nb ? (k0 + chunk) : nb;
    int64_t local = 0;
    for (int k = k0; k < k1; k++) local += cnt[k];
    int64_t tot;
    int64_t run = block_scan_excl<int64_t>(local, red, &tot);
    for (int k = k0; k < k1; k++) {
      const int c = cnt[k];
      cnt[k] = (int32_t)run;
      tp[r0 + k] = base + (int32_t)run;
      run += c;
    }
    __syncthreads();
    // entries to their transposed rows
    for (int64_t p0 = e0 + tid; p0 < e1; p0 += (int64_t)T * kU) {
      int c[kU], li[kU];
      double v[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const int64_t p = p0 + (int64_t)u * T;
        const bool ok = p < e1;
        c[u] = ok ? col[p] : -1;
        li[u] = ok ? rid[p] : 0;
        v[u] = ok ? val[p] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        if (c[u] < 0) continue;
        const int q = base + atomicAdd(&cnt[c[u]], 1);
        bcol[q] = li[u];
        bval[q] = v[u];
      }
    }
    __syncthreads();
  }
}
void launch_batch_tblock(int32_t K, int32_t N, int32_t max_nt, int64_t nnz_t, const int32_t* blk_row,
                         const int64_t* blk_nnz, const uint8_t* blk_op, const int32_t* toff, const int32_t* rid,
                         const int32_t* col, const double* val, int32_t* tp, int32_t* bcol, double* bval,
                         hipStream_t s) {
  if (K <= 0) return;
  const size_t lds = sizeof(int32_t) * ((size_t)max_nt + 1);
  set_dyn_lds((const void*)k_batch_tblock, lds);
  hipLaunchKernelGGL(k_batch_tblock, dim3((unsigned)K), dim3(1024), lds, s, K, blk_row, blk_nnz, blk_op, toff, rid,
                     col, val, tp, bcol, bval);
  launch_fill_i32(tp + N, 1, (int32_t)nnz_t, s);
}

// Per-matrix row pointers of the block-diagonal C: matrix b's n_b + 1
// entries start at crp[blk_row[b] + b]; blkoff[b] = first C entry of b.
__global__ void k_batch_local_rowptr(int32_t N, int32_t K, const int64_t* __restrict__ rowoff,
                                     const int32_t* __restrict__ rowbase, const int32_t* __restrict__ rowblk,
                                     const int32_t* __restrict__ blk_row, int32_t* __restrict__ crp,
                                     int64_t* __restrict__ blkoff) {
  const int64_t total = (int64_t)N + K + 1;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    if (t < N) {
      crp[t + rowblk[t]] = (int32_t)(rowoff[t] - rowoff[rowbase[t]]);
    } else if (t < (int64_t)N + K) {
      const int b = (int)(t - N);
      const int r0 = blk_row[b], r1 = blk_row[b + 1];
      crp[r1 + b] = (int32_t)(rowoff[r1] - rowoff[r0]);
      blkoff[b] = rowoff[r0];
    } else {
      blkoff[K] = rowoff[N];
    }
  }
}
void launch_batch_local_rowptr(int32_t N, int32_t K, const int64_t* rowoff, const int32_t* rowbase,
                               const int32_t* rowblk, const int32_t* blk_row, int32_t* crp, int64_t* blkoff,
                               hipStream_t s) {
  const int64_t total = (int64_t)N + K + 1;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_batch_local_rowptr, dim3((unsigned)blocks), dim3(256), 0, s, N, K, rowoff, rowbase, rowblk,
                     blk_row, crp, blkoff);
}

// out[b] = sum of rowf over matrix b's rows (its product count).
__global__ void k_batch_block_sum(int32_t K, const int32_t* __restrict__ blk_row, const int64_t* __restrict__ rowf,
                                  int64_t* __restrict__ out) {
  __shared__ int64_t red[17];
  for (int b = blockIdx.x; b < K; b += gridDim.x) {
    int64_t v = 0;
    for (int r = blk_row[b] + threadIdx.x; r < blk_row[b + 1]; r += blockDim.x) v += rowf[r];
    int64_t tot;
    (void)block_scan_excl<int64_t>(v, red, &tot);
    if (threadIdx.x == 0) out[b] = tot;
  }
}
void launch_batch_block_sum(int32_t K, const int32_t* blk_row, const int64_t* rowf, int64_t* out, hipStream_t s) {
  if (K <= 0) return;
  const unsigned blocks = (unsigned)(K < 4096 ? K : 4096);
  hipLaunchKernelGGL(k_batch_block_sum, dim3(blocks), dim3(256), 0, s, K, blk_row, rowf, out);
}

// Copy K separate device matrices into the packed buffers, one block per
// matrix at a time: rows get global entry offsets, entries are copied as is.
__global__ void k_batch_gather(int32_t K, const BatchSrc* __restrict__ src, const int32_t* __restrict__ blk_row,
                               const int64_t* __restrict__ blk_nnz, int32_t* __restrict__ rp,
                               int32_t* __restrict__ col, double* __restrict__ val) {
  for (int b = blockIdx.x; b < K; b += gridDim.x) {
    const BatchSrc m = src[b];
    const int32_t r0 = blk_row[b];
    const int64_t e0 = blk_nnz[b];
    for (int r = threadIdx.x; r < m.n; r += blockDim.x) rp[r0 + r] = (int32_t)(e0 + m.rowptr[r]);
    for (int64_t q = threadIdx.x; q < m.nnz; q += blockDim.x) {
      col[e0 + q] = m.col[q];
      val[e0 + q] = m.val[q];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) rp[blk_row[K]] = (int32_t)blk_nnz[K];
}
void launch_batch_gather(int32_t K, const BatchSrc* src, const int32_t* blk_row, const int64_t* blk_nnz, int32_t* rp,
                         int32_t* col, double* val, hipStream_t s) {
  if (K <= 0) return;
  const unsigned blocks = (unsigned)(K < 4096 ? K : 4096);
  hipLaunchKernelGGL(k_batch_gather, dim3(blocks), dim3(256), 0, s, K, src, blk_row, blk_nnz, rp, col, val);
}

}  // namespace appt
