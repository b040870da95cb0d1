// descriptors.hip -- structural descriptors on the GPU.
//
// Restates the host descriptor pass of bench/dispatch.cu:26-37 (deg_cv with
// ddof=0, fill_aa = sum_p deg(col_p)/nnz, fill_aat = sum_k colcount_k^2/nnz),
// the psym estimate of src/fetch_and_prep.py:42-46 (stored pattern), the
// decision rule pick_ours of bench/dispatch.cu:14-17, and adds bandwidth,
// products and the number of row bins the engine would launch.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.hpp"
#include "engine.hpp"

namespace appt {

static unsigned long long* as_ull(void* p) { return static_cast<unsigned long long*>(p); }

// acc[0] = sum d^2, acc[1] = max d, acc[2] = bandwidth, acc[3] = rows not
// strictly ascending, acc[4] = sum_p deg(col_p), acc[5] = max row products
__global__ void k_desc_rows(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                            unsigned long long* __restrict__ acc) {
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
      const unsigned long long d = (unsigned long long)(e - s);
      sd2 += d * d;
      mx = d > mx ? d : mx;
      bw = b > bw ? b : bw;
      uns += u;
      prod += f;
      pmax = f > pmax ? f : pmax;
    }
  }
  if (lane == 0) {
    if (sd2) atomicAdd(&acc[0], sd2);
    if (mx) atomicMax(&acc[1], mx);
    if (bw) atomicMax(&acc[2], bw);
    if (uns) atomicAdd(&acc[3], uns);
    if (prod) atomicAdd(&acc[4], prod);
    if (pmax) atomicMax(&acc[5], pmax);
  }
}

__global__ void k_sumsq(int32_t n, const int32_t* __restrict__ cnt, unsigned long long* __restrict__ out) {
  unsigned long long s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long c = (unsigned long long)cnt[i];
    s += c * c;
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
  if (threadIdx.x % 64 == 0 && s) atomicAdd(out, s);
}

// Count entries (i,j) whose transpose (j,i) is stored: binary search in row j
// when every row is strictly ascending, linear scan otherwise.
__global__ void k_psym(int32_t n, const int32_t* __restrict__ ap, const int32_t* __restrict__ ac, int sorted,
                       unsigned long long* __restrict__ out) {
  constexpr int G = 16;
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  unsigned long long hits = 0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * gpb) {
    for (int p = ap[i] + lane; p < ap[i + 1]; p += G) {
      const int j = ac[p];
      int lo = ap[j], hi = ap[j + 1];
      bool found = false;
      if (sorted) {
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          const int v = ac[mid];
          if (v == (int)i) { found = true; break; }
          if (v < (int)i) lo = mid + 1; else hi = mid;
        }
      } else {
        for (int q = lo; q < hi; q++) if (ac[q] == (int)i) { found = true; break; }
      }
      hits += found ? 1 : 0;
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) hits += __shfl_xor(hits, m, 64);
  if (threadIdx.x % 64 == 0 && hits) atomicAdd(out, hits);
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
  unsigned long long* acc = as_ull(ensure(ws_.stats, sizeof(unsigned long long) * 16));
  int32_t* cnt = (int32_t*)ensure(ws_.cursor, sizeof(int32_t) * ((size_t)n + 1));
  launch_fill_i32((int32_t*)acc, 32, 0, s);
  launch_fill_i32(cnt, n, 0, s);
  int64_t blocks = std::min<int64_t>(((int64_t)n * 16 + 255) / 256, 4096);
  hipLaunchKernelGGL(k_desc_rows, dim3((unsigned)blocks), dim3(256), 0, s, n, A.rowptr, A.col, acc);
  launch_colcount(nnz, A.col, cnt, s);
  int64_t b2 = std::min<int64_t>(((int64_t)n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_sumsq, dim3((unsigned)b2), dim3(256), 0, s, n, cnt, acc + 6);
  unsigned long long hacc[8];
  APPT_HIP(hipMemcpyAsync(hacc, acc, sizeof(hacc), hipMemcpyDeviceToHost, s));
  APPT_HIP(hipStreamSynchronize(s));
  const int sorted = hacc[3] == 0 ? 1 : 0;
  hipLaunchKernelGGL(k_psym, dim3((unsigned)blocks), dim3(256), 0, s, n, A.rowptr, A.col, sorted, acc + 7);
  APPT_HIP(hipMemcpyAsync(&hacc[7], acc + 7, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  APPT_HIP(hipStreamSynchronize(s));

  const double dmean = (double)nnz / n;
  const double var = (double)hacc[0] / n - dmean * dmean;
  out->deg_mean = dmean;
  out->deg_cv = std::sqrt(var > 0 ? var : 0.0) / (dmean > 1e-9 ? dmean : 1e-9);
  out->deg_max = (int32_t)hacc[1];
  out->bandwidth = (int32_t)hacc[2];
  out->products_aa = (int64_t)hacc[4];
  out->max_row_products_aa = (int32_t)std::min<unsigned long long>(hacc[5], (unsigned long long)INT32_MAX);
  out->products_aat = (int64_t)hacc[6];
  out->fill_aa = nnz > 0 ? (double)hacc[4] / nnz : 0.0;
  out->fill_aat = nnz > 0 ? (double)hacc[6] / nnz : 0.0;
  out->psym = nnz > 0 ? (double)hacc[7] / (double)nnz : 0.0;
  const double fill = op == APPT_OP_AA ? out->fill_aa : out->fill_aat;
  out->pick_ours = pick_ours_rule(out->deg_cv, fill, nnz) ? 1 : 0;
  Operands O;
  Plan P;
  stage_operands(ws_, A, op, O);
  stage_prepass(ws_, O, cfg_);
  read_sym(ws_, P);
  out->plan_bins = P.bins_used;
}

}  // namespace appt
