// engine.cpp -- orchestration of one SpGEMM call on one HIP stream.
//
// Replaces run_ours() of the reference (src/spgemm_engine.cuh:132-210):
//   reference                              here
//   build_At (:140-146)                    colcount + scan + scatter (AAt/AAtS)
//   k_flops + k_caps + arena scan (:156-162) rowstats + prepass (products, span,
//                                          bin) + bin partition
//   k_symbolic (:167)                      one launch per non-empty bin
//   exclusive_scan -> Coff (:168)          scan to int64 offsets (+ int32 copy)
//   k_numeric (:189)                       one launch per non-empty bin, rows
//                                          written sorted straight into CSR
//   k_mirror + radix sort (:190-196)       full-row structure, values for j >= i,
//                                          lower entries pulled from the upper part
//   cub::DeviceRadixSort (:195)            not needed: rows come out sorted
// Host synchronisations: 2 (bin sizes; nnz(C)).
#include "engine.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace appt {

// ---------------------------------------------------------------------------
// Pool
// ---------------------------------------------------------------------------
Pool::~Pool() { trim(); for (auto& kv : size_of_) (void)hipFree(kv.first); }

void* Pool::get(size_t bytes) {
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~(size_t)255;
  auto it = free_.lower_bound(bytes);
  if (it != free_.end() && it->first <= 2 * bytes) {
    void* p = it->second;
    free_.erase(it);
    return p;
  }
  if (budget_ > 0 && allocated_ + (int64_t)bytes > budget_) {
    trim();
    if (allocated_ + (int64_t)bytes > budget_)
      throw Error(APPT_ERR_OOM, "engine HBM budget exceeded: need " + std::to_string(bytes) + " B, " +
                                    std::to_string(allocated_) + " of " + std::to_string(budget_) + " B in use");
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(APPT_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + ") failed");
    }
  }
  size_of_[p] = bytes;
  allocated_ += (int64_t)bytes;
  return p;
}

void Pool::put(void* p) {
  if (!p) return;
  auto it = size_of_.find(p);
  if (it == size_of_.end()) return;
  free_.emplace(it->second, p);
}

void Pool::trim() {
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    allocated_ -= (int64_t)size_of_[kv.second];
    size_of_.erase(kv.second);
  }
  free_.clear();
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
static BinParams default_params() {
  BinParams P{};
  P.ncls = 3;
  P.cls_max[0] = 128; P.cls_max[1] = 512; P.cls_max[2] = 2048;
  P.cls_bytes[0] = 16 * 1024; P.cls_bytes[1] = 24 * 1024; P.cls_bytes[2] = 32 * 1024;
  P.glen_max[0] = 8; P.glen_max[1] = 16; P.glen_max[2] = 32;
  P.dense_ratio = 4;
  return P;
}

Engine::Engine(const appt_options* opt) : pool_(0) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    throw Error(APPT_ERR_NO_DEVICE, "no HIP device visible");
  }
  device_ = opt ? opt->device : 0;
  if (device_ < 0 || device_ >= ndev) throw Error(APPT_ERR_INVALID, "device ordinal out of range");
  APPT_HIP(hipSetDevice(device_));
  if (opt && opt->stream) {
    stream_ = (hipStream_t)opt->stream;
  } else {
    APPT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  int64_t budget = opt ? opt->hbm_budget_bytes : 0;
  if (budget <= 0) {
    size_t fr = 0, tot = 0;
    APPT_HIP(hipMemGetInfo(&fr, &tot));
    budget = (int64_t)(0.9 * (double)fr);
  }
  pool_.set_budget(budget);
  int lds = 0;
  APPT_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device_));
  lds_max_ = lds > 0 ? lds : 65536;
  int ncu = 0, lds_cu = 0;
  APPT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device_));
  APPT_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device_));
  set_num_cus(ncu);
  set_lds_per_cu(lds_cu > 0 ? lds_cu : lds_max_);
  APPT_HIP(hipHostMalloc((void**)&hs_, sizeof(HostStats), hipHostMallocDefault));
  for (auto& e : ev_) APPT_HIP(hipEventCreate(&e));

  sym_params = default_params();
  num_params = default_params();
  int64_t group_budget = 64 * 1024;
  if (opt) {
    const appt_tuning& t = opt->tuning;
    int nc = 0;
    while (nc < APPT_MAX_CLASSES && nc < kNumCls && t.f_class_max[nc] > 0) nc++;
    if (nc > 0) {
      sym_params.ncls = num_params.ncls = nc;
      for (int c = 0; c < nc; c++) { sym_params.cls_max[c] = t.f_class_max[c]; num_params.cls_max[c] = t.f_class_max[c]; }
    }
    if (t.lds_budget_bytes > 0) group_budget = t.lds_budget_bytes;
    if (t.bitmap_words_factor > 0) num_params.dense_ratio = t.bitmap_words_factor;
    for (int g = 0; g < 3; g++)
      if (t.group_for_len[g] > 0) sym_params.glen_max[g] = num_params.glen_max[g] = t.group_for_len[g];
  }
  if (group_budget > lds_max_) group_budget = lds_max_;
  set_lds_budget(group_budget);
  // workgroup-per-row bins may use the device's LDS per block minus static LDS
  const int32_t wg = (int32_t)(lds_max_ - 8192);
  sym_params.wg_bytes = num_params.wg_bytes = wg;
  if (opt && opt->tuning.wg_lds_bytes > 0) {
    // cap the workgroup bins (rows above go to the heavy-row bin)
    sym_params.wg_bytes = num_params.wg_bytes = std::min<int32_t>(wg, opt->tuning.wg_lds_bytes);
  }
  if (opt && opt->tuning.heavy_window_cols > 0) heavy_cap_ = opt->tuning.heavy_window_cols;
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  Scratch* all[] = {&ws_.bptr,   &ws_.bcol,     &ws_.bval,    &ws_.cursor, &ws_.bmin,   &ws_.bmax,
                    &ws_.rowf,   &ws_.rowlo,    &ws_.rowhi,   &ws_.symbin, &ws_.numbin, &ws_.rowcnt,
                    &ws_.rowoff, &ws_.symrows,  &ws_.numrows, &ws_.scan_tmp, &ws_.stats, &ws_.rowwords,
                    &ws_.wlist,  &ws_.statpart, &ws_.ctl,     &ws_.cur};
  for (Scratch* s : all) free_scratch(*s);
  pool_.trim();
  for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
  if (hs_) (void)hipHostFree(hs_);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void* Engine::ensure(Scratch& s, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (s.cap >= bytes) return s.p;
  if (s.p) pool_.put(s.p);
  size_t want = bytes + bytes / 8;
  s.p = pool_.get(want);
  s.cap = (want + 255) & ~(size_t)255;
  return s.p;
}

void Engine::free_scratch(Scratch& s) {
  if (s.p) pool_.put(s.p);
  s.p = nullptr;
  s.cap = 0;
}

void Engine::record(int idx) {
  if (timing_) APPT_HIP(hipEventRecord(ev_[idx], stream_));
}

void Engine::validate(const appt_csr_view& A) {
  if (A.n < 0) throw Error(APPT_ERR_INVALID, "n < 0");
  if (A.nnz < 0 || A.nnz > INT32_MAX) throw Error(APPT_ERR_INVALID, "nnz outside [0, 2^31)");
  if (A.n > 0 && !A.rowptr) throw Error(APPT_ERR_INVALID, "rowptr is NULL");
  if (A.nnz > 0 && (!A.col || !A.val)) throw Error(APPT_ERR_INVALID, "col/val is NULL");
}

template <typename T>
static T* as(void* p) { return static_cast<T*>(p); }

void Engine::prepass_stage(const appt_csr_view& A, appt_op op, const int32_t*& bp, const int32_t*& bc,
                           const double*& bv) {
  const int32_t n = A.n;
  const int64_t nnz = A.nnz;
  hipStream_t s = stream_;
  bp = A.rowptr;
  bc = A.col;
  bv = A.val;
  if (op != APPT_OP_AA) {
    int32_t* tp = as<int32_t>(ensure(ws_.bptr, sizeof(int32_t) * ((size_t)n + 1)));
    int32_t* tc = as<int32_t>(ensure(ws_.bcol, sizeof(int32_t) * (size_t)nnz));
    double* tv = as<double>(ensure(ws_.bval, sizeof(double) * (size_t)nnz));
    int32_t* cnt = as<int32_t>(ensure(ws_.cursor, sizeof(int32_t) * ((size_t)n + 1)));
    int64_t* off64 = as<int64_t>(ensure(ws_.rowoff, sizeof(int64_t) * ((size_t)n + 1)));
    int64_t* tmp = as<int64_t>(ensure(ws_.scan_tmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max(n, 1))));
    launch_fill_i32(cnt, n, 0, s);
    launch_colcount(nnz, A.col, cnt, s);
    launch_scan(cnt, n, off64, tp, tmp, s);
    if (n > 0) APPT_HIP(hipMemcpyAsync(cnt, tp, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    launch_transpose_scatter(n, A.rowptr, A.col, A.val, cnt, tc, tv, s);
    bp = tp; bc = tc; bv = tv;
  }
  record(1);
  int4* bmeta = as<int4>(ensure(ws_.bmin, sizeof(int4) * (size_t)std::max(n, 1)));
  int64_t* rowf = as<int64_t>(ensure(ws_.rowf, sizeof(int64_t) * (size_t)std::max(n, 1)));
  int32_t* rowlo = as<int32_t>(ensure(ws_.rowlo, sizeof(int32_t) * (size_t)std::max(n, 1)));
  int32_t* rowhi = as<int32_t>(ensure(ws_.rowhi, sizeof(int32_t) * (size_t)std::max(n, 1)));
  uint8_t* symbin = as<uint8_t>(ensure(ws_.symbin, (size_t)std::max(n, 1)));
  // stats layout: [sym stats | num stats | products(u64) | cursors]
  const size_t stat_words = (size_t)kNumBins * kStatFields;
  int32_t* dstats = as<int32_t>(ensure(ws_.stats, sizeof(int32_t) * (stat_words * 2 + 4 + kNumBins)));
  unsigned long long* dprod = reinterpret_cast<unsigned long long*>(dstats + stat_words * 2);
  launch_fill_i32(dstats, (int64_t)(stat_words * 2 + 4 + kNumBins), 0, s);
  const int upper = op == APPT_OP_AATS ? 1 : 0;
  // control words: [0] / [1] heavy-row claim counters, [2] B-unsorted flag
  int32_t* ctl = as<int32_t>(ensure(ws_.ctl, sizeof(int32_t) * 4));
  launch_fill_i32(ctl, 4, 0, s);
  launch_rowstats(n, bp, bc, bmeta, ctl + 2, s);
  int32_t* part = as<int32_t>(ensure(ws_.statpart, sizeof(int32_t) * (size_t)stats_part_elems() +
                                                      sizeof(unsigned long long) * kStatsMaxBlocks));
  launch_prepass(n, A.rowptr, A.col, bmeta, upper, sym_params, rowf, rowlo, rowhi, symbin, dstats, dprod, part, s);
  APPT_HIP(hipMemcpyAsync(hs_->sym, dstats, sizeof(int32_t) * stat_words, hipMemcpyDeviceToHost, s));
  APPT_HIP(hipMemcpyAsync(&hs_->products, dprod, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  APPT_HIP(hipStreamSynchronize(s));
}

int Engine::plan_bins(const appt_csr_view& A, appt_op op) {
  const int32_t* bp; const int32_t* bc; const double* bv;
  prepass_stage(A, op, bp, bc, bv);
  int used = 0;
  for (int b = 1; b < kNumBins; b++) used += hs_->sym[b * kStatFields] > 0 ? 1 : 0;
  return used;
}

void Engine::compute(const appt_csr_view& A, appt_op op, appt_csr_result* C) {
  validate(A);
  if (!C) throw Error(APPT_ERR_INVALID, "result pointer is NULL");
  if (op != APPT_OP_AA && op != APPT_OP_AAT && op != APPT_OP_AATS) throw Error(APPT_ERR_INVALID, "bad op");
  APPT_HIP(hipSetDevice(device_));
  std::memset(C, 0, sizeof(*C));
  times_ = appt_phase_times{};
  const int32_t n = A.n;
  hipStream_t s = stream_;
  int syncs = 0;
  record(0);

  const int32_t* bp = nullptr;
  const int32_t* bc = nullptr;
  const double* bv = nullptr;
  prepass_stage(A, op, bp, bc, bv);
  syncs++;
  const int upper = op == APPT_OP_AATS ? 1 : 0;
  int32_t* rowcnt = as<int32_t>(ensure(ws_.rowcnt, sizeof(int32_t) * (size_t)std::max(n, 1)));
  int32_t* rowwords = as<int32_t>(ensure(ws_.rowwords, sizeof(int32_t) * (size_t)std::max(n, 1)));
  int64_t* rowoff = as<int64_t>(ensure(ws_.rowoff, sizeof(int64_t) * ((size_t)n + 1)));
  int32_t* symrows = as<int32_t>(ensure(ws_.symrows, sizeof(int32_t) * (size_t)std::max(n, 1)));
  int32_t* numrows = as<int32_t>(ensure(ws_.numrows, sizeof(int32_t) * (size_t)std::max(n, 1)));
  int64_t* tmp = as<int64_t>(ensure(ws_.scan_tmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max(n, 1))));
  const int64_t* rowf = as<int64_t>(ws_.rowf.p);
  const int32_t* rowlo = as<int32_t>(ws_.rowlo.p);
  const int32_t* rowhi = as<int32_t>(ws_.rowhi.p);
  const uint8_t* symbin = as<uint8_t>(ws_.symbin.p);
  const int4* bmeta = as<int4>(ws_.bmin.p);
  uint2* wlist = as<uint2>(ensure(ws_.wlist, sizeof(uint2) * kWordListCap * (size_t)std::max(n, 1)));
  uint8_t* numbin = as<uint8_t>(ensure(ws_.numbin, (size_t)std::max(n, 1)));
  const size_t stat_words = (size_t)kNumBins * kStatFields;
  int32_t* dstats = as<int32_t>(ws_.stats.p);
  int32_t* dnum = dstats + stat_words;
  int32_t* dcursor = dstats + stat_words * 2 + 4;
  const int64_t products = (int64_t)hs_->products;
  int32_t symoff[kNumBins + 1];
  {
    int32_t run = 0;
    for (int b = 0; b < kNumBins; b++) {
      symoff[b] = run;
      hs_->cursor[b] = run;
      if (b > 0) run += hs_->sym[b * kStatFields];
    }
    symoff[kNumBins] = run;
  }
  APPT_HIP(hipMemcpyAsync(dcursor, hs_->cursor, sizeof(int32_t) * kNumBins, hipMemcpyHostToDevice, s));
  launch_bin_scatter(n, symbin, dcursor, symrows, s);
  launch_fill_i32(rowcnt, n, 0, s);
  launch_fill_i32(rowwords, n, 0, s);
  record(2);

  // heavy-row bin: LDS column windows, a cursor per A entry
  HeavyArgs heavy{};
  auto setup_heavy = [&]() {
    if (heavy.work) return;
    // dynamic LDS: the block's LDS less the kernel's static staging (~21 KB)
    const int64_t dyn = lds_max_ - 24 * 1024;
    heavy.wc_sym = heavy_window_cols(false, dyn, heavy_cap_);
    heavy.wc_num = heavy_window_cols(true, dyn, heavy_cap_);
    heavy.work = as<int32_t>(ws_.ctl.p);
    heavy.bflag = heavy.work + 2;
    heavy.cur = as<int32_t>(ensure(ws_.cur, sizeof(int32_t) * (size_t)std::max<int64_t>(A.nnz, 1)));
  };
  if (hs_->sym[kBinHeavy * kStatFields] > 0) setup_heavy();

  // ---- symbolic -------------------------------------------------------------
  int nbins_used = 0;
  for (int b = 1; b < kNumBins; b++) {
    const int32_t cnt = hs_->sym[b * kStatFields];
    if (cnt == 0) continue;
    nbins_used++;
    const int32_t mmax = hs_->sym[b * kStatFields + 1];
    const int32_t wmax = hs_->sym[b * kStatFields + 2];
    RowArgs ra{symrows + symoff[b], cnt, A.rowptr, A.col, A.val, bp, bc, bv, rowf, rowlo, rowhi, rowwords, 0, A.nnz, bmeta, wlist};
    const int64_t rb = symbolic_row_lds_bytes(b, mmax, wmax);
    if (rb > lds_max_ - 8192) throw Error(APPT_ERR_UNSUPPORTED, "symbolic row exceeds LDS");
    launch_symbolic_bin(b, ra, wmax, rowcnt, heavy, s);
  }
  record(3);

  // ---- row offsets, numeric binning ----------------------------------------
  int32_t* rp32 = as<int32_t>(pool_.get(sizeof(int32_t) * ((size_t)n + 1)));
  launch_scan(rowcnt, n, rowoff, rp32, tmp, s);
  launch_numbin(n, A.rowptr, rowf, rowcnt, rowwords, rowlo, rowhi, num_params, numbin, dnum,
                as<int32_t>(ws_.statpart.p), s);
  APPT_HIP(hipMemcpyAsync(hs_->num, dnum, sizeof(int32_t) * stat_words, hipMemcpyDeviceToHost, s));
  APPT_HIP(hipMemcpyAsync(&hs_->total, rowoff + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  APPT_HIP(hipStreamSynchronize(s));
  syncs++;
  const int64_t unnz = hs_->total;
  int32_t numoff[kNumBins + 1];
  {
    int32_t run = 0;
    for (int b = 0; b < kNumBins; b++) {
      numoff[b] = run;
      hs_->cursor[b] = run;
      if (b > 0) run += hs_->num[b * kStatFields];
    }
    numoff[kNumBins] = run;
  }
  APPT_HIP(hipMemcpyAsync(dcursor, hs_->cursor, sizeof(int32_t) * kNumBins, hipMemcpyHostToDevice, s));
  launch_bin_scatter(n, numbin, dcursor, numrows, s);
  record(4);

  // output buffers (the symbolic pass sized C exactly)
  int32_t* ccol = as<int32_t>(pool_.get(sizeof(int32_t) * (size_t)std::max<int64_t>(unnz, 1)));
  double* cval = as<double>(pool_.get(sizeof(double) * (size_t)std::max<int64_t>(unnz, 1)));

  // ---- numeric --------------------------------------------------------------
  // AAtS: values are formed for j >= i only; the structure is the full row
  for (int b = 1; b < kNumBins; b++) {
    const int32_t cnt = hs_->num[b * kStatFields];
    if (cnt == 0) continue;
    const int32_t cmax = hs_->num[b * kStatFields + 1];
    const int32_t wmax = hs_->num[b * kStatFields + 2];
    RowArgs ra{numrows + numoff[b], cnt, A.rowptr, A.col, A.val, bp, bc, bv, rowf, rowlo, rowhi, rowwords, upper, A.nnz, bmeta, wlist};
    const int64_t rb = numeric_row_lds_bytes(b, cmax, wmax);
    if (rb > lds_max_ - 8192) throw Error(APPT_ERR_UNSUPPORTED, "numeric row exceeds LDS");
    if (b == kBinHeavy) setup_heavy();   // rows may become heavy only in the numeric binning
    launch_numeric_bin(b, ra, wmax, cmax, rowoff, ccol, cval, heavy, s);
  }
  record(5);

  // ---- AAtS: mirror the upper triangle into the lower entries -------------
  if (upper) launch_mirror_pull(n, rowoff, ccol, cval, s);
  const int64_t cnnz = unnz;
  void* rowptr_out = rp32;
  if (cnnz > INT32_MAX) {
    pool_.put(rp32);
    int64_t* rp64 = as<int64_t>(pool_.get(sizeof(int64_t) * ((size_t)n + 1)));
    launch_copy_i64(rowoff, (int64_t)n + 1, rp64, s);
    rowptr_out = rp64;
  }
  record(6);
  APPT_HIP(hipGetLastError());
  APPT_HIP(hipStreamSynchronize(s));

  Result* R = new Result{rowptr_out, ccol, cval};
  C->n = n;
  C->nnz = cnnz;
  C->rowptr_is_64 = cnnz > INT32_MAX ? 1 : 0;
  C->rowptr = rowptr_out;
  C->col = ccol;
  C->val = cval;
  C->_priv = R;

  times_.products = products;
  times_.nnz_c = cnnz;
  times_.n_bins_used = nbins_used;
  times_.sync_points = syncs + 1;
  if (timing_) {
    float ms[6];
    for (int k = 0; k < 6; k++) APPT_HIP(hipEventElapsedTime(&ms[k], ev_[k], ev_[k + 1]));
    times_.transpose_ms = ms[0];
    times_.prepass_ms = ms[1];
    times_.symbolic_ms = ms[2];
    times_.scan_ms = ms[3];
    times_.numeric_ms = ms[4];
    times_.mirror_ms = ms[5];
    APPT_HIP(hipEventElapsedTime(&times_.total_ms, ev_[0], ev_[6]));
  }
}

void Engine::release(appt_csr_result* C) {
  if (!C || !C->_priv) return;
  Result* R = static_cast<Result*>(C->_priv);
  pool_.put(R->rowptr);
  pool_.put(R->col);
  pool_.put(R->val);
  delete R;
  std::memset(C, 0, sizeof(*C));
}

void Engine::run_timed(const char* task, const appt_csr_view& A, int reps, appt_eng_result* out) {
  appt_op op;
  if (appt_op_from_task(task, &op) != APPT_OK) throw Error(APPT_ERR_INVALID, std::string("bad task ") + (task ? task : "(null)"));
  if (reps < 0) throw Error(APPT_ERR_INVALID, "reps < 0");
  APPT_HIP(hipSetDevice(device_));
  // checked call: checksums over the row-major sorted values (host, in order)
  appt_csr_result C;
  compute(A, op, &C);
  std::vector<double> hv((size_t)C.nnz);
  if (C.nnz) APPT_HIP(hipMemcpy(hv.data(), C.val, sizeof(double) * (size_t)C.nnz, hipMemcpyDeviceToHost));
  double sm = 0, as_ = 0, sq = 0;
  for (double x : hv) { sm += x; as_ += x < 0 ? -x : x; sq += x * x; }
  const int64_t nnz_c = C.nnz;
  release(&C);
  APPT_HIP(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  APPT_HIP(hipEventCreate(&e0));
  APPT_HIP(hipEventCreate(&e1));
  APPT_HIP(hipEventRecord(e0, stream_));
  for (int r = 0; r < reps; r++) {
    compute(A, op, &C);
    release(&C);
  }
  APPT_HIP(hipEventRecord(e1, stream_));
  APPT_HIP(hipEventSynchronize(e1));
  float ms = 0;
  APPT_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  out->ms = reps > 0 ? ms / (float)reps : 0.f;
  out->nnz = nnz_c;
  out->s = sm;
  out->as = as_;
  out->sq = sq;
}

}  // namespace appt
