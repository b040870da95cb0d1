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
//
// The pipeline is split into stages that only enqueue kernels; compute()
// runs them with the two synchronisations between, and a batch replays them
// inside a HIP graph with the launch shapes of its first run.
#include "engine.hpp"
#include "noalloc.hpp"

#include <chrono>

#include <algorithm>
#include <cstring>
#include <vector>

namespace appt {

// ---------------------------------------------------------------------------
// Pool
// ---------------------------------------------------------------------------
Pool::~Pool() {
  trim_locked();
  for (auto& kv : size_of_) (void)hipFree(kv.first);
  if (tally_) tally_->add(-allocated_.load());
}

void* Pool::get(size_t bytes) {
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = free_.lower_bound(bytes);
  if (it != free_.end() && it->first <= 2 * bytes) {
    void* p = it->second;
    free_.erase(it);
    return p;
  }
  if (budget_ > 0 && allocated_ + (int64_t)bytes > budget_) {
    trim_locked();
    if (allocated_ + (int64_t)bytes > budget_)
      throw Error(APPT_ERR_OOM, "engine HBM budget exceeded: need " + std::to_string(bytes) + " B, " +
                                    std::to_string(allocated_) + " of " + std::to_string(budget_) + " B in use");
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim_locked();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(APPT_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + ") failed");
    }
  }
  size_of_[p] = bytes;
  allocated_ += (int64_t)bytes;
  if (tally_) tally_->add((int64_t)bytes);
  return p;
}

void Pool::set_tally(PoolTally* t) {
  std::lock_guard<std::mutex> lk(mu_);
  if (tally_) tally_->add(-allocated_.load());
  tally_ = t;
  if (tally_) tally_->add(allocated_.load());
}

void Pool::put(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = size_of_.find(p);
  if (it == size_of_.end()) return;
  free_.emplace(it->second, p);
}

void Pool::trim() {
  std::lock_guard<std::mutex> lk(mu_);
  trim_locked();
}

void Pool::trim_locked() {
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    allocated_ -= (int64_t)size_of_[kv.second];
    if (tally_) tally_->add(-(int64_t)size_of_[kv.second]);
    size_of_.erase(kv.second);
  }
  free_.clear();
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
// Scattered rows: the register sort beats a window once the span has this
// many times more words than the row has products (symbolic: by span alone,
// so only far sparser spans qualify).
constexpr int32_t kDefaultSortRatio = 2, kDefaultSymSortRatio = 16;
// Device bytes the symbolic workgroup-bitmap rows may keep for the numeric pass.
constexpr int64_t kWgBitsBudget = (int64_t)1 << 30;

static BinParams default_params() {
  BinParams P{};
  P.ncls = 3;
  P.cls_max[0] = 128; P.cls_max[1] = 512; P.cls_max[2] = 2048;
  P.cls_bytes[0] = 16 * 1024; P.cls_bytes[1] = 24 * 1024; P.cls_bytes[2] = 32 * 1024;
  P.glen_max[0] = 8; P.glen_max[1] = 16; P.glen_max[2] = 32;
  P.dense_ratio = 4;
  P.lds_budget = 64 * 1024;
  P.sort_ratio = kDefaultSortRatio;
  P.sort_max = kSortMax;
  return P;
}

Config make_config(const appt_tuning* t, int64_t lds_max) {
  Config c;
  c.sym = default_params();
  c.num = default_params();
  c.sym.sort_ratio = kDefaultSymSortRatio;
  int64_t group_budget = 64 * 1024;
  // workgroup-per-row bins may use the device's LDS per block minus static LDS
  int32_t wg = (int32_t)(lds_max - 8192);
  if (t) {
    int nc = 0;
    while (nc < APPT_MAX_CLASSES && nc < kNumCls && t->f_class_max[nc] > 0) nc++;
    if (nc > 0) {
      c.sym.ncls = c.num.ncls = nc;
      for (int k = 0; k < nc; k++) c.sym.cls_max[k] = c.num.cls_max[k] = t->f_class_max[k];
    }
    if (t->lds_budget_bytes > 0) group_budget = t->lds_budget_bytes;
    if (t->bitmap_words_factor > 0) c.num.dense_ratio = t->bitmap_words_factor;
    for (int g = 0; g < 3; g++)
      if (t->group_for_len[g] > 0) c.sym.glen_max[g] = c.num.glen_max[g] = t->group_for_len[g];
    if (t->wg_lds_bytes > 0) wg = std::min<int32_t>(wg, t->wg_lds_bytes);
    if (t->heavy_window_cols > 0) c.heavy_cap = t->heavy_window_cols;
    if (t->transpose_radix_min_nnz > 0) c.radix_min_nnz = t->transpose_radix_min_nnz;
    if (t->transpose_radix_min_nnz < 0) c.radix_min_nnz = INT64_MAX;
    if (t->mirror_radix_min_nnz > 0) c.radix_mirror_min_nnz = t->mirror_radix_min_nnz;
    if (t->mirror_radix_min_nnz < 0) c.radix_mirror_min_nnz = INT64_MAX;
    if (t->sort_span_factor != 0) c.num.sort_ratio = t->sort_span_factor > 0 ? t->sort_span_factor : 0;
    if (t->sort_span_factor_symbolic != 0)
      c.sym.sort_ratio = t->sort_span_factor_symbolic > 0 ? t->sort_span_factor_symbolic : 0;
    for (int k = 0; k < kNumCls; k++)
      if (t->class_lds_bytes[k] > 0) c.sym.cls_bytes[k] = c.num.cls_bytes[k] = t->class_lds_bytes[k];
    if (t->sort_max_products != 0) {
      int32_t m = 0;  // a power of two in [16, kSortMax], or no sort bins
      if (t->sort_max_products > 0) {
        m = 16;
        while (m < t->sort_max_products && m < kSortMax) m <<= 1;
      }
      c.sym.sort_max = c.num.sort_max = m;
    }
    if (t->rowptr64_min_nnz > 0) c.rowptr64_min_nnz = t->rowptr64_min_nnz;
    if (t->row_templates < 0) c.templates = false;
    if (t->row_templates > 0) c.templates_min_nnz = t->row_templates;
    c.gram = t->gram;
  }
  if (group_budget > lds_max) group_budget = lds_max;
  c.sym.lds_budget = c.num.lds_budget = (int32_t)group_budget;
  c.sym.wg_bytes = c.num.wg_bytes = wg;
  return c;
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
  // device address of the no-allocation entry point's static scratch, taken
  // once per device while it is current
  APPT_HIP(static_scratch_init(device_));
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
  APPT_HIP(hipEventCreateWithFlags(&ev_wait_, hipEventDisableTiming));
  cfg_ = make_config(opt ? &opt->tuning : nullptr, lds_max_);
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  free_workspace(ws_);
  pool_.trim();
  for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
  if (ev_wait_) (void)hipEventDestroy(ev_wait_);
  if (hs_) (void)hipHostFree(hs_);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Engine::wait_stream() {
  APPT_HIP(hipEventRecord(ev_wait_, stream_));
  const auto t0 = std::chrono::steady_clock::now();
  while (true) {
    const hipError_t e = hipEventQuery(ev_wait_);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) APPT_HIP(e);
    if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(2000)) break;
  }
  APPT_HIP(hipEventSynchronize(ev_wait_));
}

void* Engine::ensure(Scratch& s, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (s.cap >= bytes) return s.p;
  if (arena_) {
    const size_t want = (bytes + 255) & ~(size_t)255;
    if (arena_->used + want > arena_->cap)
      throw Error(APPT_ERR_UNSUPPORTED, "static scratch exhausted: " + std::to_string(arena_->used + want) +
                                            " of " + std::to_string(arena_->cap) + " bytes");
    s.p = arena_->base + arena_->used;
    s.cap = want;
    arena_->used += want;
    return s.p;
  }
  if (s.p) pool_.put(s.p);
  size_t want = bytes + bytes / 8;
  s.p = pool_.get(want);
  s.cap = (want + 255) & ~(size_t)255;
  return s.p;
}

void Engine::free_workspace(Workspace& W) {
  W.each([&](Scratch& s) {
    if (s.p) pool_.put(s.p);
    s.p = nullptr;
    s.cap = 0;
  });
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

int64_t Engine::work_bound(int32_t rows, int32_t nb, int64_t nnz) {
  // per output row: pre-pass, bins, offsets, word list, template ids and
  // int32 row pointers; per B row: B-row metadata, transpose pointers and
  // cursors, template classes; per entry: the transposed column and value,
  // radix keys / rows / values and histograms, heavy-row cursors.  Buffers
  // grow by an eighth beyond their request; fixed: bin statistics and the
  // template tables.
  const int64_t per_row = 700, per_brow = 64, per_entry = 48;
  const int64_t fixed = (int64_t)24 << 20;
  return per_row * ((int64_t)rows + 2) + per_brow * ((int64_t)nb + 2) + per_entry * nnz + fixed;
}

void Engine::trim_all() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  free_workspace(ws_);
  plan_key_ = RowPlanKey{};
  pool_.trim();
}

int64_t result_bytes(int32_t n, int64_t nnz) {
  return (nnz > INT32_MAX ? 8 : 4) * ((int64_t)n + 1) + 12 * nnz;
}

static size_t stat_words() { return (size_t)kNumBins * kStatFields; }
// stats layout: [sym stats | num stats | products (u64) | nnz(C) (i64) | sym cursors | num cursors];
// the first four parts mirror HostStats, so each host read is one copy
static size_t stats_ints() { return stat_words() * 2 + 4 + 2 * kNumBins; }

// ---------------------------------------------------------------------------
// stages
// ---------------------------------------------------------------------------
void Engine::stage_operands(Workspace& W, const appt_csr_view& A, appt_op op, Operands& O) {
  const int32_t n = A.n;
  const int64_t nnz = A.nnz;
  hipStream_t s = stream_;
  O = Operands{};
  O.n = n;
  O.nb = n;
  O.annz = nnz;
  O.ap = A.rowptr;
  O.ac = A.col;
  O.av = A.val;
  O.bc = A.col;
  O.bv = A.val;
  O.bnnz = nnz;
  O.upper = op == APPT_OP_AATS ? 1 : 0;
  O.ncols = n;
  O.a_is_b = op == APPT_OP_AA;
  const int32_t* bptr = A.rowptr;
  int32_t* ctl = reset_ctl(W);
  if (op != APPT_OP_AA) {
    // B = A^T with sorted rows (heavy rows then carry per-entry cursors
    // across windows): a symmetric pattern is gathered in A's layout; any
    // other takes the column histogram, scan, atomic-cursor scatter and
    // per-row sort (all skipped on the device when the gather succeeded)
    int32_t* tp = as<int32_t>(ensure(W.bptr, sizeof(int32_t) * ((size_t)n + 1)));
    int32_t* tc = as<int32_t>(ensure(W.bcol, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
    double* tv = as<double>(ensure(W.bval, sizeof(double) * (size_t)std::max<int64_t>(nnz, 1)));
    int32_t* cnt = as<int32_t>(ensure(W.cursor, sizeof(int32_t) * ((size_t)n + 1)));
    int64_t* off64 = as<int64_t>(ensure(W.rowoff, sizeof(int64_t) * ((size_t)n + 1)));
    int64_t* tmp = as<int64_t>(ensure(W.scan_tmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max(n, 1))));
    int32_t* slist = as<int32_t>(ensure(W.sortlist, sizeof(int32_t) * (size_t)sort_rows_scratch_elems(n)));
    int32_t* asym = ctl + 3;
    RadixScratch rx{};
    const bool radix = nnz >= cfg_.radix_min_nnz;
    if (radix) {
      const size_t nz = (size_t)std::max<int64_t>(nnz, 1);
      const int64_t hn = ((int64_t)1 << radix_bits(n)) * radix_tiles(nnz);
      rx.keys = as<int32_t>(ensure(W.rkeys, sizeof(int32_t) * nz));
      rx.rows = as<int32_t>(ensure(W.rrows, sizeof(int32_t) * nz));
      rx.vals = as<double>(ensure(W.rvals, sizeof(double) * nz));
      rx.keys2 = as<int32_t>(ensure(W.rkeys2, sizeof(int32_t) * nz));
      rx.hist = as<int32_t>(ensure(W.rhist, sizeof(int32_t) * (size_t)std::max<int64_t>(hn, 1)));
      rx.hoff = as<int64_t>(ensure(W.rhoff, sizeof(int64_t) * (size_t)(hn + 1)));
      rx.htmp = as<int64_t>(ensure(W.rtmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max<int64_t>(hn, 1))));
    }
    // the gather's outcome is read back before the general transpose is
    // launched: a synchronisation costs less than the general path's
    // launches (up to 17 for the radix sort), which would otherwise be
    // queued and find their gate closed on the device
    launch_transpose_gather(n, A.rowptr, A.col, A.val, tp, tc, tv, asym, s);
    APPT_HIP(hipMemcpyAsync(&hs_->ctl[3], asym, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    wait_stream();
    if (hs_->ctl[3] != 0)
      launch_transpose_general(n, nnz, A.rowptr, A.col, A.val, tp, tc, tv, cnt, off64, tmp, slist,
                               radix ? &rx : nullptr, asym, s);
    // a symmetric pattern: B has A's row pointers and column array, so
    // template detection takes the A*A path (one fused class and template
    // pass; templates concern patterns only, the values stay B's)
    else
      O.a_is_b = true;
    O.bc = tc;
    O.bv = tv;
    bptr = tp;
  }
  // strictly increasing rows are implied when the symmetric-pattern gather
  // succeeded (it checks every row for descents and duplicates), so the
  // check runs on the device only when that transpose failed
  if (op == APPT_OP_AATS) launch_rows_strict(n, A.rowptr, A.col, ctl + 4, s, ctl + 3);
  O.bptr = bptr;
  // when row templates will be looked for, their class pass over B's rows
  // also writes the B-row metadata (a Gram needs it at once)
  O.bmeta_pending = cfg_.templates && !arena_ && n > 0 && nnz >= cfg_.templates_min_nnz &&
                    !(op == APPT_OP_AATS && gram_mode(A) != kGramNone);
  if (!O.bmeta_pending) stage_bmeta(W, O, bptr, nullptr, nullptr, 0);
}

// control words: see kCtlWords
int32_t* Engine::reset_ctl(Workspace& W) {
  int32_t* ctl = as<int32_t>(ensure(W.ctl, sizeof(int32_t) * kCtlWords));
  launch_fill_i32(ctl, kCtlWords, 0, stream_);
  return ctl;
}

void Engine::stage_bmeta(Workspace& W, const Operands& O, const int32_t* ptr, const int32_t* ptr2,
                         const uint8_t* sel, int64_t off2) {
  int4* bmeta = as<int4>(ensure(W.bmeta, sizeof(int4) * (size_t)std::max(O.nb, 1)));
  int32_t* ctl = as<int32_t>(W.ctl.p);
  launch_rowstats(O.nb, ptr, ptr2, sel, off2, O.bc, bmeta, ctl + 2, stream_);
}

void Engine::stage_prepass(Workspace& W, const Operands& O, const Config& cfg) {
  const int32_t n = O.n;
  hipStream_t s = stream_;
  const size_t nn = (size_t)std::max(n, 1);
  // row templates: not in the static scratch (no-allocation calls) nor where
  // a batch's columns are local to its blocks (cfg.templates off), nor for a
  // Gram, whose B rows are per-entry suffixes, not matrix rows
  W.tpl_on = cfg.templates && !arena_ && !O.gram && n > 0 && O.annz >= cfg.templates_min_nnz;
  // B-row metadata left by the operand stage: the class pass writes it, or
  // it is built here
  const bool meta_in_classes = O.bmeta_pending && W.tpl_on;
  if (O.bmeta_pending && !W.tpl_on) stage_bmeta(W, O, O.bptr, nullptr, nullptr, 0);
  if (meta_in_classes) (void)ensure(W.bmeta, sizeof(int4) * (size_t)std::max(O.nb, 1));
  int4* bmeta = as<int4>(W.bmeta.p);
  int64_t* rowf = as<int64_t>(ensure(W.rowf, sizeof(int64_t) * nn));
  int32_t* rowlo = as<int32_t>(ensure(W.rowlo, sizeof(int32_t) * nn));
  int32_t* rowhi = as<int32_t>(ensure(W.rowhi, sizeof(int32_t) * nn));
  uint8_t* symbin = as<uint8_t>(ensure(W.symbin, nn));
  int32_t* symrows = as<int32_t>(ensure(W.symrows, sizeof(int32_t) * nn));
  int32_t* rowcnt = as<int32_t>(ensure(W.rowcnt, sizeof(int32_t) * nn));
  int32_t* rowwords = as<int32_t>(ensure(W.rowwords, sizeof(int32_t) * nn));
  int32_t* dstats = as<int32_t>(ensure(W.stats, sizeof(int32_t) * stats_ints()));
  unsigned long long* dprod = reinterpret_cast<unsigned long long*>(dstats + stat_words() * 2);
  int32_t* dcursor = dstats + stat_words() * 2 + 4;
  int32_t* part = as<int32_t>(ensure(W.statpart, sizeof(int32_t) * (size_t)stats_part_elems() +
                                                      sizeof(unsigned long long) * kStatsMaxBlocks));
  launch_fill3_i32(dstats, (int64_t)stats_ints(), rowcnt, n, rowwords, n, s);
  if (W.tpl_on) {
    TplArgs& T = W.tpl;
    const size_t nb = (size_t)std::max(O.nb, 1);
    T.ckey = as<unsigned long long>(ensure(W.ckey, sizeof(unsigned long long) * kClassSlots));
    T.crep = as<int32_t>(ensure(W.crep, sizeof(int32_t) * kClassSlots));
    T.cpat = as<int32_t>(ensure(W.cpat, sizeof(int32_t) * (size_t)kClassSlots * (kTplMaxLen + 1)));
    T.bclass = as<int32_t>(ensure(W.bclass, sizeof(int32_t) * nb));
    const bool a_is_b = O.a_is_b && O.row0 == 0 && O.n == O.nb;
    T.aclass = a_is_b ? T.bclass : as<int32_t>(ensure(W.aclass, sizeof(int32_t) * nn));
    T.akey = a_is_b ? T.ckey : as<unsigned long long>(ensure(W.akey, sizeof(unsigned long long) * kClassSlots));
    T.arep = a_is_b ? T.crep : as<int32_t>(ensure(W.arep, sizeof(int32_t) * kClassSlots));
    T.tkey = as<unsigned long long>(ensure(W.tkey, sizeof(unsigned long long) * kTplSlots));
    T.trep = as<int32_t>(ensure(W.trep, sizeof(int32_t) * kTplSlots));
    T.tplid = as<int32_t>(ensure(W.tplid, sizeof(int32_t) * nn));
    T.info = as<int4>(ensure(W.tinfo, sizeof(int4) * kTplSlots));
    T.posoff = as<int32_t>(ensure(W.tposoff, sizeof(int32_t) * (size_t)kTplSlots * (kTplMaxLen + 1)));
    T.pos = as<uint16_t>(ensure(W.tpos, sizeof(uint16_t) * (size_t)kTplSlots * kTplMaxProducts));
    T.soff = as<int32_t>(ensure(W.tsoff, sizeof(int32_t) * (size_t)kTplSlots * kTplMaxProducts));
    T.tuple = as<int32_t>(ensure(W.ttuple, sizeof(int32_t) * (size_t)kTplSlots * kTplMaxLen));
    T.pos2 = as<uint32_t>(ensure(W.tpos2, sizeof(uint32_t) * (size_t)kTplSlots * 2 * kTplMaxLen * kTplPairs));
    T.grid = as<int2>(ensure(W.tgrid, sizeof(int2) * kTplSlots));
    T.matched = as<int32_t>(ensure(W.tmatched, sizeof(int32_t) * 4));
    T.row0 = O.row0;
    launch_tpl_detect(T, n, O.ap, O.ac, O.nb, bmeta, O.bc, a_is_b, meta_in_classes ? O.bptr : nullptr,
                      as<int32_t>(W.ctl.p) + 2, as<int32_t>(W.ctl.p) + kCtlLongRows, as<int32_t>(W.ctl.p) + 3, s);
  }
  launch_prepass(n, O.ap, O.ac, bmeta, O.upper, cfg.sym, rowf, rowlo, rowhi, symbin, dstats, dprod, part,
                 W.tpl_on ? &W.tpl : nullptr, rowcnt, rowwords, s);
  launch_bin_scatter(n, symbin, dstats, dcursor, symrows, nullptr, nullptr, s);
}

void Engine::read_sym(Workspace& W, Plan& P, bool want_ctl) {
  hipStream_t s = stream_;
  // a launch that failed is reported here, before its statistics are used
  APPT_HIP(hipGetLastError());
  const int32_t* dstats = as<int32_t>(W.stats.p);
  // symbolic stats and the products count in one copy (the numeric stats
  // and total between them are not read here)
  APPT_HIP(hipMemcpyAsync(hs_->sym, dstats, offsetof(HostStats, total), hipMemcpyDeviceToHost, s));
  if (want_ctl) APPT_HIP(hipMemcpyAsync(hs_->ctl, W.ctl.p, sizeof(hs_->ctl), hipMemcpyDeviceToHost, s));
  else hs_->ctl[4] = 1;
  wait_stream();
  std::memcpy(P.sym, hs_->sym, sizeof(P.sym));
  P.products = (int64_t)hs_->products;
  P.a_strict = hs_->ctl[4] == 0;
  P.bins_used = 0;
  for (int b = 1; b < kNumBins; b++) P.bins_used += P.sym[b * kStatFields] > 0 ? 1 : 0;
}

HeavyArgs Engine::heavy_args(Workspace& W, const Operands& O, const Config& cfg) {
  HeavyArgs h{};
  // dynamic LDS: the block's LDS less the heavy kernel's static staging (~21 KB)
  const int64_t dyn = lds_max_ - 24 * 1024;
  h.wc_sym = heavy_window_cols(dyn, cfg.heavy_cap);
  heavy_num_shape(dyn, cfg.heavy_cap, &h.wc_num, &h.vcap, &h.dcols);
  h.adapt = cfg.heavy_cap == 0;
  h.work = as<int32_t>(W.ctl.p);
  h.bflag = h.work + 2;
  h.fault = h.work + kCtlHeavyFault;
  h.cur = as<int32_t>(ensure(W.cur, sizeof(int32_t) * (size_t)std::max<int64_t>(O.annz, 1)));
  return h;
}

static RowArgs row_args(const Workspace& W, const Operands& O, const int32_t* rows, int32_t cnt, int upper,
                        const BinParams& bp) {
  RowArgs ra{};
  ra.rows = rows;
  ra.nrows = cnt;
  ra.ap = O.ap;
  ra.ac = O.ac;
  ra.av = O.av;
  ra.bc = O.bc;
  ra.bv = O.bv;
  ra.rowf = as<const int64_t>(W.rowf.p);
  ra.rowlo = as<const int32_t>(W.rowlo.p);
  ra.rowhi = as<const int32_t>(W.rowhi.p);
  ra.rowwords = as<int32_t>(W.rowwords.p);
  ra.upper = upper;
  ra.bnnz = O.bnnz;
  ra.bmeta = as<const int4>(W.bmeta.p);
  ra.wlist = as<uint2>(W.wlist.p);
  ra.lds_budget = bp.lds_budget;
  ra.ncols = O.ncols > 0 ? O.ncols : O.nb;
  if (W.wg_stride > 0) {
    ra.wgbits = as<uint32_t>(W.wgbits.p);
    ra.wgslot = as<int32_t>(W.wgslot.p);
    ra.wgstride = W.wg_stride;
  }
  return ra;
}

void Engine::stage_symbolic(Workspace& W, const Operands& O, const Config& cfg, const Plan& P) {
  hipStream_t s = stream_;
  (void)ensure(W.wlist, sizeof(uint2) * kWordListCap * (size_t)std::max(O.n, 1));
  // the workgroup-bitmap rows keep their bitmaps for the numeric workgroup
  // pass (which then skips its structure pass) when they fit the budget
  W.wg_stride = 0;
  {
    const int64_t cnt = P.sym[kBinWgBitmap * kStatFields], wmax = P.sym[kBinWgBitmap * kStatFields + 2];
    if (!arena_ && cnt > 0 && wmax > 0 && cnt * wmax * 4 <= kWgBitsBudget) {
      const size_t need = sizeof(uint32_t) * (size_t)(cnt * wmax);
      if (admit_ && W.wgbits.cap < need) admit_((int64_t)(need + need / 8 + 256), false);
      (void)ensure(W.wgbits, need);
      int32_t* slot = as<int32_t>(ensure(W.wgslot, sizeof(int32_t) * (size_t)std::max(O.n, 1)));
      launch_fill_i32(slot, O.n, -1, s);
      W.wg_stride = (int32_t)wmax;
    }
  }
  const int32_t* symrows = as<int32_t>(W.symrows.p);
  int32_t* rowcnt = as<int32_t>(W.rowcnt.p);
  HeavyArgs heavy{};
  if (P.sym[kBinHeavy * kStatFields] > 0) heavy = heavy_args(W, O, cfg);
  int32_t off = 0;
  for (int b = 1; b < kNumBins; b++) {
    const int32_t cnt = P.sym[b * kStatFields];
    if (cnt == 0) continue;
    if (b == kBinTemplate) {  // counts written by the pre-pass
      off += cnt;
      continue;
    }
    const int32_t mmax = P.sym[b * kStatFields + 1];
    const int32_t wmax = P.sym[b * kStatFields + 2];
    const int64_t rb = symbolic_row_lds_bytes(b, mmax, wmax);
    if (rb > lds_max_ - 8192) throw Error(APPT_ERR_UNSUPPORTED, "symbolic row exceeds LDS");
    // structure passes compute full rows (AAtS mirrors values only)
    RowArgs ra = row_args(W, O, symrows + off, cnt, 0, cfg.sym);
    if (b != kBinWgBitmap) {
      ra.wgbits = nullptr;
      ra.wgslot = nullptr;
      ra.wgstride = 0;
    }
    if (b == kBinHeavy && heavy_plan_cnt_ != nullptr)
      launch_take_plan(ra, heavy_plan_r0_, heavy_plan_cnt_, heavy_plan_words_, rowcnt, s);
    else
      launch_symbolic_bin(b, ra, wmax, rowcnt, heavy, s);
    off += cnt;
  }
}

void Engine::stage_offsets(Workspace& W, const Operands& O, const Config& cfg, int32_t* rp32) {
  const int32_t n = O.n;
  hipStream_t s = stream_;
  const size_t nn = (size_t)std::max(n, 1);
  int64_t* rowoff = as<int64_t>(ensure(W.rowoff, sizeof(int64_t) * ((size_t)n + 1)));
  int64_t* tmp = as<int64_t>(ensure(W.scan_tmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max(n, 1))));
  uint8_t* numbin = as<uint8_t>(ensure(W.numbin, nn));
  int32_t* numrows = as<int32_t>(ensure(W.numrows, sizeof(int32_t) * nn));
  int32_t* dstats = as<int32_t>(W.stats.p);
  int32_t* dnum = dstats + stat_words();
  int32_t* dcursor = dstats + stat_words() * 2 + 4 + kNumBins;
  launch_scan(as<int32_t>(W.rowcnt.p), n, rowoff, rp32, tmp, s);
  launch_numbin(n, O.ap, as<int64_t>(W.rowf.p), as<int32_t>(W.rowcnt.p), as<int32_t>(W.rowwords.p),
                as<int32_t>(W.rowlo.p), as<int32_t>(W.rowhi.p), cfg.num, numbin, dnum, as<int32_t>(W.statpart.p),
                W.tpl_on ? &W.tpl : nullptr, s);
  // the scatter also parks nnz(C) next to the products count (for n = 0 the
  // zeroed statistics hold it already)
  launch_bin_scatter(n, numbin, dnum, dcursor, numrows, rowoff + n,
                     reinterpret_cast<int64_t*>(dstats + stat_words() * 2 + 2), s);
}

void Engine::read_num(Workspace& W, const Operands& O, Plan& P) {
  hipStream_t s = stream_;
  APPT_HIP(hipGetLastError());
  const int32_t* dstats = as<int32_t>(W.stats.p);
  // numeric stats, products and nnz(C) in one copy
  APPT_HIP(hipMemcpyAsync(hs_->num, dstats + stat_words(),
                          offsetof(HostStats, ctl) - offsetof(HostStats, num), hipMemcpyDeviceToHost, s));
  wait_stream();
  (void)O;
  std::memcpy(P.num, hs_->num, sizeof(P.num));
  P.nnz = hs_->total;
}

void Engine::stage_numeric(Workspace& W, const Operands& O, const Config& cfg, const Plan& P, int32_t* ccol,
                           double* cval) {
  hipStream_t s = stream_;
  const int32_t* numrows = as<int32_t>(W.numrows.p);
  const int64_t* rowoff = as<int64_t>(W.rowoff.p);
  HeavyArgs heavy{};
  if (P.num[kBinHeavy * kStatFields] > 0) heavy = heavy_args(W, O, cfg);
  int32_t off = 0;
  for (int b = 1; b < kNumBins; b++) {
    const int32_t cnt = P.num[b * kStatFields];
    if (cnt == 0) continue;
    const int32_t cmax = P.num[b * kStatFields + 1];
    const int32_t wmax = P.num[b * kStatFields + 2];
    const int64_t rb = numeric_row_lds_bytes(b, cmax, wmax);
    if (rb > lds_max_ - 8192) throw Error(APPT_ERR_UNSUPPORTED, "numeric row exceeds LDS");
    // AAtS: values are formed for j >= i only (unless computed directly);
    // the structure is the full row
    RowArgs ra = row_args(W, O, numrows + off, cnt, P.direct_sym ? 0 : O.upper, cfg.num);
    if (b == kBinTemplate)
      launch_tpl_num(ra, W.tpl, wmax, cmax, (O.row0 == 0 && O.n == O.nb) ? O.n : 0, rowoff, ccol, cval,
                     num_cus(), lds_per_cu(), s);
    else
      launch_numeric_bin(b, ra, wmax, cmax, rowoff, ccol, cval, heavy, s);
    off += cnt;
  }
}

int Engine::gram_mode(const appt_csr_view& A) const {
  // the static scratch holds no per-entry descriptors or triangle.  By
  // default AAtS takes full rows: the Gram halves the products, but the
  // transposition of the mirrored half (a radix sort of nnz(C)/2 entries)
  // costs more than they do on every measured input, banded to power-law
  if (arena_ || A.n == 0 || cfg_.gram <= 0) return kGramNone;
  return cfg_.gram == 1 ? kGramUpper : kGramLower;
}

void Engine::stage_gram_operand(Workspace& W, Operands& O, bool lower) {
  hipStream_t s = stream_;
  const int64_t na = std::max<int64_t>(O.annz, 1);
  int4* vmeta = as<int4>(ensure(W.vmeta, sizeof(int4) * (size_t)na));
  int32_t* vac = as<int32_t>(ensure(W.vac, sizeof(int32_t) * (size_t)na));
  auto* full = reinterpret_cast<unsigned long long*>(as<int32_t>(W.ctl.p) + kCtlGramProducts);
  launch_gram_operand(O.n, O.ap, O.ac, O.nb, as<const int4>(W.bmeta.p), O.bc, lower ? 1 : 0, vmeta, vac, full, s);
  // the per-entry descriptors take the B-row metadata's place
  std::swap(W.bmeta, W.vmeta);
  O.ac = vac;
  O.nb = (int32_t)O.annz;
  O.upper = 0;  // every product formed is one of the triangle
  O.a_is_b = false;
  O.gram = true;
  O.gram_lower = lower;
}

void Engine::stage_gram_expand(Workspace& W, const Operands& O, const Plan& P, const int32_t* ucol,
                               const double* uval, int64_t nc, int32_t* rp32, int32_t* ccol, double* cval) {
  hipStream_t s = stream_;
  const int32_t n = O.n;
  if (n <= 0) return;
  const size_t nn = (size_t)n;
  const int64_t nu = P.nnz;
  // strictly-upper entries: U's less each non-empty row's diagonal
  const int64_t ns = (nc - ((int64_t)n - P.num[0])) / 2;
  const int64_t* uoff = as<int64_t>(W.rowoff.p);
  const int32_t* ucnt = as<int32_t>(W.rowcnt.p);
  int32_t* lower = as<int32_t>(ensure(W.mlow, sizeof(int32_t) * nn));
  int32_t* strict = as<int32_t>(ensure(W.mup, sizeof(int32_t) * nn));
  int32_t* total = as<int32_t>(ensure(W.gtot, sizeof(int32_t) * nn));
  int64_t* sup_off = as<int64_t>(ensure(W.mup_off, sizeof(int64_t) * (nn + 1)));
  int64_t* lo_off = as<int64_t>(ensure(W.mlo_off, sizeof(int64_t) * (nn + 1)));
  int64_t* coff = as<int64_t>(ensure(W.goff, sizeof(int64_t) * (nn + 1)));
  int64_t* tmp = as<int64_t>(ensure(W.scan_tmp, sizeof(int64_t) * (size_t)scan_tmp_elems(n)));
  // radix buffers sized by U (an upper bound of the strict entries)
  const size_t nz = (size_t)std::max<int64_t>(nu, 1);
  RadixScratch rx{};
  const int64_t hn = ((int64_t)1 << radix_bits(n)) * radix_tiles((int64_t)nz);
  rx.keys = as<int32_t>(ensure(W.rkeys, sizeof(int32_t) * nz));
  rx.rows = as<int32_t>(ensure(W.rrows, sizeof(int32_t) * nz));
  rx.vals = as<double>(ensure(W.rvals, sizeof(double) * nz));
  rx.keys2 = as<int32_t>(ensure(W.rkeys2, sizeof(int32_t) * nz));
  rx.hist = as<int32_t>(ensure(W.rhist, sizeof(int32_t) * (size_t)std::max<int64_t>(hn, 1)));
  rx.hoff = as<int64_t>(ensure(W.rhoff, sizeof(int64_t) * (size_t)(hn + 1)));
  rx.htmp = as<int64_t>(ensure(W.rtmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max<int64_t>(hn, 1))));
  int32_t* rows2 = as<int32_t>(ensure(W.rrows2, sizeof(int32_t) * nz));
  double* vals2 = as<double>(ensure(W.rvals2, sizeof(double) * nz));

  // row j of C: lower[j] mirrored entries, then U row j
  launch_fill_i32(lower, n, 0, s);
  launch_gram_mirrored(n, nu, uoff, ucol, lower, s);
  launch_gram_counts(n, ucnt, lower, strict, total, s);
  launch_scan(strict, n, sup_off, nullptr, tmp, s);
  launch_scan(lower, n, lo_off, nullptr, tmp, s);
  launch_scan(total, n, coff, nc <= INT32_MAX ? rp32 : nullptr, tmp, s);
  // the strictly-upper entries in row-major order, stably sorted by column:
  // column j's entries come out with their rows ascending, row j's lower part
  const bool odd = radix_passes(n) % 2 == 1;
  const int lt = O.gram_lower ? 1 : 0;
  launch_gram_extract(n, nu, uoff, ucol, uval, sup_off, lt, odd ? rx.keys : rx.keys2, odd ? rx.rows : rows2,
                      odd ? rx.vals : vals2, s);
  launch_radix_pairs(n, ns, rx, rows2, vals2, s);
  launch_gram_assemble(n, nc, coff, lt, lower, lo_off, rows2, vals2, uoff, ucol, uval, ccol, cval, s);
  // the last offset is nnz(C): checked against nc once the call has synchronised
  APPT_HIP(hipMemcpyAsync(&hs_->total, coff + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
}

void Engine::stage_mirror(Workspace& W, const Operands& O, const Plan& P, const int32_t* ccol, double* cval) {
  // AAtS: lower entries take the value of their upper mirror image
  if (!O.upper || P.direct_sym) return;
  const int32_t n = O.n;
  const int64_t* off = as<int64_t>(W.rowoff.p);
  // (the static scratch holds no result-sized mirror buffers)
  if (arena_ || P.nnz < cfg_.radix_mirror_min_nnz) {
    launch_mirror_pull(n, P.nnz, off, ccol, cval, stream_);
    return;
  }
  // large results: the strictly-upper entries sorted by column into the lower halves
  hipStream_t s = stream_;
  const size_t nn = (size_t)std::max(n, 1);
  int32_t* lower = as<int32_t>(ensure(W.mlow, sizeof(int32_t) * nn));
  int32_t* upper = as<int32_t>(ensure(W.mup, sizeof(int32_t) * nn));
  int64_t* lo_off = as<int64_t>(ensure(W.mlo_off, sizeof(int64_t) * (nn + 1)));
  int64_t* up_off = as<int64_t>(ensure(W.mup_off, sizeof(int64_t) * (nn + 1)));
  int64_t* tmp = as<int64_t>(ensure(W.scan_tmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max(n, 1))));
  launch_mirror_counts(n, off, ccol, lower, upper, s);
  launch_scan(lower, n, lo_off, nullptr, tmp, s);
  launch_scan(upper, n, up_off, nullptr, tmp, s);
  APPT_HIP(hipMemcpyAsync(&hs_->total, up_off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  APPT_HIP(hipStreamSynchronize(s));
  const int64_t count = hs_->total;
  if (count <= 0) return;
  const size_t nz = (size_t)count;
  const int64_t hn = ((int64_t)1 << radix_bits(n)) * radix_tiles(count);
  RadixScratch rx{};
  rx.keys = as<int32_t>(ensure(W.rkeys, sizeof(int32_t) * nz));
  rx.vals = as<double>(ensure(W.rvals, sizeof(double) * nz));
  rx.keys2 = as<int32_t>(ensure(W.rkeys2, sizeof(int32_t) * nz));
  rx.hist = as<int32_t>(ensure(W.rhist, sizeof(int32_t) * (size_t)std::max<int64_t>(hn, 1)));
  rx.hoff = as<int64_t>(ensure(W.rhoff, sizeof(int64_t) * (size_t)(hn + 1)));
  rx.htmp = as<int64_t>(ensure(W.rtmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max<int64_t>(hn, 1))));
  double* vals2 = as<double>(ensure(W.rvals2, sizeof(double) * nz));
  launch_mirror_radix(n, count, off, ccol, cval, upper, up_off, lo_off, rx, vals2, s);
}

// ---------------------------------------------------------------------------
// one call
// ---------------------------------------------------------------------------
void Engine::compute(const appt_csr_view& A, appt_op op, appt_csr_result* C) {
  compute_rows(A, op, 0, A.n, C, false);
}

void Engine::symbolic_rowptr(const appt_csr_view& A, appt_op op, int64_t* rowptr_host) {
  validate(A);
  if (!rowptr_host) throw Error(APPT_ERR_INVALID, "rowptr pointer is NULL");
  if (op != APPT_OP_AA && op != APPT_OP_AAT && op != APPT_OP_AATS) throw Error(APPT_ERR_INVALID, "bad op");
  // the symmetric A*A^T has A*A^T's rows (its panels run as A*A^T)
  if (op == APPT_OP_AATS) op = APPT_OP_AAT;
  APPT_HIP(hipSetDevice(device_));
  Workspace& W = ws_;
  Operands O;
  Plan P;
  stage_operands(W, A, op, O);
  stage_prepass(W, O, cfg_);
  read_sym(W, P, O.upper != 0);
  plan_key_ = RowPlanKey{};
  stage_symbolic(W, O, cfg_, P);
  int32_t* rp32 = as<int32_t>(pool_.get(sizeof(int32_t) * ((size_t)A.n + 1)));
  try {
    stage_offsets(W, O, cfg_, rp32);
    // keep the rows' counts and distinct words for the panels that follow
    const size_t rb = sizeof(int32_t) * (size_t)std::max(A.n, 1);
    APPT_HIP(hipMemcpyAsync(ensure(W.plancnt, rb), W.rowcnt.p, rb, hipMemcpyDeviceToDevice, stream_));
    APPT_HIP(hipMemcpyAsync(ensure(W.planwords, rb), W.rowwords.p, rb, hipMemcpyDeviceToDevice, stream_));
    APPT_HIP(hipMemcpyAsync(rowptr_host, W.rowoff.p, sizeof(int64_t) * ((size_t)A.n + 1), hipMemcpyDeviceToHost,
                            stream_));
    APPT_HIP(hipStreamSynchronize(stream_));
  } catch (...) {
    pool_.put(rp32);
    throw;
  }
  pool_.put(rp32);
  plan_key_ = RowPlanKey{A.rowptr, A.col, A.val, A.n, A.nnz, (int)op};
}

void Engine::compute_rows(const appt_csr_view& A, appt_op op, int32_t r0, int32_t r1, appt_csr_result* C,
                          bool use_plan) {
  validate(A);
  if (!C) throw Error(APPT_ERR_INVALID, "result pointer is NULL");
  if (op != APPT_OP_AA && op != APPT_OP_AAT && op != APPT_OP_AATS) throw Error(APPT_ERR_INVALID, "bad op");
  if (r0 < 0 || r1 < r0 || r1 > A.n) throw Error(APPT_ERR_INVALID, "row range outside [0, n]");
  // a row panel of the symmetric A*A^T holds complete rows, which A*A^T's
  // own rows are: no upper triangle to mirror from
  if (op == APPT_OP_AATS && (r0 != 0 || r1 != A.n)) op = APPT_OP_AAT;
  APPT_HIP(hipSetDevice(device_));
  std::memset(C, 0, sizeof(*C));
  times_ = appt_phase_times{};
  const int32_t n = r1 - r0;
  if (admit_) admit_(work_bound(n, A.n, A.nnz), false);
  hipStream_t s = stream_;
  Workspace& W = ws_;
  Operands O;
  Plan P;
  record(0);
  stage_operands(W, A, op, O);
  // a panel: output rows r0..r1-1 of A against all of B
  O.ap = A.rowptr + r0;
  O.n = n;
  O.row0 = r0;
  const int gmode = op == APPT_OP_AATS ? gram_mode(A) : kGramNone;
  if (gmode == kGramUpper || gmode == kGramLower) stage_gram_operand(W, O, gmode == kGramLower);
  record(1);
  stage_prepass(W, O, cfg_);
  read_sym(W, P, O.upper != 0 || O.gram);
  if (O.gram) {
    unsigned long long full;
    std::memcpy(&full, hs_->ctl + kCtlGramProducts, sizeof(full));
    P.gram_products = (int64_t)full;
  }
  record(2);
  {
    // heavy rows reuse the row-pointer plan of this view and op, if any
    const bool planned = use_plan && plan_key_.rowptr == A.rowptr && plan_key_.col == A.col &&
                         plan_key_.val == A.val && plan_key_.n == A.n && plan_key_.nnz == A.nnz &&
                         plan_key_.op == (int)op;
    struct Reset {
      Engine* e;
      ~Reset() { e->heavy_plan_cnt_ = nullptr; e->heavy_plan_words_ = nullptr; e->heavy_plan_r0_ = 0; }
    } reset{this};
    if (planned) {
      heavy_plan_cnt_ = as<const int32_t>(W.plancnt.p);
      heavy_plan_words_ = as<const int32_t>(W.planwords.p);
      heavy_plan_r0_ = r0;
    }
    stage_symbolic(W, O, cfg_, P);
  }
  record(3);
  int32_t* rp32 = as<int32_t>(pool_.get(sizeof(int32_t) * ((size_t)n + 1)));
  int32_t* ccol = nullptr;
  double* cval = nullptr;
  int64_t cnnz = 0;
  void* rowptr_out = rp32;
  try {
    stage_offsets(W, O, cfg_, rp32);
    read_num(W, O, P);
    // AAtS on strictly increasing rows with every numeric row in an
    // order-exact bin: c_ij and c_ji sum the same products a_ik * a_jk in the
    // same ascending k order, so computing the lower entries directly gives
    // the mirror's values bit for bit, without the mirror pass
    P.direct_sym = O.upper && P.a_strict && P.num[kBinWgBitmap * kStatFields] == 0 &&
                   P.num[kBinWgHash * kStatFields] == 0 && P.num[kBinHeavy * kStatFields] == 0;
    record(4);
    // a Gram's C: U's off-diagonal entries twice, each non-empty row's
    // diagonal once
    cnnz = O.gram ? 2 * P.nnz - ((int64_t)n - P.num[0]) : P.nnz;
    if (admit_) admit_(result_bytes(n, cnnz), true);
    // output buffers (the symbolic pass sized C exactly)
    ccol = as<int32_t>(pool_.get(sizeof(int32_t) * (size_t)std::max<int64_t>(cnnz, 1)));
    cval = as<double>(pool_.get(sizeof(double) * (size_t)std::max<int64_t>(cnnz, 1)));
    if (O.gram) {
      int32_t* ucol = as<int32_t>(ensure(W.ucol, sizeof(int32_t) * (size_t)std::max<int64_t>(P.nnz, 1)));
      double* uval = as<double>(ensure(W.uval, sizeof(double) * (size_t)std::max<int64_t>(P.nnz, 1)));
      stage_numeric(W, O, cfg_, P, ucol, uval);
      record(5);
      stage_gram_expand(W, O, P, ucol, uval, cnnz, rp32, ccol, cval);
    } else {
      stage_numeric(W, O, cfg_, P, ccol, cval);
      record(5);
      stage_mirror(W, O, P, ccol, cval);
    }
    if (cfg_.rowptr64(cnnz)) {
      int64_t* rp64 = as<int64_t>(pool_.get(sizeof(int64_t) * ((size_t)n + 1)));
      rowptr_out = rp64;
      launch_copy_i64(as<int64_t>(O.gram ? W.goff.p : W.rowoff.p), (int64_t)n + 1, rp64, s);
    }
    record(6);
    APPT_HIP(hipGetLastError());
    const bool heavy = P.num[kBinHeavy * kStatFields] > 0;
    if (heavy) APPT_HIP(hipMemcpyAsync(hs_->ctl, W.ctl.p, sizeof(hs_->ctl), hipMemcpyDeviceToHost, s));
    wait_stream();
    if (heavy && hs_->ctl[kCtlHeavyFault] != 0)
      throw Error(APPT_ERR_HIP, "heavy-row window could not advance (window shape below one word of outputs)");
    if (O.gram && hs_->total != cnnz)
      throw Error(APPT_ERR_HIP, "Gram expansion produced " + std::to_string(hs_->total) + " entries, expected " +
                                    std::to_string(cnnz));
  } catch (...) {
    // nothing enqueued may still write these buffers once they are reused
    (void)hipStreamSynchronize(s);
    if (rowptr_out != rp32) pool_.put(rowptr_out);
    pool_.put(rp32);
    pool_.put(ccol);
    pool_.put(cval);
    throw;
  }
  if (rowptr_out != rp32) pool_.put(rp32);

  Result* R = new Result{rowptr_out, ccol, cval};
  C->n = n;
  C->nnz = cnnz;
  C->rowptr_is_64 = rowptr_out != rp32 ? 1 : 0;
  C->rowptr = rowptr_out;
  C->col = ccol;
  C->val = cval;
  C->_priv = R;

  bins_ = appt_bin_report{};
  bins_.n_bins = kNumBins;
  for (int b = 0; b < kNumBins; b++) {
    bins_.sym_rows[b] = P.sym[b * kStatFields];
    bins_.num_rows[b] = P.num[b * kStatFields];
    bins_.sym_metric_max[b] = P.sym[b * kStatFields + 1];
    bins_.num_metric_max[b] = P.num[b * kStatFields + 1];
  }
  times_.products = O.gram ? P.gram_products : P.products;
  times_.nnz_c = cnnz;
  times_.n_bins_used = P.bins_used;
  times_.sync_points = 3;
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

void Engine::compute_noalloc(const appt_csr_view& A, appt_op op, int32_t* rowptr, int32_t* col, double* val,
                             int64_t capacity, int64_t* nnz_out) {
  validate(A);
  if (op != APPT_OP_AA && op != APPT_OP_AAT && op != APPT_OP_AATS) throw Error(APPT_ERR_INVALID, "bad op");
  // sizes first: an input too large for the static scratch is refused before any work
  const int64_t need = static_scratch_need(A.n, A.nnz);
  if (need > (int64_t)kStaticScratchBytes)
    throw Error(APPT_ERR_UNSUPPORTED, "n = " + std::to_string(A.n) + ", nnz = " + std::to_string(A.nnz) +
                                          " needs up to " + std::to_string(need) + " bytes of scratch; the "
                                          "no-allocation entry point has " +
                                          std::to_string(kStaticScratchBytes) + " (use appt_spgemm_compute)");
  if (!nnz_out || (A.n > 0 && !rowptr) || capacity < 0 || (capacity > 0 && (!col || !val)))
    throw Error(APPT_ERR_INVALID, "output buffers missing");
  APPT_HIP(hipSetDevice(device_));
  void* base = static_scratch(device_);
  if (!base) throw Error(APPT_ERR_HIP, "static scratch of this device has no address");
  *nnz_out = 0;
  times_ = appt_phase_times{};
  // every buffer is carved afresh from the start of the array: the previous
  // call's contents are stale and each stage clears what it needs
  Arena arena{static_cast<unsigned char*>(base), kStaticScratchBytes, 0};
  ws_static_.each([](Scratch& x) { x.p = nullptr; x.cap = 0; });
  ws_static_.wg_stride = 0;
  struct Leave {
    Engine* e;
    ~Leave() { e->arena_ = nullptr; }
  } leave{this};
  arena_ = &arena;
  const int32_t n = A.n;
  hipStream_t s = stream_;
  Workspace& W = ws_static_;
  Operands O;
  Plan P;
  try {
    record(0);
    stage_operands(W, A, op, O);
    record(1);
    stage_prepass(W, O, cfg_);
    read_sym(W, P, O.upper != 0);
    record(2);
    stage_symbolic(W, O, cfg_, P);
    record(3);
    stage_offsets(W, O, cfg_, rowptr);
    read_num(W, O, P);
    *nnz_out = P.nnz;
    if (P.nnz > INT32_MAX) throw Error(APPT_ERR_OVERFLOW, "nnz(C) needs 64-bit row pointers");
    if (P.nnz > capacity)
      throw Error(APPT_ERR_OVERFLOW, "nnz(C) = " + std::to_string(P.nnz) + " exceeds the capacity " +
                                         std::to_string(capacity));
    P.direct_sym = O.upper && P.a_strict && P.num[kBinWgBitmap * kStatFields] == 0 &&
                   P.num[kBinWgHash * kStatFields] == 0 && P.num[kBinHeavy * kStatFields] == 0;
    record(4);
    stage_numeric(W, O, cfg_, P, col, val);
    record(5);
    stage_mirror(W, O, P, col, val);
    record(6);
    APPT_HIP(hipGetLastError());
    const bool heavy = P.num[kBinHeavy * kStatFields] > 0;
    if (heavy) APPT_HIP(hipMemcpyAsync(hs_->ctl, W.ctl.p, sizeof(hs_->ctl), hipMemcpyDeviceToHost, s));
    wait_stream();
    if (heavy && hs_->ctl[kCtlHeavyFault] != 0)
      throw Error(APPT_ERR_HIP, "heavy-row window could not advance (window shape below one word of outputs)");
  } catch (...) {
    // the next call reuses the array: nothing of this one may still run
    (void)hipStreamSynchronize(s);
    throw;
  }
  bins_ = appt_bin_report{};
  bins_.n_bins = kNumBins;
  for (int b = 0; b < kNumBins; b++) {
    bins_.sym_rows[b] = P.sym[b * kStatFields];
    bins_.num_rows[b] = P.num[b * kStatFields];
    bins_.sym_metric_max[b] = P.sym[b * kStatFields + 1];
    bins_.num_metric_max[b] = P.num[b * kStatFields + 1];
  }
  times_.products = P.products;
  times_.nnz_c = P.nnz;
  times_.n_bins_used = P.bins_used;
  times_.sync_points = 3;
  (void)n;
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

void Engine::digest(const appt_csr_result& C, uint64_t out[2]) {
  if (C.n < 0 || C.nnz < 0 || (C.nnz > 0 && (!C.col || !C.val)) || !C.rowptr)
    throw Error(APPT_ERR_INVALID, "not a result");
  APPT_HIP(hipSetDevice(device_));
  auto* d = as<unsigned long long>(ensure(ws_.digest, 2 * sizeof(unsigned long long)));
  launch_digest(C.n, C.nnz, C.rowptr, C.rowptr_is_64 != 0, C.col, C.val, d, stream_);
  APPT_HIP(hipGetLastError());
  APPT_HIP(hipMemcpyAsync(&hs_->total, d, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
  APPT_HIP(hipMemcpyAsync(&hs_->products, d + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
  APPT_HIP(hipStreamSynchronize(stream_));
  out[0] = (uint64_t)hs_->total;
  out[1] = (uint64_t)hs_->products;
}

void Engine::check(const appt_csr_view& A) {
  validate(A);
  APPT_HIP(hipSetDevice(device_));
  int32_t* flags = reset_ctl(ws_) + 5;
  launch_csr_check(A.n, A.nnz, A.rowptr, A.col, flags, stream_);
  APPT_HIP(hipMemcpyAsync(hs_->ctl, ws_.ctl.p, sizeof(hs_->ctl), hipMemcpyDeviceToHost, stream_));
  APPT_HIP(hipStreamSynchronize(stream_));
  if (hs_->ctl[5]) throw Error(APPT_ERR_INVALID, "rowptr[0] != 0 or rowptr[n] != nnz");
  if (hs_->ctl[6]) throw Error(APPT_ERR_INVALID, "row pointers decrease");
  if (hs_->ctl[7]) throw Error(APPT_ERR_INVALID, "column index outside [0, n)");
}

void Engine::eliminate_zeros(appt_csr_result* C) {
  if (!C || !C->_priv) throw Error(APPT_ERR_INVALID, "not an engine-owned result");
  APPT_HIP(hipSetDevice(device_));
  hipStream_t s = stream_;
  const int32_t n = C->n;
  const size_t n1 = (size_t)n + 1;
  Workspace& W = ws_;
  int32_t* cnt = as<int32_t>(ensure(W.rowcnt, sizeof(int32_t) * n1));
  int64_t* off = as<int64_t>(ensure(W.rowoff, sizeof(int64_t) * n1));
  int64_t* tmp = as<int64_t>(ensure(W.scan_tmp, sizeof(int64_t) * (size_t)scan_tmp_elems(std::max(n, 1))));
  launch_count_nonzero(n, C->rowptr, C->rowptr_is_64 != 0, C->val, cnt, s);
  launch_scan(cnt, n, off, nullptr, tmp, s);
  int64_t nnz = 0;
  APPT_HIP(hipMemcpyAsync(&hs_->total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  APPT_HIP(hipStreamSynchronize(s));
  nnz = hs_->total;
  if (nnz == C->nnz) return;  // nothing to drop
  const bool rp64 = cfg_.rowptr64(nnz);
  void* rp = nullptr;
  int32_t* col = nullptr;
  double* val = nullptr;
  try {
    rp = pool_.get((rp64 ? sizeof(int64_t) : sizeof(int32_t)) * n1);
    col = as<int32_t>(pool_.get(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
    val = as<double>(pool_.get(sizeof(double) * (size_t)std::max<int64_t>(nnz, 1)));
    launch_compact_nonzero(n, C->rowptr, C->rowptr_is_64 != 0, C->col, C->val, off, col, val, s);
    if (rp64) launch_copy_i64(off, (int64_t)n + 1, as<int64_t>(rp), s);
    else launch_i64_to_i32(off, (int64_t)n + 1, as<int32_t>(rp), s);
    APPT_HIP(hipGetLastError());
    APPT_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    pool_.put(rp);
    pool_.put(col);
    pool_.put(val);
    throw;
  }
  release(C);
  Result* R = new Result{rp, col, val};
  C->n = n;
  C->nnz = nnz;
  C->rowptr_is_64 = rp64 ? 1 : 0;
  C->rowptr = rp;
  C->col = col;
  C->val = val;
  C->_priv = R;
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
