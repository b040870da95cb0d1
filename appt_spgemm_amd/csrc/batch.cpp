// batch.cpp -- many independent matrices as one block-diagonal product.
//
// The packed rows of all matrices form one operand; the engine's stages run
// over it unchanged (B rows are addressed through the B-row metadata, so a
// block's B side may be its own rows or its transpose).  Planning runs the
// stages once with the engine's two synchronisations to size C and fix the
// launch shapes; the same launches are then captured into a HIP graph and
// replayed with no host synchronisation inside.
#include <algorithm>
#include <cstring>

#include "engine.hpp"

namespace appt {

appt_tuning tuning_of(const Config& c) {
  appt_tuning t{};
  for (int k = 0; k < c.sym.ncls && k < APPT_MAX_CLASSES; k++) t.f_class_max[k] = c.sym.cls_max[k];
  t.wg_lds_bytes = c.sym.wg_bytes;
  t.lds_budget_bytes = c.sym.lds_budget;
  t.bitmap_words_factor = c.num.dense_ratio;
  for (int g = 0; g < 3; g++) t.group_for_len[g] = c.sym.glen_max[g];
  t.heavy_window_cols = c.heavy_cap;
  t.sort_span_factor = c.num.sort_ratio > 0 ? c.num.sort_ratio : -1;
  t.sort_span_factor_symbolic = c.sym.sort_ratio > 0 ? c.sym.sort_ratio : -1;
  t.transpose_radix_min_nnz = c.radix_min_nnz > INT32_MAX ? -1 : (int32_t)c.radix_min_nnz;
  t.mirror_radix_min_nnz = c.radix_mirror_min_nnz > INT32_MAX ? -1 : (int32_t)c.radix_mirror_min_nnz;
  for (int k = 0; k < kNumCls; k++) t.class_lds_bytes[k] = c.sym.cls_bytes[k];
  t.sort_max_products = c.sym.sort_max > 0 ? c.sym.sort_max : -1;
  t.rowptr64_min_nnz = c.rowptr64_min_nnz > INT32_MAX ? 0 : (int32_t)c.rowptr64_min_nnz;
  t.row_templates = !c.templates ? -1 : (c.templates_min_nnz > INT32_MAX ? -1 : (int32_t)c.templates_min_nnz);
  t.gram = c.gram;
  return t;
}

template <typename T>
static T* as(void* p) { return static_cast<T*>(p); }

void* Batch::alloc(size_t bytes) {
  void* p = E_.pool().get(bytes);
  owned_.push_back(p);
  return p;
}

Batch::Batch(Engine& e, int32_t count, const int32_t* n_each, const int64_t* nnz_each, const appt_op* ops)
    : E_(e), cfg_(e.config()) {
  // block-local columns: relative row patterns are not comparable across blocks
  cfg_.templates = false;
  if (count < 0 || (count > 0 && (!n_each || !nnz_each || !ops))) throw Error(APPT_ERR_INVALID, "bad batch arguments");
  K_ = count;
  blk_row_h_.assign((size_t)K_ + 1, 0);
  blk_nnz_h_.assign((size_t)K_ + 1, 0);
  std::vector<uint8_t> op_h((size_t)std::max(K_, 1), 0);
  std::vector<int32_t> toff_h((size_t)K_ + 1, 0);
  int64_t rows = 0, nnz = 0, tnnz = 0;
  for (int32_t k = 0; k < K_; k++) {
    if (n_each[k] < 0 || nnz_each[k] < 0) throw Error(APPT_ERR_INVALID, "negative batch matrix size");
    if (ops[k] != APPT_OP_AA && ops[k] != APPT_OP_AAT)
      throw Error(APPT_ERR_UNSUPPORTED, "batched ops are A*A and A*A^T");
    blk_row_h_[k] = (int32_t)rows;
    blk_nnz_h_[k] = nnz;
    toff_h[k] = (int32_t)tnnz;
    max_n_ = std::max(max_n_, n_each[k]);
    if (ops[k] == APPT_OP_AAT) max_nt_ = std::max(max_nt_, n_each[k]);
    op_h[k] = ops[k] == APPT_OP_AAT ? 1 : 0;
    rows += n_each[k];
    nnz += nnz_each[k];
    if (ops[k] == APPT_OP_AAT) tnnz += nnz_each[k];
    if (rows + K_ >= INT32_MAX || nnz + tnnz >= INT32_MAX)
      throw Error(APPT_ERR_OVERFLOW, "batch exceeds int32 rows or entries");
  }
  blk_row_h_[K_] = (int32_t)rows;
  blk_nnz_h_[K_] = nnz;
  toff_h[K_] = (int32_t)tnnz;
  N_ = (int32_t)rows;
  nnz_a_ = nnz;
  nnz_t_ = tnnz;
  APPT_HIP(hipSetDevice(E_.device()));
  hipStream_t s = E_.stream();
  const size_t n1 = (size_t)N_ + 1, nn = (size_t)std::max(N_, 1);
  in_rp_ = as<int32_t>(alloc(sizeof(int32_t) * n1));
  in_col_ = as<int32_t>(alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz_a_ + nnz_t_, 1)));
  in_val_ = as<double>(alloc(sizeof(double) * (size_t)std::max<int64_t>(nnz_a_ + nnz_t_, 1)));
  blk_row_ = as<int32_t>(alloc(sizeof(int32_t) * ((size_t)K_ + 1)));
  blk_nnz_ = as<int64_t>(alloc(sizeof(int64_t) * ((size_t)K_ + 1)));
  blk_op_ = as<uint8_t>(alloc((size_t)std::max(K_, 1)));
  blk_toff_ = as<int32_t>(alloc(sizeof(int32_t) * ((size_t)K_ + 1)));
  rowbase_ = as<int32_t>(alloc(sizeof(int32_t) * nn));
  rowblk_ = as<int32_t>(alloc(sizeof(int32_t) * nn));
  rowop_ = as<uint8_t>(alloc(nn));
  acg_ = as<int32_t>(alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz_a_, 1)));
  tcnt_ = as<int32_t>(alloc(sizeof(int32_t) * n1));
  tp32_ = as<int32_t>(alloc(sizeof(int32_t) * n1));
  tcur_ = as<int32_t>(alloc(sizeof(int32_t) * n1));
  toff_ = as<int64_t>(alloc(sizeof(int64_t) * n1));
  tscan_ = as<int64_t>(alloc(sizeof(int64_t) * (size_t)scan_tmp_elems(std::max(N_, 1))));
  rp32_ = as<int32_t>(alloc(sizeof(int32_t) * n1));
  crp_ = as<int32_t>(alloc(sizeof(int32_t) * ((size_t)N_ + K_ + 1)));
  blkoff_ = as<int64_t>(alloc(sizeof(int64_t) * ((size_t)K_ + 1)));
  blkprod_ = as<int64_t>(alloc(sizeof(int64_t) * ((size_t)K_ + 1)));
  src_ = as<BatchSrc>(alloc(sizeof(BatchSrc) * (size_t)std::max(K_, 1)));
  APPT_HIP(hipMemcpyAsync(blk_row_, blk_row_h_.data(), sizeof(int32_t) * ((size_t)K_ + 1), hipMemcpyHostToDevice, s));
  APPT_HIP(hipMemcpyAsync(blk_nnz_, blk_nnz_h_.data(), sizeof(int64_t) * ((size_t)K_ + 1), hipMemcpyHostToDevice, s));
  if (K_ > 0) APPT_HIP(hipMemcpyAsync(blk_op_, op_h.data(), (size_t)K_, hipMemcpyHostToDevice, s));
  APPT_HIP(hipMemcpyAsync(blk_toff_, toff_h.data(), sizeof(int32_t) * ((size_t)K_ + 1), hipMemcpyHostToDevice, s));
  if (K_ > 0) launch_batch_rows(N_, K_, blk_row_, blk_op_, rowbase_, rowblk_, rowop_, s);
  APPT_HIP(hipStreamSynchronize(s));
}

Batch::~Batch() {
  (void)hipSetDevice(E_.device());
  (void)hipStreamSynchronize(E_.stream());
  drop_graph();
  E_.free_workspace(W_);
  if (ccol_) E_.pool().put(ccol_);
  if (cval_) E_.pool().put(cval_);
  for (void* p : owned_)
    if (p) E_.pool().put(p);
}

void Batch::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

void Batch::inputs(appt_batch_buffers* out) const {
  out->n_total = N_;
  out->nnz = nnz_a_;
  out->rowptr = in_rp_;
  out->col = in_col_;
  out->val = in_val_;
}

void Batch::upload(const appt_host_csr& h) {
  if (h.n != N_ || h.nnz != nnz_a_) throw Error(APPT_ERR_INVALID, "packed input does not match the batch sizes");
  if (!h.rowptr || (h.nnz > 0 && (!h.col || !h.val))) throw Error(APPT_ERR_INVALID, "packed input has NULL arrays");
  APPT_HIP(hipSetDevice(E_.device()));
  APPT_HIP(hipMemcpy(in_rp_, h.rowptr, sizeof(int32_t) * ((size_t)N_ + 1), hipMemcpyHostToDevice));
  if (h.nnz > 0) {
    APPT_HIP(hipMemcpy(in_col_, h.col, sizeof(int32_t) * (size_t)h.nnz, hipMemcpyHostToDevice));
    APPT_HIP(hipMemcpy(in_val_, h.val, sizeof(double) * (size_t)h.nnz, hipMemcpyHostToDevice));
  }
}

void Batch::gather(const appt_csr_view* each) {
  if (K_ == 0) return;
  if (!each) throw Error(APPT_ERR_INVALID, "matrix list is NULL");
  std::vector<BatchSrc> h((size_t)K_);
  for (int32_t k = 0; k < K_; k++) {
    const appt_csr_view& v = each[k];
    if (v.n != blk_row_h_[k + 1] - blk_row_h_[k] || v.nnz != blk_nnz_h_[k + 1] - blk_nnz_h_[k])
      throw Error(APPT_ERR_INVALID, "matrix " + std::to_string(k) + " does not match the batch sizes");
    h[k] = BatchSrc{v.rowptr, v.col, v.val, v.n, v.nnz};
  }
  APPT_HIP(hipSetDevice(E_.device()));
  hipStream_t s = E_.stream();
  APPT_HIP(hipMemcpyAsync(src_, h.data(), sizeof(BatchSrc) * (size_t)K_, hipMemcpyHostToDevice, s));
  launch_batch_gather(K_, src_, blk_row_, blk_nnz_, in_rp_, in_col_, in_val_, s);
  APPT_HIP(hipStreamSynchronize(s));
}

void Batch::operands() {
  hipStream_t s = E_.stream();
  O_ = Operands{};
  O_.n = N_;
  O_.nb = N_;
  O_.annz = nnz_a_;
  O_.ap = in_rp_;
  O_.ac = acg_;
  O_.av = in_val_;
  O_.bc = in_col_;
  O_.bv = in_val_;
  O_.bnnz = nnz_a_ + nnz_t_;
  O_.upper = 0;
  O_.ncols = std::max(max_n_, 1);
  (void)E_.reset_ctl(W_);
  // A*A^T matrices of at most kBatchTBlockN rows: one workgroup transposes
  // each (LDS column counts); otherwise global counts and cursors
  const bool tblock = nnz_t_ > 0 && max_nt_ <= kBatchTBlockN;
  if (!tblock) launch_fill_i32(tcnt_, (int64_t)N_ + 1, 0, s);
  // row ids of the entries (local to their matrix) for the workgroup transposes
  int32_t* rid = tblock ? static_cast<int32_t*>(E_.ensure(W_.rkeys, sizeof(int32_t) * (size_t)nnz_a_)) : nullptr;
  launch_batch_acg(N_, in_rp_, in_col_, rowbase_, rowop_, acg_, tblock ? nullptr : tcnt_, rid, s);
  if (tblock) {
    launch_batch_tblock(K_, N_, max_nt_, nnz_t_, blk_row_, blk_nnz_, blk_op_, blk_toff_, rid, in_col_, in_val_,
                        tp32_, in_col_ + nnz_a_, in_val_ + nnz_a_, s);
    int32_t* slist = static_cast<int32_t*>(E_.ensure(W_.sortlist, sizeof(int32_t) * (size_t)sort_rows_scratch_elems(N_)));
    launch_sort_rows(N_, tp32_, in_col_ + nnz_a_, in_val_ + nnz_a_, slist, nullptr, s);
  } else if (nnz_t_ > 0) {
    launch_scan(tcnt_, N_, toff_, tp32_, tscan_, s);
    launch_i64_to_i32(toff_, N_, tcur_, s);
    launch_batch_tscatter(N_, in_rp_, acg_, rowbase_, rowop_, tcur_, nnz_a_, in_col_, in_val_, s);
    int32_t* slist = static_cast<int32_t*>(E_.ensure(W_.sortlist, sizeof(int32_t) * (size_t)sort_rows_scratch_elems(N_)));
    launch_sort_rows(N_, tp32_, in_col_ + nnz_a_, in_val_ + nnz_a_, slist, nullptr, s);
  } else {
    launch_fill_i32(tp32_, (int64_t)N_ + 1, 0, s);
  }
  E_.stage_bmeta(W_, O_, in_rp_, tp32_, rowop_, nnz_a_);
}

void Batch::stages() {
  hipStream_t s = E_.stream();
  operands();
  E_.stage_prepass(W_, O_, cfg_);
  E_.stage_symbolic(W_, O_, cfg_, plan_);
  E_.stage_offsets(W_, O_, cfg_, rp32_);
  E_.stage_numeric(W_, O_, cfg_, plan_, ccol_, cval_);
  launch_batch_local_rowptr(N_, K_, as<int64_t>(W_.rowoff.p), rowbase_, rowblk_, blk_row_, crp_, blkoff_, s);
}

void Batch::run_planning() {
  hipStream_t s = E_.stream();
  operands();
  E_.stage_prepass(W_, O_, cfg_);
  E_.read_sym(W_, plan_);
  E_.stage_symbolic(W_, O_, cfg_, plan_);
  E_.stage_offsets(W_, O_, cfg_, rp32_);
  E_.read_num(W_, O_, plan_);
  if (E_.admission()) E_.admission()(result_bytes(N_, plan_.nnz) + 4 * (int64_t)K_, true);
  const size_t cap = (size_t)std::max<int64_t>(plan_.nnz, 1);
  if (ccol_) E_.pool().put(ccol_);
  if (cval_) E_.pool().put(cval_);
  ccol_ = nullptr;
  cval_ = nullptr;
  ccol_ = as<int32_t>(E_.pool().get(sizeof(int32_t) * cap));
  cval_ = as<double>(E_.pool().get(sizeof(double) * cap));
  E_.stage_numeric(W_, O_, cfg_, plan_, ccol_, cval_);
  launch_batch_local_rowptr(N_, K_, as<int64_t>(W_.rowoff.p), rowbase_, rowblk_, blk_row_, crp_, blkoff_, s);
  launch_batch_block_sum(K_, blk_row_, as<int64_t>(W_.rowf.p), blkprod_, s);
  blkoff_h_.assign((size_t)K_ + 1, 0);
  blkprod_h_.assign((size_t)K_ + 1, 0);
  APPT_HIP(hipMemcpyAsync(blkoff_h_.data(), blkoff_, sizeof(int64_t) * ((size_t)K_ + 1), hipMemcpyDeviceToHost, s));
  if (K_ > 0)
    APPT_HIP(hipMemcpyAsync(blkprod_h_.data(), blkprod_, sizeof(int64_t) * (size_t)K_, hipMemcpyDeviceToHost, s));
  APPT_HIP(hipGetLastError());
  APPT_HIP(hipStreamSynchronize(s));
  for (int32_t k = 0; k < K_; k++)
    if (blkoff_h_[k + 1] - blkoff_h_[k] > INT32_MAX)
      throw Error(APPT_ERR_OVERFLOW, "a batch result exceeds 2^31 entries");
}

void Batch::plan(bool use_graph) {
  APPT_HIP(hipSetDevice(E_.device()));
  drop_graph();
  planned_ = false;
  run_planning();
  use_graph_ = use_graph;
  if (use_graph_) {
    hipStream_t s = E_.stream();
    APPT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    try {
      stages();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    APPT_HIP(hipStreamEndCapture(s, &graph_));
    APPT_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
  planned_ = true;
}

void Batch::execute() {
  if (!planned_) throw Error(APPT_ERR_INVALID, "batch is not planned");
  APPT_HIP(hipSetDevice(E_.device()));
  hipStream_t s = E_.stream();
  if (exec_) APPT_HIP(hipGraphLaunch(exec_, s));
  else stages();
  APPT_HIP(hipGetLastError());
  APPT_HIP(hipStreamSynchronize(s));
}

float Batch::time_eager(int reps) {
  hipStream_t s = E_.stream();
  hipEvent_t e0, e1;
  APPT_HIP(hipEventCreate(&e0));
  APPT_HIP(hipEventCreate(&e1));
  APPT_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) stages();
  APPT_HIP(hipEventRecord(e1, s));
  APPT_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  APPT_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / (float)(reps > 0 ? reps : 1);
}

void Batch::autotune(int32_t reps, appt_tuning* chosen) {
  APPT_HIP(hipSetDevice(E_.device()));
  if (reps < 1) reps = 1;
  const bool graph = use_graph_;
  drop_graph();
  planned_ = false;
  // candidates: products classes x lanes-per-row rule x group-bin LDS budget
  static const int32_t kCls[3][3] = {{128, 512, 2048}, {64, 256, 1024}, {256, 1024, 4096}};
  static const int32_t kLen[3][3] = {{8, 16, 32}, {4, 8, 16}, {16, 32, 64}};
  static const int32_t kBudget[2] = {64 * 1024, 32 * 1024};
  const Config base = cfg_;
  Config best = base;
  float best_ms = -1.f;
  for (const auto& cls : kCls)
    for (const auto& len : kLen)
      for (int32_t budget : kBudget) {
        Config c = base;
        for (int k = 0; k < 3; k++) {
          c.sym.cls_max[k] = c.num.cls_max[k] = cls[k];
          c.sym.glen_max[k] = c.num.glen_max[k] = len[k];
        }
        c.sym.ncls = c.num.ncls = 3;
        c.sym.lds_budget = c.num.lds_budget = (int32_t)std::min<int64_t>(budget, E_.lds_max());
        cfg_ = c;
        run_planning();
        const float ms = time_eager(reps);
        if (best_ms < 0.f || ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
  cfg_ = best;
  plan(graph);
  if (chosen) *chosen = tuning_of(cfg_);
}

void Batch::detach_results(void** crp, void** ccol, void** cval) {
  *crp = crp_;
  *ccol = ccol_;
  *cval = cval_;
  for (auto& p : owned_)
    if (p == crp_) p = nullptr;
  ccol_ = nullptr;
  cval_ = nullptr;
  crp_ = nullptr;
  planned_ = false;
  drop_graph();
}

void Batch::result(int32_t k, appt_csr_result* out) const {
  if (!planned_) throw Error(APPT_ERR_INVALID, "batch is not planned");
  if (k < 0 || k >= K_) throw Error(APPT_ERR_INVALID, "batch index out of range");
  std::memset(out, 0, sizeof(*out));
  out->n = blk_row_h_[k + 1] - blk_row_h_[k];
  out->nnz = blkoff_h_[k + 1] - blkoff_h_[k];
  out->rowptr_is_64 = 0;
  out->rowptr = crp_ + blk_row_h_[k] + k;
  out->col = ccol_ + blkoff_h_[k];
  out->val = cval_ + blkoff_h_[k];
  out->_priv = nullptr;
}

void Batch::info(appt_batch_info* out) const {
  std::memset(out, 0, sizeof(*out));
  out->count = K_;
  out->n_total = N_;
  out->nnz_a = nnz_a_;
  out->nnz_c = planned_ ? plan_.nnz : 0;
  out->products = planned_ ? plan_.products : 0;
  out->graph = exec_ ? 1 : 0;
  out->bins_used = planned_ ? plan_.bins_used : 0;
  out->tuning = tuning_of(cfg_);
}

}  // namespace appt
