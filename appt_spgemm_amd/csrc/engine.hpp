// engine.hpp -- host-side engine state (C++), behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <atomic>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/appt_spgemm.h"
#include "kernels.hpp"

namespace appt {

struct Error : std::runtime_error {
  appt_status code;
  Error(appt_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define APPT_HIP(x)                                                                                      \
  do {                                                                                                   \
    hipError_t e_ = (x);                                                                                 \
    if (e_ != hipSuccess)                                                                                \
      throw ::appt::Error(e_ == hipErrorOutOfMemory ? APPT_ERR_OOM : APPT_ERR_HIP,                      \
                          std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" +       \
                          std::to_string(__LINE__) + ")");                                              \
  } while (0)

// Device memory pool: exact-capacity buffers kept on a free list and reused
// when a request fits in [bytes, 2*bytes].  Accounts every byte against the
// engine budget.  Thread-safe (a server releases results from other threads).
// Device bytes held by a set of pools (a server's slots), now and at peak.
struct PoolTally {
  std::atomic<int64_t> now{0}, peak{0};
  void add(int64_t b) {
    const int64_t v = now.fetch_add(b) + b;
    int64_t p = peak.load();
    while (v > p && !peak.compare_exchange_weak(p, v)) {
    }
  }
};

class Pool {
 public:
  explicit Pool(int64_t budget) : budget_(budget) {}
  ~Pool();
  void* get(size_t bytes);
  void put(void* p);
  void trim();
  int64_t allocated() const { return allocated_; }
  int64_t budget() const { return budget_; }
  void set_budget(int64_t b) { budget_ = b; }
  // count this pool's hipMalloc'd bytes (already held ones included) in t
  void set_tally(PoolTally* t);

 private:
  void trim_locked();
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> size_of_;
  std::atomic<int64_t> allocated_{0};
  int64_t budget_;
  PoolTally* tally_ = nullptr;
};

// A pooled buffer that grows on demand and is kept across calls.
struct Scratch {
  void* p = nullptr;
  size_t cap = 0;
};

// Bin thresholds and launch shapes of one pipeline: the engine's defaults,
// or a batch's own (autotuned) copy.
struct Config {
  BinParams sym{}, num{};
  int32_t heavy_cap = 0;  // column-window cap of the heavy-row bin (0: as wide as LDS allows)
  int64_t radix_min_nnz = 1 << 18;  // asymmetric transposes of at least this many entries use the radix sort
  int64_t radix_mirror_min_nnz = 1 << 20;  // AAtS mirrors of results this large use the radix sort
  int64_t rowptr64_min_nnz = (int64_t)INT32_MAX + 1;  // results this large get int64 row pointers
  // row templates (templates.hip) for rows sharing a relative pattern, looked
  // for when A has at least this many entries (small calls would only pay
  // for the detection)
  bool templates = true;
  int64_t templates_min_nnz = (int64_t)1 << 18;
  // AAtS: 1 / 2 = Gram over the upper / lower triangle (gram.hip), otherwise
  // full rows (direct, or a mirror where rows accumulate with atomics)
  int32_t gram = 0;
  bool rowptr64(int64_t nnz) const { return nnz > INT32_MAX || nnz >= rowptr64_min_nnz; }
};

// Device scratch of one pipeline.  Buffers only grow, so a HIP graph that
// captured their addresses stays valid while the owner lives.
struct Workspace {
  Scratch bptr, bcol, bval, cursor, bmeta, rowf, rowlo, rowhi, symbin, numbin, rowcnt, rowoff, symrows,
      numrows, scan_tmp, stats, rowwords, wlist, statpart, ctl, cur, sortlist, rkeys, rrows, rvals, rkeys2,
      rhist, rhoff, rtmp, rvals2, mlow, mup, mlo_off, mup_off, wgbits, wgslot, plancnt, planwords, digest,
      ckey, crep, cpat, akey, arep, bclass, aclass, tkey, trep, tplid, tinfo, tposoff, tpos, tsoff, ttuple, tpos2, tmatched, tgrid,
      vmeta, vac, ucol, uval, rrows2, goff, gtot;
  // words per row of the stored workgroup-bin bitmaps of the current call
  // (0: none stored)
  int32_t wg_stride = 0;
  // row templates of the current call (tpl_on: detection ran)
  TplArgs tpl{};
  bool tpl_on = false;
  template <typename F>
  void each(F&& f) {
    Scratch* all[] = {&bptr,   &bcol,    &bval,    &cursor,   &bmeta, &rowf,     &rowlo,    &rowhi, &symbin, &numbin,
                      &rowcnt, &rowoff,  &symrows, &numrows,  &scan_tmp, &stats, &rowwords, &wlist, &statpart, &ctl,
                      &cur,    &sortlist, &rkeys,  &rrows,    &rvals, &rkeys2, &rhist,    &rhoff, &rtmp,
                      &rvals2, &mlow,     &mup,    &mlo_off,  &mup_off, &wgbits, &wgslot, &plancnt, &planwords,
                      &digest, &ckey,    &crep,  &cpat,  &akey,   &arep,   &bclass,  &aclass, &tkey,  &trep,   &tplid,
                      &tinfo,  &tposoff, &tpos,  &tsoff,  &ttuple, &tpos2, &tmatched, &tgrid, &vmeta, &vac,
                      &ucol,   &uval,    &rrows2, &goff,  &gtot};
    for (Scratch* s : all) f(*s);
  }
};

// Operands of one pipeline run.  A side: output row i owns entries p in
// [ap[i], ap[i+1]) with B-row id ac[p] and value av[p].  B side: row k is
// bc/bv[bmeta[k].x, bmeta[k].y), bmeta built by the operand stage.
struct Operands {
  int32_t n = 0;   // output rows (A-side rows; a panel of A's rows)
  int32_t nb = 0;  // B rows
  int64_t annz = 0;  // entries addressed by A-side positions (cursor array size)
  const int32_t* ap = nullptr;
  const int32_t* ac = nullptr;
  const double* av = nullptr;
  const int32_t* bc = nullptr;
  const double* bv = nullptr;
  int64_t bnnz = 0;
  int upper = 0;  // AAtS: values formed for j >= i, lower part mirrored
  int32_t ncols = 0;  // bound on output column indices (a batch's are local to each block)
  int32_t row0 = 0;   // global row index of A-side row 0 (row panels)
  bool a_is_b = false;  // the A-side rows have B's rows' structure (A*A, or A*A^T of a symmetric pattern)
  bool gram = false;    // AAtS as a triangle-only Gram: B rows are per-entry column ranges (gram.hip)
  bool gram_lower = false;  // the Gram's triangle: j <= i (else j >= i)
  // B's row pointers; bmeta_pending: the B-row metadata is left to the
  // pre-pass (the template class pass writes it from the same loads)
  const int32_t* bptr = nullptr;
  bool bmeta_pending = false;
};

// Bin statistics read back at the two host synchronisations; they fix the
// launch shapes of every later replay of the same structure.
struct Plan {
  int32_t sym[kNumBins * kStatFields] = {};
  int32_t num[kNumBins * kStatFields] = {};
  int64_t products = 0;
  int64_t nnz = 0;
  int bins_used = 0;
  bool a_strict = false;    // A's rows strictly increasing (read for AAtS)
  bool direct_sym = false;  // AAtS: lower entries computed directly, no mirror
  int64_t gram_products = 0;  // AAtS as a Gram: products of the full A*A^T (products counts U's)
};

struct Result {  // what appt_csr_result::_priv points to
  void* rowptr = nullptr;
  void* col = nullptr;
  void* val = nullptr;
};

// Control words of a pipeline run (zeroed by Engine::reset_ctl):
// [0..1] heavy-row claim counters (symbolic, numeric); [2] B rows unsorted;
// [3] the symmetric-pattern transpose gather failed; [4] A rows not strictly
// increasing; [5..7] structure check failures (ends, row pointer order,
// columns); [8] a heavy numeric window could not advance; [10] a B row longer than the
// template class pass describes; [12..13] AAtS as
// an upper-triangle Gram: the full A*A^T product count (u64).
constexpr int kCtlWords = 16;
constexpr int kCtlHeavyFault = 8;
constexpr int kCtlLongRows = 10;   // some B row has more than kTplMaxLen entries (template class pass)
constexpr int kCtlGramProducts = 12;

struct HostStats {  // pinned staging for the small device->host reads; the
                    // fields up to ctl mirror the device stats layout
  int32_t sym[kNumBins * kStatFields];
  int32_t num[kNumBins * kStatFields];
  unsigned long long products;
  int64_t total;
  int32_t ctl[kCtlWords];
};
static_assert(offsetof(HostStats, products) == sizeof(int32_t) * 2 * kNumBins * kStatFields &&
                  offsetof(HostStats, total) == offsetof(HostStats, products) + 8,
              "HostStats mirrors the device stats layout");

class Engine {
 public:
  explicit Engine(const appt_options* opt);
  ~Engine();

  void compute(const appt_csr_view& A, appt_op op, appt_csr_result* C);
  // Rows [r0, r1) of op(A) (a row panel): C has r1 - r0 rows, global columns.
  // use_plan: heavy rows may take their counts from the last symbolic_rowptr
  // of the same view and op (the panel API; compute() never does).
  void compute_rows(const appt_csr_view& A, appt_op op, int32_t r0, int32_t r1, appt_csr_result* C,
                    bool use_plan = true);
  // Exact row pointers of op(A) (symbolic phase only), copied to the host.
  void symbolic_rowptr(const appt_csr_view& A, appt_op op, int64_t* rowptr_host);
  // C = op(A) into caller-owned device buffers (int32 row pointers, n + 1;
  // col / val of `capacity` entries), with every scratch buffer taken from
  // the device's static scratch array: no device allocation.  *nnz_out
  // receives nnz(C); APPT_ERR_OVERFLOW (nothing written to col / val) when it
  // exceeds capacity.
  void compute_noalloc(const appt_csr_view& A, appt_op op, int32_t* rowptr, int32_t* col, double* val,
                       int64_t capacity, int64_t* nnz_out);
  void release(appt_csr_result* C);
  // Drop entries whose value is exactly zero (numeric semantics); C must be
  // an engine-owned result.
  void eliminate_zeros(appt_csr_result* C);
  // Structure and value digests of a device result (see launch_digest).
  void digest(const appt_csr_result& C, uint64_t out[2]);
  // Device-side structure check of A (throws APPT_ERR_INVALID).
  void check(const appt_csr_view& A);
  void run_timed(const char* task, const appt_csr_view& A, int reps, appt_eng_result* out);
  void descriptors(const appt_csr_view& A, appt_op op, appt_descriptors* out);

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }
  Pool& pool() { return pool_; }
  void set_timing(bool t) { timing_ = t; }
  const appt_phase_times& times() const { return times_; }
  const appt_bin_report& bin_report() const { return bins_; }
  const Config& config() const { return cfg_; }
  void set_config(const Config& c) { cfg_ = c; }
  int64_t lds_max() const { return lds_max_; }
  static void validate(const appt_csr_view& A);

  // Scratch buffer of at least `bytes`: from the pool, or, inside
  // compute_noalloc, carved from the static scratch array.
  void* ensure(Scratch& s, size_t bytes);
  void free_workspace(Workspace& W);
  // Called with device bytes a call is about to allocate, before it does:
  // its working set (result = false: before the operands are staged, bounded
  // by work_bound; again before the symbolic pass keeps workgroup bitmaps),
  // then C itself (result = true, the exact bytes once the symbolic phase has
  // sized it).  May block, or throw to refuse the call.
  using Admission = std::function<void(int64_t bytes, bool result)>;
  void set_admission(Admission f) { admit_ = std::move(f); }
  const Admission& admission() const { return admit_; }
  // Upper bound of the device memory one call on `rows` output rows of an
  // operand with nb rows and nnz entries may newly take from the pool, C
  // excluded (the workgroup bitmaps are admitted separately): every scratch
  // buffer at its growth size, the transpose and its radix scratch, the row
  // templates' tables, the int32 row pointers.
  static int64_t work_bound(int32_t rows, int32_t nb, int64_t nnz);
  // Return the workspace and the pool's free blocks to HIP.
  void trim_all();

  // ---- pipeline stages: kernel launches only.  No host synchronisation and,
  // once the workspace has grown to the problem, no allocation, so a replay
  // of the same structure can be captured into a HIP graph.
  // B = A or A^T for one matrix, and the B-row metadata.
  void stage_operands(Workspace& W, const appt_csr_view& A, appt_op op, Operands& O);
  // Control words [0] / [1] heavy-row claim counters, [2] B-unsorted flag,
  // [3] asymmetric-pattern flag of the transpose; zeroed here.
  int32_t* reset_ctl(Workspace& W);
  // B-row metadata from B's row starts/ends: row k is [ptr[k], ptr[k+1]) or,
  // where sel[k] != 0, [off2 + ptr2[k], off2 + ptr2[k+1]).
  void stage_bmeta(Workspace& W, const Operands& O, const int32_t* ptr, const int32_t* ptr2, const uint8_t* sel,
                   int64_t off2);
  void stage_prepass(Workspace& W, const Operands& O, const Config& cfg);
  void stage_symbolic(Workspace& W, const Operands& O, const Config& cfg, const Plan& P);
  void stage_offsets(Workspace& W, const Operands& O, const Config& cfg, int32_t* rp32);
  void stage_numeric(Workspace& W, const Operands& O, const Config& cfg, const Plan& P, int32_t* ccol,
                     double* cval);
  void stage_mirror(Workspace& W, const Operands& O, const Plan& P, const int32_t* ccol, double* cval);
  // AAtS as an upper-triangle Gram (gram.hip): B rows become per-entry
  // suffixes (j >= i) of the transposed operand ...
  void stage_gram_operand(Workspace& W, Operands& O, bool lower);
  // AAtS: full rows, or a Gram over the upper / lower triangle
  enum { kGramNone = 0, kGramUpper = 1, kGramLower = 2 };
  int gram_mode(const appt_csr_view& A) const;
  // ... and U (ucol / uval, offsets in W.rowoff) expands into the full C
  // (nc entries; final offsets in W.goff and rp32)
  void stage_gram_expand(Workspace& W, const Operands& O, const Plan& P, const int32_t* ucol, const double* uval,
                         int64_t nc, int32_t* rp32, int32_t* ccol, double* cval);
  // host synchronisations: bin statistics of the symbolic / numeric binning
  void read_sym(Workspace& W, Plan& P, bool want_ctl = false);
  void read_num(Workspace& W, const Operands& O, Plan& P);

 private:
  void record(int idx);
  HeavyArgs heavy_args(Workspace& W, const Operands& O, const Config& cfg);

  // Bump allocator over the static scratch array (compute_noalloc only).
  struct Arena {
    unsigned char* base = nullptr;
    size_t cap = 0, used = 0;
  };
  Arena* arena_ = nullptr;
  Workspace ws_static_;  // scratch views into the static array, reset every call

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  Pool pool_;
  Workspace ws_;
  // Row-pointer plan of the last symbolic_rowptr call: the A view and op it
  // was made for, and its per-row counts and distinct words (ws_.plancnt,
  // ws_.planwords).  compute_rows on the same view and op takes its heavy
  // rows' symbolic results from it (heavy_plan_ set for that call only).
  struct RowPlanKey {
    const void* rowptr = nullptr;
    const void* col = nullptr;
    const void* val = nullptr;
    int32_t n = -1;
    int64_t nnz = -1;
    int op = -1;
  } plan_key_;
  const int32_t* heavy_plan_cnt_ = nullptr;
  const int32_t* heavy_plan_words_ = nullptr;
  int32_t heavy_plan_r0_ = 0;
  Config cfg_;
  HostStats* hs_ = nullptr;
  bool timing_ = false;
  hipEvent_t ev_[8] = {};
  hipEvent_t ev_wait_ = nullptr;  // completion marker the host polls at a synchronisation
  // Wait for everything enqueued on the engine's stream: poll a marker
  // event (the engine's waits are short), blocking after about 2 ms.
  void wait_stream();
  appt_phase_times times_ = {};
  appt_bin_report bins_ = {};
  int64_t lds_max_ = 65536;
  Admission admit_;
};

Config make_config(const appt_tuning* t, int64_t lds_max);
// Device bytes of a CSR result with n rows and nnz entries.
int64_t result_bytes(int32_t n, int64_t nnz);
// Square CSR structure check of a host matrix (gen.cpp).
bool host_csr_valid(const appt_host_csr& m);
appt_tuning tuning_of(const Config& c);

// Many independent matrices as one block-diagonal product (A*A or A*A^T per
// matrix): one pipeline over the packed rows, planned once with the two
// synchronisations, then replayed from a captured HIP graph.
class Batch {
 public:
  Batch(Engine& e, int32_t count, const int32_t* n_each, const int64_t* nnz_each, const appt_op* ops);
  ~Batch();
  void inputs(appt_batch_buffers* out) const;
  void upload(const appt_host_csr& packed);
  void gather(const appt_csr_view* each);
  void plan(bool use_graph);
  void execute();
  void result(int32_t k, appt_csr_result* out) const;
  // Hand the C buffers (row pointers, columns, values) to the caller, who
  // returns them to the engine's pool; the batch must be planned again.
  void detach_results(void** crp, void** ccol, void** cval);
  void autotune(int32_t reps, appt_tuning* chosen);
  void info(appt_batch_info* out) const;
  int64_t products(int32_t k) const { return blkprod_h_.at((size_t)k); }

 private:
  void* alloc(size_t bytes);
  void operands();           // kernels: B-row ids, transposes of A*A^T blocks, B-row metadata
  void stages();             // kernels: operands .. numeric, per-matrix row pointers
  void run_planning();       // stages with the synchronisations; sizes C
  void drop_graph();
  float time_eager(int reps);

  Engine& E_;
  Config cfg_;
  Workspace W_;
  Plan plan_{};
  Operands O_{};
  int32_t K_ = 0, N_ = 0;
  int32_t max_n_ = 0;   // largest block: bounds the (local) output columns
  int32_t max_nt_ = 0;  // largest A*A^T block
  int64_t nnz_a_ = 0, nnz_t_ = 0;
  std::vector<int32_t> blk_row_h_;
  std::vector<int64_t> blk_nnz_h_, blkoff_h_, blkprod_h_;
  std::vector<void*> owned_;
  // device
  int32_t *in_rp_ = nullptr, *in_col_ = nullptr, *blk_row_ = nullptr, *rowbase_ = nullptr, *rowblk_ = nullptr;
  int32_t *acg_ = nullptr, *tcnt_ = nullptr, *tp32_ = nullptr, *tcur_ = nullptr, *rp32_ = nullptr, *crp_ = nullptr;
  int32_t* blk_toff_ = nullptr;  // per block: its A*A^T entries' offset in the transposed region
  int32_t* ccol_ = nullptr;
  double *in_val_ = nullptr, *cval_ = nullptr;
  int64_t *blk_nnz_ = nullptr, *toff_ = nullptr, *tscan_ = nullptr, *blkoff_ = nullptr, *blkprod_ = nullptr;
  uint8_t *blk_op_ = nullptr, *rowop_ = nullptr;
  BatchSrc* src_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  bool planned_ = false;
  bool use_graph_ = true;
};

void set_last_error(const std::string& m);

}  // namespace appt
