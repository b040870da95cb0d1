// engine.hpp -- host-side engine state (C++), behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <map>
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
// engine budget.
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

 private:
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> size_of_;
  int64_t allocated_ = 0;
  int64_t budget_;
};

// A pooled buffer that grows on demand and is kept across calls.
struct Scratch {
  void* p = nullptr;
  size_t cap = 0;
};

struct Workspace {
  Scratch bptr, bcol, bval, cursor, bmin, bmax, rowf, rowlo, rowhi, symbin, numbin, rowcnt, rowoff,
      symrows, numrows, scan_tmp, stats, rowwords, wlist, statpart, ctl, cur;
};

struct Result {  // what appt_csr_result::_priv points to
  void* rowptr = nullptr;
  void* col = nullptr;
  void* val = nullptr;
};

struct HostStats {  // pinned staging for the small device->host reads
  int32_t sym[kNumBins * kStatFields];
  int32_t num[kNumBins * kStatFields];
  unsigned long long products;
  int64_t total;
  int32_t cursor[kNumBins];
};

class Engine {
 public:
  explicit Engine(const appt_options* opt);
  ~Engine();

  void compute(const appt_csr_view& A, appt_op op, appt_csr_result* C);
  void release(appt_csr_result* C);
  void run_timed(const char* task, const appt_csr_view& A, int reps, appt_eng_result* out);
  void descriptors(const appt_csr_view& A, appt_op op, appt_descriptors* out);

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }
  Pool& pool() { return pool_; }
  void set_timing(bool t) { timing_ = t; }
  const appt_phase_times& times() const { return times_; }

  BinParams sym_params, num_params;

 private:
  void* ensure(Scratch& s, size_t bytes);
  void free_scratch(Scratch& s);
  void record(int idx);
  void validate(const appt_csr_view& A);
  // transpose (if op needs it), B-row spans, pre-pass and symbolic binning;
  // ends with one host synchronisation (bin statistics in hs_->sym).
  void prepass_stage(const appt_csr_view& A, appt_op op, const int32_t*& bp, const int32_t*& bc,
                     const double*& bv);
  int plan_bins(const appt_csr_view& A, appt_op op);

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  Pool pool_;
  Workspace ws_;
  HostStats* hs_ = nullptr;
  bool timing_ = false;
  hipEvent_t ev_[8] = {};
  appt_phase_times times_ = {};
  int64_t lds_max_ = 65536;
  int32_t heavy_cap_ = 0;  // column-window cap of the heavy-row bin (0: LDS-limited)
};

void set_lds_budget(int64_t b);
void set_last_error(const std::string& m);

}  // namespace appt
