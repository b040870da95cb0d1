// capi.cpp -- the extern "C" boundary.  Every entry point catches all C++
// exceptions and returns a status; nothing here exits the process (the
// reference's CK macro exits, src/spgemm_engine.cuh:18).
#include <cstring>
#include <functional>
#include <memory>
#include <string>

#include "engine.hpp"

namespace appt {
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace appt

using appt::Engine;
using appt::Error;

struct appt_engine {
  std::unique_ptr<Engine> impl;
};
struct appt_batch {
  std::unique_ptr<appt::Batch> impl;
};

template <typename F>
static appt_status guard(F&& f) {
  try {
    f();
    return APPT_OK;
  } catch (const Error& e) {
    appt::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    appt::set_last_error("host allocation failed");
    return APPT_ERR_OOM;
  } catch (const std::exception& e) {
    appt::set_last_error(e.what());
    return APPT_ERR_HIP;
  } catch (...) {
    appt::set_last_error("unknown error");
    return APPT_ERR_HIP;
  }
}

namespace appt {
appt_status guard_call(const std::function<void()>& f) { return guard(f); }
}  // namespace appt

extern "C" {

const char* appt_last_error(void) { return appt::g_last_error.c_str(); }

appt_status appt_op_from_task(const char* task, appt_op* out) {
  if (!task || !out) return APPT_ERR_INVALID;
  if (std::strcmp(task, "AA") == 0) { *out = APPT_OP_AA; return APPT_OK; }
  if (std::strcmp(task, "AAt") == 0) { *out = APPT_OP_AAT; return APPT_OK; }
  if (std::strcmp(task, "AAtS") == 0) { *out = APPT_OP_AATS; return APPT_OK; }
  appt::set_last_error(std::string("unknown task ") + task);
  return APPT_ERR_INVALID;
}

appt_status appt_engine_create(const appt_options* opt, appt_engine** out) {
  if (!out) return APPT_ERR_INVALID;
  *out = nullptr;
  return guard([&] {
    auto* e = new appt_engine;
    try {
      e->impl.reset(new Engine(opt));
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

appt_status appt_engine_destroy(appt_engine* eng) {
  if (!eng) return APPT_OK;
  return guard([&] { delete eng; });
}

appt_status appt_engine_set_timing(appt_engine* eng, int enable) {
  if (!eng) return APPT_ERR_INVALID;
  eng->impl->set_timing(enable != 0);
  return APPT_OK;
}

appt_status appt_engine_bin_report(const appt_engine* eng, appt_bin_report* out) {
  if (!eng || !out) return APPT_ERR_INVALID;
  *out = eng->impl->bin_report();
  return APPT_OK;
}

const char* appt_bin_name(int32_t bin) {
  static const char* const names[appt::kNumBins] = {
      "empty",
      "window/c0/G8", "window/c0/G16", "window/c0/G32", "window/c0/G64",
      "hash/c0/G8", "hash/c0/G16", "hash/c0/G32", "hash/c0/G64",
      "window/c1/G8", "window/c1/G16", "window/c1/G32", "window/c1/G64",
      "hash/c1/G8", "hash/c1/G16", "hash/c1/G32", "hash/c1/G64",
      "window/c2/G8", "window/c2/G16", "window/c2/G32", "window/c2/G64",
      "hash/c2/G8", "hash/c2/G16", "hash/c2/G32", "hash/c2/G64",
      "wg-window", "wg-hash", "heavy",
      "list/c0/G8", "list/c0/G16", "list/c0/G32", "list/c0/G64",
      "list/c1/G8", "list/c1/G16", "list/c1/G32", "list/c1/G64",
      "list/c2/G8", "list/c2/G16", "list/c2/G32", "list/c2/G64",
      "sort/N16", "sort/N32", "sort/N64", "sort/N128", "sort/N256", "sort/N512", "sort/N1024", "sort/N2048",
      "template"};
  static_assert(appt::kNumBins == 49, "bin names cover every bin id");
  if (bin < 0 || bin >= appt::kNumBins) return nullptr;
  return names[bin];
}

appt_status appt_engine_phase_times(const appt_engine* eng, appt_phase_times* out) {
  if (!eng || !out) return APPT_ERR_INVALID;
  *out = eng->impl->times();
  return APPT_OK;
}

appt_status appt_engine_trim(appt_engine* eng) {
  if (!eng) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->pool().trim(); });
}

int64_t appt_engine_device_bytes(const appt_engine* eng) { return eng ? eng->impl->pool().allocated() : 0; }

appt_status appt_spgemm_compute(appt_engine* eng, const appt_csr_view* A, appt_op op, appt_csr_result* C) {
  if (!eng || !A || !C) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->compute(*A, op, C); });
}

appt_status appt_spgemm_compute_noalloc(appt_engine* eng, const appt_csr_view* A, appt_op op, int32_t* rowptr,
                                        int32_t* col, double* val, int64_t capacity, int64_t* nnz_out) {
  if (!eng || !A || !nnz_out) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->compute_noalloc(*A, op, rowptr, col, val, capacity, nnz_out); });
}

appt_status appt_spgemm_compute_rows(appt_engine* eng, const appt_csr_view* A, appt_op op, int32_t row_begin,
                                     int32_t row_end, appt_csr_result* C) {
  if (!eng || !A || !C) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->compute_rows(*A, op, row_begin, row_end, C); });
}

appt_status appt_spgemm_symbolic_rowptr(appt_engine* eng, const appt_csr_view* A, appt_op op, int64_t* rowptr) {
  if (!eng || !A || !rowptr) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->symbolic_rowptr(*A, op, rowptr); });
}

appt_status appt_csr_eliminate_zeros(appt_engine* eng, appt_csr_result* C) {
  if (!eng || !C) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->eliminate_zeros(C); });
}

appt_status appt_csr_release(appt_engine* eng, appt_csr_result* C) {
  if (!eng || !C) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->release(C); });
}

appt_status appt_engine_run_timed(appt_engine* eng, const char* task, const appt_csr_view* A, int32_t reps,
                                  appt_eng_result* out) {
  if (!eng || !A || !out) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->run_timed(task, *A, reps, out); });
}

appt_status appt_spgemm_run_timed(const char* task, const int32_t* dOff, const int32_t* dCol, const double* dVal,
                                  int32_t n, int64_t nnz, int32_t reps, appt_eng_result* out) {
  if (!out) return APPT_ERR_INVALID;
  return guard([&] {
    // one engine per host thread, created on first use (the reference
    // allocates per call, src/spgemm_engine.cuh:136-181, outside its timing)
    static thread_local std::unique_ptr<Engine> eng;
    if (!eng) {
      int dev = 0;
      (void)hipGetDevice(&dev);
      appt_options o{};
      o.device = dev;
      eng.reset(new Engine(&o));
    }
    appt_csr_view A{n, nnz, dOff, dCol, dVal};
    eng->run_timed(task, A, reps, out);
  });
}

appt_status appt_dispatch_descriptors(appt_engine* eng, const appt_csr_view* A, appt_op op, appt_descriptors* out) {
  if (!eng || !A || !out) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->descriptors(*A, op, out); });
}

appt_status appt_batch_create(appt_engine* eng, int32_t count, const int32_t* n_each, const int64_t* nnz_each,
                              const appt_op* ops, appt_batch** out) {
  if (!eng || !out) return APPT_ERR_INVALID;
  *out = nullptr;
  return guard([&] {
    auto* b = new appt_batch;
    try {
      b->impl.reset(new appt::Batch(*eng->impl, count, n_each, nnz_each, ops));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

appt_status appt_batch_destroy(appt_batch* b) {
  if (!b) return APPT_OK;
  return guard([&] { delete b; });
}

appt_status appt_batch_inputs(appt_batch* b, appt_batch_buffers* out) {
  if (!b || !out) return APPT_ERR_INVALID;
  return guard([&] { b->impl->inputs(out); });
}

appt_status appt_batch_upload(appt_batch* b, const appt_host_csr* packed) {
  if (!b || !packed) return APPT_ERR_INVALID;
  return guard([&] { b->impl->upload(*packed); });
}

appt_status appt_batch_gather(appt_batch* b, const appt_csr_view* each) {
  if (!b) return APPT_ERR_INVALID;
  return guard([&] { b->impl->gather(each); });
}

appt_status appt_batch_plan(appt_batch* b, int32_t use_graph) {
  if (!b) return APPT_ERR_INVALID;
  return guard([&] { b->impl->plan(use_graph != 0); });
}

appt_status appt_batch_autotune(appt_batch* b, int32_t reps, appt_tuning* chosen) {
  if (!b) return APPT_ERR_INVALID;
  return guard([&] { b->impl->autotune(reps, chosen); });
}

appt_status appt_batch_execute(appt_batch* b) {
  if (!b) return APPT_ERR_INVALID;
  return guard([&] { b->impl->execute(); });
}

appt_status appt_batch_result(const appt_batch* b, int32_t k, appt_csr_result* out) {
  if (!b || !out) return APPT_ERR_INVALID;
  return guard([&] { b->impl->result(k, out); });
}

appt_status appt_batch_info_get(const appt_batch* b, appt_batch_info* out) {
  if (!b || !out) return APPT_ERR_INVALID;
  return guard([&] { b->impl->info(out); });
}

appt_status appt_csr_check(appt_engine* eng, const appt_csr_view* A) {
  if (!eng || !A) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->check(*A); });
}

appt_status appt_device_upload(appt_engine* eng, const appt_host_csr* h, appt_csr_view* v) {
  if (!eng || !h || !v) return APPT_ERR_INVALID;
  if (!appt::host_csr_valid(*h)) {
    appt::set_last_error("host CSR structure invalid (row pointers or column indices)");
    return APPT_ERR_INVALID;
  }
  return guard([&] {
    APPT_HIP(hipSetDevice(eng->impl->device()));
    int32_t *rp = nullptr, *c = nullptr;
    double* x = nullptr;
    APPT_HIP(hipMalloc(&rp, sizeof(int32_t) * ((size_t)h->n + 1)));
    APPT_HIP(hipMalloc(&c, sizeof(int32_t) * (size_t)(h->nnz > 0 ? h->nnz : 1)));
    APPT_HIP(hipMalloc(&x, sizeof(double) * (size_t)(h->nnz > 0 ? h->nnz : 1)));
    APPT_HIP(hipMemcpy(rp, h->rowptr, sizeof(int32_t) * ((size_t)h->n + 1), hipMemcpyHostToDevice));
    if (h->nnz > 0) {
      APPT_HIP(hipMemcpy(c, h->col, sizeof(int32_t) * (size_t)h->nnz, hipMemcpyHostToDevice));
      APPT_HIP(hipMemcpy(x, h->val, sizeof(double) * (size_t)h->nnz, hipMemcpyHostToDevice));
    }
    v->n = h->n;
    v->nnz = h->nnz;
    v->rowptr = rp;
    v->col = c;
    v->val = x;
  });
}

appt_status appt_device_free_view(appt_engine* eng, appt_csr_view* v) {
  if (!eng || !v) return APPT_ERR_INVALID;
  return guard([&] {
    APPT_HIP(hipSetDevice(eng->impl->device()));
    (void)hipFree((void*)v->rowptr);
    (void)hipFree((void*)v->col);
    (void)hipFree((void*)v->val);
    std::memset(v, 0, sizeof(*v));
  });
}

appt_status appt_csr_download_rows(appt_engine* eng, const appt_csr_result* C, int32_t row_begin, int32_t row_end,
                                   int64_t* rowptr, int32_t* col, double* val) {
  if (!C || !rowptr) return APPT_ERR_INVALID;
  if (row_begin < 0 || row_end < row_begin || row_end > C->n) return APPT_ERR_INVALID;
  return guard([&] {
    if (eng) APPT_HIP(hipSetDevice(eng->impl->device()));
    const size_t n1 = (size_t)(row_end - row_begin) + 1;
    if (C->rowptr_is_64) {
      APPT_HIP(hipMemcpy(rowptr, (const int64_t*)C->rowptr + row_begin, sizeof(int64_t) * n1, hipMemcpyDeviceToHost));
    } else {
      std::unique_ptr<int32_t[]> t(new int32_t[n1]);
      APPT_HIP(hipMemcpy(t.get(), (const int32_t*)C->rowptr + row_begin, sizeof(int32_t) * n1, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n1; i++) rowptr[i] = t[i];
    }
    const int64_t base = rowptr[0], cnt = rowptr[n1 - 1] - rowptr[0];
    for (size_t i = 0; i < n1; i++) rowptr[i] -= base;
    if (col && cnt) APPT_HIP(hipMemcpy(col, C->col + base, sizeof(int32_t) * (size_t)cnt, hipMemcpyDeviceToHost));
    if (val && cnt) APPT_HIP(hipMemcpy(val, C->val + base, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost));
  });
}

appt_status appt_csr_digest(appt_engine* eng, const appt_csr_result* C, uint64_t* out) {
  if (!eng || !C || !out) return APPT_ERR_INVALID;
  return guard([&] { eng->impl->digest(*C, out); });
}

appt_status appt_csr_download(appt_engine* eng, const appt_csr_result* C, int64_t* rowptr, int32_t* col,
                              double* val) {
  if (!C) return APPT_ERR_INVALID;
  return guard([&] {
    if (eng) APPT_HIP(hipSetDevice(eng->impl->device()));
    const size_t n1 = (size_t)C->n + 1;
    if (rowptr) {
      if (C->rowptr_is_64) {
        APPT_HIP(hipMemcpy(rowptr, C->rowptr, sizeof(int64_t) * n1, hipMemcpyDeviceToHost));
      } else {
        std::unique_ptr<int32_t[]> t(new int32_t[n1]);
        APPT_HIP(hipMemcpy(t.get(), C->rowptr, sizeof(int32_t) * n1, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n1; i++) rowptr[i] = t[i];
      }
    }
    if (col && C->nnz) APPT_HIP(hipMemcpy(col, C->col, sizeof(int32_t) * (size_t)C->nnz, hipMemcpyDeviceToHost));
    if (val && C->nnz) APPT_HIP(hipMemcpy(val, C->val, sizeof(double) * (size_t)C->nnz, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
