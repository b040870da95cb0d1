// server.cpp -- request loop with memory admission and continuous batching.
//
// Worker threads, one per slot, each own an Engine (own stream, own pool and
// scratch).  A worker takes the oldest queued job; when it is small, it also
// takes the other small jobs waiting at that moment and runs them together
// as one block-diagonal batch.  Admission happens inside the run, at the
// point where the symbolic phase has fixed nnz(C): the exact result bytes
// plus a working-set estimate are reserved against the budget, waiting while
// other jobs may still return memory, refusing when the job cannot fit.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace appt {

using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Working memory of one job beside its result: transpose, per-row metadata,
// bin lists, cursors (bytes; generous).
static int64_t work_bytes(int32_t n, int64_t nnz) { return 64 * (int64_t)n + 28 * nnz + (8 << 20); }

class Server {
 public:
  explicit Server(const appt_server_options& o);
  ~Server();
  uint64_t submit(const appt_csr_view& A, appt_op op);
  void poll(uint64_t id, appt_job_info* out);
  void wait(uint64_t id, appt_job_info* out);
  void release(uint64_t id);
  void stats(appt_server_stats* out);

 private:
  struct Job {
    appt_csr_view A{};
    appt_op op = APPT_OP_AA;
    appt_job_state state = APPT_JOB_QUEUED;
    appt_status status = APPT_OK;
    int group = 1;
    Clock::time_point t_submit, t_start, t_done;
    int64_t products = 0;
    int64_t reserved = 0;  // result bytes held until release
    appt_csr_result C{};
    int slot = -1;
  };
  struct Slot {
    std::unique_ptr<Engine> eng;
    std::thread th;
  };

  bool small(const Job& j) const { return j.A.n <= row_limit_ && max_group_ > 1; }
  void worker(int slot);
  void run_one(int slot, uint64_t id);
  void run_group(int slot, const std::vector<uint64_t>& ids);
  // reserve `need` bytes against the budget, waiting for memory to return;
  // throws APPT_ERR_ADMISSION when refused
  void reserve(int64_t need);
  void unreserve(int64_t bytes);
  void finish(uint64_t id, appt_status st);
  void fill(const Job& j, appt_job_info* out) const;

  int device_;
  int64_t budget_;
  int max_group_;
  int32_t row_limit_;
  int wait_ms_;
  std::vector<Slot> slots_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_, cv_mem_;
  std::deque<uint64_t> queue_;
  std::unordered_map<uint64_t, Job> jobs_;
  uint64_t next_id_ = 1;
  bool stop_ = false;
  int running_ = 0;
  int64_t reserved_ = 0, peak_reserved_ = 0;
  int64_t submitted_ = 0, completed_ = 0, refused_ = 0, failed_ = 0, groups_ = 0;
  std::vector<double> latencies_;
  double flops_done_ = 0.0;
  bool any_submit_ = false;
  Clock::time_point first_submit_, last_done_;
};

Server::Server(const appt_server_options& o)
    : device_(o.device),
      budget_(o.hbm_budget_bytes > 0 ? o.hbm_budget_bytes : (int64_t)200 * 1000 * 1000 * 1000),
      max_group_(o.batch_max_jobs > 0 ? o.batch_max_jobs : 64),
      row_limit_(o.batch_row_limit > 0 ? o.batch_row_limit : 65536),
      wait_ms_(o.admission_wait_ms > 0 ? o.admission_wait_ms : 10000) {
  const int nslots = o.slots > 0 ? o.slots : 4;
  slots_.resize((size_t)nslots);
  for (int k = 0; k < nslots; k++) {
    appt_options eo{};
    eo.device = device_;
    eo.tuning = o.tuning;
    slots_[k].eng.reset(new Engine(&eo));
  }
  for (int k = 0; k < nslots; k++) slots_[k].th = std::thread([this, k] { worker(k); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
    for (uint64_t id : queue_) {
      Job& j = jobs_[id];
      j.state = APPT_JOB_DONE;
      j.status = APPT_ERR_INVALID;
      j.t_done = Clock::now();
    }
    queue_.clear();
  }
  cv_work_.notify_all();
  cv_mem_.notify_all();
  for (auto& s : slots_)
    if (s.th.joinable()) s.th.join();
  for (auto& kv : jobs_)
    if (kv.second.C._priv && kv.second.slot >= 0) slots_[kv.second.slot].eng->release(&kv.second.C);
  jobs_.clear();
}

uint64_t Server::submit(const appt_csr_view& A, appt_op op) {
  Engine::validate(A);
  if (op != APPT_OP_AA && op != APPT_OP_AAT) throw Error(APPT_ERR_UNSUPPORTED, "served ops are A*A and A*A^T");
  std::lock_guard<std::mutex> lk(mu_);
  if (stop_) throw Error(APPT_ERR_INVALID, "server is stopping");
  const uint64_t id = next_id_++;
  Job& j = jobs_[id];
  j.A = A;
  j.op = op;
  j.t_submit = Clock::now();
  if (!any_submit_) {
    any_submit_ = true;
    first_submit_ = j.t_submit;
  }
  submitted_++;
  queue_.push_back(id);
  cv_work_.notify_one();
  return id;
}

void Server::reserve(int64_t need) {
  std::unique_lock<std::mutex> lk(mu_);
  if (need > budget_)
    throw Error(APPT_ERR_ADMISSION, "job needs " + std::to_string(need) + " B, budget " + std::to_string(budget_) + " B");
  const auto deadline = Clock::now() + std::chrono::milliseconds(wait_ms_);
  while (reserved_ + need > budget_) {
    // memory returns when running jobs end or results are released
    if (cv_mem_.wait_until(lk, deadline) == std::cv_status::timeout && reserved_ + need > budget_)
      throw Error(APPT_ERR_ADMISSION, "no room within the HBM budget after waiting " + std::to_string(wait_ms_) +
                                          " ms (" + std::to_string(reserved_) + " B held)");
    if (stop_) throw Error(APPT_ERR_ADMISSION, "server is stopping");
  }
  reserved_ += need;
  peak_reserved_ = std::max(peak_reserved_, reserved_);
}

void Server::unreserve(int64_t bytes) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    reserved_ -= bytes;
  }
  cv_mem_.notify_all();
}

void Server::finish(uint64_t id, appt_status st) {
  Job& j = jobs_[id];
  j.state = APPT_JOB_DONE;
  j.status = st;
  j.t_done = Clock::now();
  if (st == APPT_OK) {
    completed_++;
    latencies_.push_back(ms_between(j.t_submit, j.t_done));
    flops_done_ += 2.0 * (double)j.products;
    last_done_ = std::max(last_done_, j.t_done);
  } else if (st == APPT_ERR_ADMISSION) {
    refused_++;
  } else {
    failed_++;
  }
}

void Server::worker(int slot) {
  (void)hipSetDevice(device_);
  while (true) {
    std::vector<uint64_t> ids;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;  // stopping
      const uint64_t first = queue_.front();
      queue_.pop_front();
      ids.push_back(first);
      if (small(jobs_[first])) {
        // continuous batching: every small job waiting now joins this launch
        for (auto it = queue_.begin(); it != queue_.end() && (int)ids.size() < max_group_;) {
          if (small(jobs_[*it])) {
            ids.push_back(*it);
            it = queue_.erase(it);
          } else {
            ++it;
          }
        }
      }
      const auto now = Clock::now();
      for (uint64_t id : ids) {
        Job& j = jobs_[id];
        j.state = APPT_JOB_RUNNING;
        j.t_start = now;
        j.group = (int)ids.size();
        j.slot = slot;
      }
      running_++;
      if (ids.size() > 1) groups_++;
    }
    if (ids.size() == 1) run_one(slot, ids[0]);
    else run_group(slot, ids);
    {
      std::lock_guard<std::mutex> lk(mu_);
      running_--;
    }
    cv_done_.notify_all();
    cv_mem_.notify_all();
  }
}

void Server::run_one(int slot, uint64_t id) {
  Engine& eng = *slots_[slot].eng;
  appt_csr_view A;
  appt_op op;
  {
    std::lock_guard<std::mutex> lk(mu_);
    A = jobs_[id].A;
    op = jobs_[id].op;
  }
  const int64_t work = work_bytes(A.n, A.nnz);
  int64_t held = 0;
  appt_csr_result C{};
  appt_status st = APPT_OK;
  eng.set_admission([&](int64_t cbytes) {
    reserve(cbytes + work);
    held = cbytes + work;
  });
  try {
    eng.compute(A, op, &C);
  } catch (const Error& e) {
    st = e.code;
  } catch (...) {
    st = APPT_ERR_HIP;
  }
  eng.set_admission(nullptr);
  const int64_t cbytes = st == APPT_OK ? result_bytes(C.n, C.nnz) : 0;
  if (held > cbytes) unreserve(held - cbytes);
  std::lock_guard<std::mutex> lk(mu_);
  Job& j = jobs_[id];
  if (st == APPT_OK) {
    j.C = C;
    j.reserved = cbytes;
    j.products = eng.times().products;
  }
  finish(id, st);
}

void Server::run_group(int slot, const std::vector<uint64_t>& ids) {
  Engine& eng = *slots_[slot].eng;
  const int K = (int)ids.size();
  std::vector<appt_csr_view> views((size_t)K);
  std::vector<int32_t> ns((size_t)K);
  std::vector<int64_t> nnzs((size_t)K);
  std::vector<appt_op> ops((size_t)K);
  int64_t work = 0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (int k = 0; k < K; k++) {
      const Job& j = jobs_[ids[k]];
      views[k] = j.A;
      ns[k] = j.A.n;
      nnzs[k] = j.A.nnz;
      ops[k] = j.op;
      work += work_bytes(j.A.n, j.A.nnz);
    }
  }
  int64_t held = 0;
  appt_status st = APPT_OK;
  std::vector<appt_csr_result> out((size_t)K);
  std::vector<int64_t> cb((size_t)K, 0);
  std::vector<int64_t> prods((size_t)K, 0);
  // the group's C is computed in one buffer, then copied into one result per
  // job (so each job is released on its own): reserve both while running
  eng.set_admission([&](int64_t cbytes) {
    reserve(2 * cbytes + work);
    held = 2 * cbytes + work;
  });
  try {
    Batch b(eng, K, ns.data(), nnzs.data(), ops.data());
    b.gather(views.data());
    b.plan(false);
    hipStream_t s = eng.stream();
    for (int k = 0; k < K; k++) {
      appt_csr_result v{};
      b.result(k, &v);
      appt_csr_result& r = out[k];
      r.n = v.n;
      r.nnz = v.nnz;
      r.rowptr_is_64 = 0;
      r.rowptr = eng.pool().get(sizeof(int32_t) * ((size_t)v.n + 1));
      r.col = static_cast<int32_t*>(eng.pool().get(sizeof(int32_t) * (size_t)std::max<int64_t>(v.nnz, 1)));
      r.val = static_cast<double*>(eng.pool().get(sizeof(double) * (size_t)std::max<int64_t>(v.nnz, 1)));
      r._priv = new Result{r.rowptr, r.col, r.val};
      APPT_HIP(hipMemcpyAsync(r.rowptr, v.rowptr, sizeof(int32_t) * ((size_t)v.n + 1), hipMemcpyDeviceToDevice, s));
      if (v.nnz > 0) {
        APPT_HIP(hipMemcpyAsync(r.col, v.col, sizeof(int32_t) * (size_t)v.nnz, hipMemcpyDeviceToDevice, s));
        APPT_HIP(hipMemcpyAsync(r.val, v.val, sizeof(double) * (size_t)v.nnz, hipMemcpyDeviceToDevice, s));
      }
      cb[k] = result_bytes(v.n, v.nnz);
    }
    APPT_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < K; k++) prods[k] = b.products(k);
  } catch (const Error& e) {
    st = e.code;
  } catch (...) {
    st = APPT_ERR_HIP;
  }
  eng.set_admission(nullptr);
  if (st == APPT_ERR_ADMISSION) {
    // the group does not fit as a whole: admit its jobs one by one
    for (auto& r : out)
      if (r._priv) eng.release(&r);
    if (held > 0) unreserve(held);
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (uint64_t id : ids) jobs_[id].group = 1;
    }
    for (uint64_t id : ids) run_one(slot, id);
    return;
  }
  int64_t keep = 0;
  if (st == APPT_OK)
    for (int k = 0; k < K; k++) keep += cb[k];
  if (st != APPT_OK)
    for (auto& r : out)
      if (r._priv) eng.release(&r);
  if (held > keep) unreserve(held - keep);
  std::lock_guard<std::mutex> lk(mu_);
  for (int k = 0; k < K; k++) {
    Job& j = jobs_[ids[k]];
    if (st == APPT_OK) {
      j.C = out[k];
      j.reserved = cb[k];
      j.products = prods[k];
    }
    finish(ids[k], st);
  }
}

void Server::fill(const Job& j, appt_job_info* out) const {
  std::memset(out, 0, sizeof(*out));
  out->state = j.state;
  out->status = j.status;
  out->group = j.group;
  const bool started = j.state != APPT_JOB_QUEUED;
  const bool done = j.state == APPT_JOB_DONE || j.state == APPT_JOB_RELEASED;
  out->queue_ms = started ? ms_between(j.t_submit, j.t_start) : ms_between(j.t_submit, Clock::now());
  out->run_ms = done ? ms_between(j.t_start, j.t_done) : 0.0;
  out->latency_ms = done ? ms_between(j.t_submit, j.t_done) : 0.0;
  out->products = j.products;
  out->reserved_bytes = j.reserved;
  out->C = j.C;
}

void Server::poll(uint64_t id, appt_job_info* out) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw Error(APPT_ERR_INVALID, "unknown job");
  fill(it->second, out);
}

void Server::wait(uint64_t id, appt_job_info* out) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw Error(APPT_ERR_INVALID, "unknown job");
  cv_done_.wait(lk, [&] { return it->second.state == APPT_JOB_DONE || it->second.state == APPT_JOB_RELEASED; });
  fill(it->second, out);
}

void Server::release(uint64_t id) {
  appt_csr_result C{};
  int slot = -1;
  int64_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) throw Error(APPT_ERR_INVALID, "unknown job");
    Job& j = it->second;
    if (j.state != APPT_JOB_DONE) throw Error(APPT_ERR_INVALID, "job has not finished");
    C = j.C;
    slot = j.slot;
    bytes = j.reserved;
    std::memset(&j.C, 0, sizeof(j.C));
    j.reserved = 0;
    j.state = APPT_JOB_RELEASED;
  }
  if (C._priv && slot >= 0) {
    Engine& eng = *slots_[slot].eng;
    (void)hipSetDevice(device_);
    eng.release(&C);
    // large results go back to HIP so another slot can allocate them
    if (bytes > ((int64_t)1 << 30)) eng.pool().trim();
  }
  unreserve(bytes);
}

void Server::stats(appt_server_stats* out) {
  std::lock_guard<std::mutex> lk(mu_);
  std::memset(out, 0, sizeof(*out));
  out->submitted = submitted_;
  out->completed = completed_;
  out->refused = refused_;
  out->failed = failed_;
  if (!latencies_.empty()) {
    std::vector<double> v = latencies_;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const double pos = q * (double)(v.size() - 1);
      const size_t lo = (size_t)pos;
      const size_t hi = std::min(lo + 1, v.size() - 1);
      return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    double s = 0;
    for (double x : v) s += x;
    out->mean_ms = s / (double)v.size();
    const double span = ms_between(first_submit_, last_done_) * 1e-3;
    if (span > 0) {
      out->gflops = flops_done_ / span / 1e9;
      out->jobs_per_s = (double)completed_ / span;
    }
  }
  out->bytes_reserved = reserved_;
  out->peak_bytes_reserved = peak_reserved_;
  out->groups_run = groups_;
}

}  // namespace appt

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct appt_server {
  std::unique_ptr<appt::Server> impl;
};

namespace appt {
appt_status guard_call(const std::function<void()>& f);
}

extern "C" {

appt_status appt_server_create(const appt_server_options* opt, appt_server** out) {
  if (!out) return APPT_ERR_INVALID;
  *out = nullptr;
  return appt::guard_call([&] {
    appt_server_options o{};
    if (opt) o = *opt;
    auto* s = new appt_server;
    try {
      s->impl.reset(new appt::Server(o));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

appt_status appt_server_destroy(appt_server* s) {
  if (!s) return APPT_OK;
  return appt::guard_call([&] { delete s; });
}

appt_status appt_server_submit(appt_server* s, const appt_csr_view* A, appt_op op, uint64_t* job) {
  if (!s || !A || !job) return APPT_ERR_INVALID;
  return appt::guard_call([&] { *job = s->impl->submit(*A, op); });
}

appt_status appt_server_poll(appt_server* s, uint64_t job, appt_job_info* out) {
  if (!s || !out) return APPT_ERR_INVALID;
  return appt::guard_call([&] { s->impl->poll(job, out); });
}

appt_status appt_server_wait(appt_server* s, uint64_t job, appt_job_info* out) {
  if (!s || !out) return APPT_ERR_INVALID;
  return appt::guard_call([&] { s->impl->wait(job, out); });
}

appt_status appt_server_release(appt_server* s, uint64_t job) {
  if (!s) return APPT_ERR_INVALID;
  return appt::guard_call([&] { s->impl->release(job); });
}

appt_status appt_server_stats_get(appt_server* s, appt_server_stats* out) {
  if (!s || !out) return APPT_ERR_INVALID;
  return appt::guard_call([&] { s->impl->stats(out); });
}

}  // extern "C"
