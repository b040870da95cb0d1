// server.cpp -- request loop with memory admission and continuous batching.
//
// Worker threads, one per slot, each own an Engine (own stream, own pool and
// scratch).  A worker takes the oldest queued job; when it is small, it also
// takes the other small jobs waiting at that moment and runs them together
// as one block-diagonal batch.
//
// Admission counts every byte the slots' pools hold against the budget: each
// slot is charged what its pool has allocated (cached working set, free
// blocks, live results).  A run asks for more before it allocates: a bound of
// its working set before the operands are staged (Engine::work_bound), the
// workgroup bitmaps before the symbolic pass keeps them, then the exact bytes
// of C once the symbolic phase has fixed nnz(C).  A request that does not fit
// first returns idle slots' caches (and its own free blocks) to HIP.  Only a
// run's first request, made while it holds nothing, then waits for running
// jobs to end or results to be released (refused when it alone exceeds the
// budget, or after admission_wait_ms); a later one that still does not fit
// makes the run give back what it holds and start again, admitted up front
// for everything it asked for -- no run waits while holding memory, so runs
// cannot wait on one another.  When a run ends, its slot's charge is set to
// what the pool really holds, and a slot whose cache outgrows its share of
// the budget returns it.
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

// Working set of a group of K jobs with N rows and nnz entries in all, C
// excluded: the batch's packed operands (A and the transposes: 24 B an
// entry, row-to-block maps and local row pointers: 64 B a row) and the
// pipeline's own working set over the block-diagonal matrix.
// Thrown by a mid-run request that does not fit: the run starts again.
struct Retry {};

static int64_t group_work(int64_t N, int64_t nnz, int K) {
  return Engine::work_bound((int32_t)N, (int32_t)N, nnz) + 64 * (N + K + 1) + 32 * nnz + 256 * (int64_t)K;
}

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
  // C of a job group: one allocation shared by its jobs' results, returned to
  // the slot's pool when the last of them is released
  struct SharedC {
    void* rp = nullptr;
    void* col = nullptr;
    void* val = nullptr;
    int64_t bytes = 0;
    int remaining = 0;
  };
  struct Job {
    appt_csr_view A{};
    appt_op op = APPT_OP_AA;
    appt_job_state state = APPT_JOB_QUEUED;
    appt_status status = APPT_OK;
    int group = 1;
    Clock::time_point t_submit, t_start, t_done;
    int64_t products = 0;
    int64_t reserved = 0;  // bytes of its result
    appt_csr_result C{};   // _priv set when the job owns C alone
    std::shared_ptr<SharedC> shared;
    int slot = -1;
  };
  struct Slot {
    std::unique_ptr<Engine> eng;
    std::thread th;
    std::mutex run;      // held by the worker while it runs, and by whoever trims an idle slot
    int64_t charge = 0;  // budget bytes charged to this slot (under mu_)
    int64_t live = 0;    // bytes of its finished, unreleased results (under mu_)
  };

  bool small(const Job& j) const { return j.A.n <= row_limit_ && max_group_ > 1; }
  void worker(int slot);
  void run_one(int slot, uint64_t id);
  void run_group(int slot, const std::vector<uint64_t>& ids);
  // charge `need` more bytes to `slot`, returning caches as needed.  `idle`
  // (a run's first request): the slot's own workspace is not in use yet and
  // may be returned too, and the request waits for memory, throwing
  // APPT_ERR_ADMISSION when refused; otherwise a request that does not fit
  // throws Retry.
  void reserve(int slot, int64_t need, bool idle);
  // Engine admission callback of one run: the first request reserves
  // `prepay` more bytes, from which later requests are served first; `asked`
  // sums the later requests (the next attempt's prepay)
  Engine::Admission admission(int slot, int64_t prepay, int64_t* asked);
  // return idle slots' workspaces and free blocks (and this slot's free
  // blocks, or its workspace too when `self_idle`) to HIP
  void reclaim(int self, bool self_idle);
  // `freed` bytes of `slot`'s pool went back to HIP
  void credit(int slot, int64_t freed);
  // end of a run: charge the slot what its pool holds, trimming a large cache
  void settle(int slot);
  void finish(uint64_t id, appt_status st);
  void fill(const Job& j, appt_job_info* out) const;

  int device_;
  int64_t budget_, cache_cap_;
  int max_group_;
  int32_t row_limit_;
  int wait_ms_;
  PoolTally tally_;  // declared before slots_: outlives the pools counted in it
  std::vector<std::unique_ptr<Slot>> slots_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_, cv_mem_;
  std::deque<uint64_t> queue_;
  std::unordered_map<uint64_t, Job> jobs_;
  uint64_t next_id_ = 1;
  bool stop_ = false;
  int running_ = 0, waiters_ = 0;
  int64_t reserved_ = 0, peak_reserved_ = 0, overruns_ = 0;
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
  cache_cap_ = budget_ / (2 * (int64_t)nslots);
  for (int k = 0; k < nslots; k++) {
    appt_options eo{};
    eo.device = device_;
    eo.tuning = o.tuning;
    std::unique_ptr<Slot> S(new Slot);
    S->eng.reset(new Engine(&eo));
    S->eng->pool().set_tally(&tally_);
    S->charge = S->eng->pool().allocated();
    reserved_ += S->charge;
    slots_.push_back(std::move(S));
  }
  peak_reserved_ = reserved_;
  for (int k = 0; k < nslots; k++) slots_[k]->th = std::thread([this, k] { worker(k); });
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
    if (s->th.joinable()) s->th.join();
  (void)hipSetDevice(device_);
  for (auto& kv : jobs_) {
    Job& j = kv.second;
    if (j.slot < 0) continue;
    Engine& e = *slots_[j.slot]->eng;
    if (j.C._priv) e.release(&j.C);
    if (j.shared && --j.shared->remaining == 0) {
      e.pool().put(j.shared->rp);
      e.pool().put(j.shared->col);
      e.pool().put(j.shared->val);
    }
  }
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

void Server::credit(int slot, int64_t freed) {
  if (freed <= 0) return;
  {
    std::lock_guard<std::mutex> lk(mu_);
    slots_[slot]->charge -= freed;
    reserved_ -= freed;
  }
  cv_mem_.notify_all();
}

void Server::reclaim(int self, bool self_idle) {
  for (int k = 0; k < (int)slots_.size(); k++) {
    Slot& S = *slots_[k];
    std::unique_lock<std::mutex> rl(S.run, std::defer_lock);
    if (k != self && !rl.try_lock()) continue;  // running: its memory comes back when it ends
    Engine& e = *S.eng;
    const int64_t before = e.pool().allocated();
    if (k != self || self_idle) e.trim_all();
    else e.pool().trim();
    credit(k, before - e.pool().allocated());
  }
}

void Server::reserve(int slot, int64_t need, bool idle) {
  if (need > budget_)
    throw Error(APPT_ERR_ADMISSION, "job needs " + std::to_string(need) + " B, budget " + std::to_string(budget_) + " B");
  const auto deadline = Clock::now() + std::chrono::milliseconds(wait_ms_);
  std::unique_lock<std::mutex> lk(mu_);
  bool reclaimed = false;
  while (reserved_ + need > budget_) {
    if (stop_) throw Error(APPT_ERR_ADMISSION, "server is stopping");
    if (!reclaimed) {
      reclaimed = true;
      lk.unlock();
      reclaim(slot, idle);
      lk.lock();
      continue;
    }
    if (!idle) throw Retry{};
    // memory returns when running jobs end or results are released
    waiters_++;
    const bool timed_out = cv_mem_.wait_until(lk, deadline) == std::cv_status::timeout;
    waiters_--;
    if (timed_out && reserved_ + need > budget_)
      throw Error(APPT_ERR_ADMISSION, "no room within the HBM budget after waiting " + std::to_string(wait_ms_) +
                                          " ms (" + std::to_string(reserved_) + " B held)");
    reclaimed = false;  // slots may have gone idle meanwhile
  }
  slots_[slot]->charge += need;
  reserved_ += need;
  peak_reserved_ = std::max(peak_reserved_, reserved_);
}

Engine::Admission Server::admission(int slot, int64_t prepay, int64_t* asked) {
  auto left = std::make_shared<int64_t>(-1);  // prepaid bytes not used yet; -1 before the first request
  return [this, slot, prepay, asked, left](int64_t bytes, bool) {
    if (*left < 0) {
      reserve(slot, bytes + prepay, true);
      *left = prepay;
      return;
    }
    *asked += bytes;
    const int64_t use = std::min(*left, bytes);
    *left -= use;
    if (bytes > use) reserve(slot, bytes - use, false);
  };
}

void Server::settle(int slot) {
  Slot& S = *slots_[slot];
  Engine& e = *S.eng;
  int64_t live;
  {
    std::lock_guard<std::mutex> lk(mu_);
    live = S.live;
  }
  if (e.pool().allocated() - live > cache_cap_) e.trim_all();
  const int64_t held = e.pool().allocated();
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (held > S.charge) overruns_++;
    reserved_ += held - S.charge;
    S.charge = held;
    peak_reserved_ = std::max(peak_reserved_, reserved_);
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
    {
      std::lock_guard<std::mutex> rl(slots_[slot]->run);
      if (ids.size() == 1) run_one(slot, ids[0]);
      else run_group(slot, ids);
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      running_--;
    }
    cv_done_.notify_all();
    cv_mem_.notify_all();
  }
}

void Server::run_one(int slot, uint64_t id) {
  Engine& eng = *slots_[slot]->eng;
  appt_csr_view A;
  appt_op op;
  {
    std::lock_guard<std::mutex> lk(mu_);
    A = jobs_[id].A;
    op = jobs_[id].op;
  }
  appt_csr_result C{};
  appt_status st = APPT_OK;
  int64_t prepay = 0;
  for (int attempt = 0;; attempt++) {
    int64_t asked = 0;
    bool retry = false;
    st = APPT_OK;
    eng.set_admission(admission(slot, prepay, &asked));
    try {
      eng.compute(A, op, &C);
    } catch (const Retry&) {
      retry = true;
    } catch (const Error& e) {
      st = e.code;
    } catch (...) {
      st = APPT_ERR_HIP;
    }
    eng.set_admission(nullptr);
    if (!retry) break;
    // the second attempt is admitted up front for all the first one asked
    // for; a third means the requests keep growing: refuse
    settle(slot);
    if (attempt >= 2) {
      st = APPT_ERR_ADMISSION;
      break;
    }
    prepay = std::max(prepay, asked);
  }
  const int64_t cbytes = st == APPT_OK ? result_bytes(C.n, C.nnz) : 0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    slots_[slot]->live += cbytes;
  }
  settle(slot);
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
  Engine& eng = *slots_[slot]->eng;
  const int K = (int)ids.size();
  std::vector<appt_csr_view> views((size_t)K);
  std::vector<int32_t> ns((size_t)K);
  std::vector<int64_t> nnzs((size_t)K);
  std::vector<appt_op> ops((size_t)K);
  int64_t N = 0, nnz = 0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (int k = 0; k < K; k++) {
      const Job& j = jobs_[ids[k]];
      views[k] = j.A;
      ns[k] = j.A.n;
      nnzs[k] = j.A.nnz;
      ops[k] = j.op;
      N += j.A.n;
      nnz += j.A.nnz;
    }
  }
  appt_status st = APPT_OK;
  std::vector<appt_csr_result> out((size_t)K);
  std::vector<int64_t> prods((size_t)K, 0);
  std::shared_ptr<SharedC> shared;
  // the batch's own requests (bitmaps, C) come after its working set
  int64_t asked = 0;
  auto admit = admission(slot, 0, &asked);
  eng.set_admission(admit);
  try {
    if (N > INT32_MAX - 2 * (int64_t)K) throw Error(APPT_ERR_ADMISSION, "group too large for one batch");
    admit(group_work(N, nnz, K), false);
    Batch b(eng, K, ns.data(), nnzs.data(), ops.data());
    b.gather(views.data());
    b.plan(false);
    APPT_HIP(hipStreamSynchronize(eng.stream()));
    int64_t cnnz = 0;
    for (int k = 0; k < K; k++) {
      b.result(k, &out[k]);
      prods[k] = b.products(k);
      cnnz += out[k].nnz;
    }
    shared = std::make_shared<SharedC>();
    b.detach_results(&shared->rp, &shared->col, &shared->val);
    shared->bytes = result_bytes((int32_t)std::min<int64_t>(N + K, INT32_MAX), cnnz);
    shared->remaining = K;
  } catch (const Retry&) {
    st = APPT_ERR_ADMISSION;
  } catch (const Error& e) {
    st = e.code;
  } catch (...) {
    st = APPT_ERR_HIP;
  }
  eng.set_admission(nullptr);
  if (st == APPT_ERR_ADMISSION) {
    // the group does not fit as a whole: admit its jobs one by one
    settle(slot);
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (uint64_t id : ids) jobs_[id].group = 1;
    }
    for (uint64_t id : ids) run_one(slot, id);
    return;
  }
  if (st == APPT_OK) {
    std::lock_guard<std::mutex> lk(mu_);
    slots_[slot]->live += shared->bytes;
  }
  settle(slot);
  std::lock_guard<std::mutex> lk(mu_);
  for (int k = 0; k < K; k++) {
    Job& j = jobs_[ids[k]];
    if (st == APPT_OK) {
      j.C = out[k];  // a view into the group's C
      j.shared = shared;
      j.reserved = result_bytes(out[k].n, out[k].nnz);
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
  std::shared_ptr<SharedC> last;  // the group's C, when this is its last job
  int slot = -1;
  bool waiting = false;
  int64_t live = 0;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) throw Error(APPT_ERR_INVALID, "unknown job");
    Job& j = it->second;
    if (j.state != APPT_JOB_DONE) throw Error(APPT_ERR_INVALID, "job has not finished");
    C = j.C;
    slot = j.slot;
    if (j.shared && --j.shared->remaining == 0) last = j.shared;
    j.shared.reset();
    if (slot >= 0) {
      if (C._priv) slots_[slot]->live -= j.reserved;
      if (last) slots_[slot]->live -= last->bytes;
    }
    std::memset(&j.C, 0, sizeof(j.C));
    j.reserved = 0;
    j.state = APPT_JOB_RELEASED;
    waiting = waiters_ > 0;
    if (slot >= 0) live = slots_[slot]->live;
  }
  if (slot >= 0 && (C._priv || last)) {
    Engine& eng = *slots_[slot]->eng;
    (void)hipSetDevice(device_);
    if (C._priv) eng.release(&C);
    if (last) {
      eng.pool().put(last->rp);
      eng.pool().put(last->col);
      eng.pool().put(last->val);
    }
    // the bytes stay in the slot's pool (and charged to it) for reuse, unless
    // a job is waiting for memory or the slot's idle cache outgrows its share
    if (waiting || eng.pool().allocated() - live > cache_cap_) {
      const int64_t before = eng.pool().allocated();
      eng.pool().trim();
      credit(slot, before - eng.pool().allocated());
    }
  }
  cv_mem_.notify_all();
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
  out->device_bytes = tally_.now.load();
  out->peak_device_bytes = tally_.peak.load();
  out->overruns = overruns_;
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
