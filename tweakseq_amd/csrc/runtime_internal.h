// Internal interfaces between the batch runtime and the serving loop.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tweakalign.h"

struct Serve;

// ta_ctx internals shared with the serving loop.
struct ta_ctx {
  int device = 0;
  int cus = 256;
  int lds_max = 65536;               // LDS bytes one workgroup may hold
  hipStream_t stream = nullptr;
  Serve *serve = nullptr;
  std::mutex tune_mu;                // the serving thread creates batches too
  std::map<int, ta_tuning> tune;     // tune_key -> tuning
  // ta_align_pair's host staging (pinned, mapped: the kernel reads the
  // letters and writes the result there), grown on demand, kept until
  // ta_destroy.
  void *pair_stage = nullptr;
  size_t pair_stage_cap = 0;
  // Resident sequences (ta_seq_add): residue codes in one device arena, each
  // sequence 16-byte aligned and padded with code 0, 32 bytes of slack after
  // the last (the walkers' windows); per sequence its offset, length,
  // alphabet and the count of each residue code (the int16 saturation bound
  // of a batch depends on its scores).  Batches hold the arena they read
  // (shared), so growing or resetting the store never frees an arena a
  // served request still reads.  Added codes wait in host memory
  // (seq_pending: arena bytes [seq_flushed, seq_used + 32)) and go up in one
  // copy when a call resolves handles, or before an add once 64 MB are
  // pending.
  // seq_mu guards the store: serving threads read it while the caller adds.
  struct SeqRec {
    uint64_t off;
    int32_t len, alphabet;
  };
  std::mutex seq_mu;
  std::shared_ptr<uint8_t> seq_arena;
  size_t seq_cap = 0, seq_used = 0, seq_flushed = 0;
  std::vector<uint8_t> seq_pending;
  std::vector<SeqRec> seqs;
  std::vector<int32_t> seq_counts;  // 32 per sequence
};

// Pairs of resident sequences resolved when a call takes them (the store may
// grow or be reset afterwards): the arena the codes are in and, per pair,
// code offsets, lengths and the int16 saturation bound of each side under the
// call's scores.
struct ResidentBatch {
  std::shared_ptr<uint8_t> arena;
  std::vector<uint64_t> a_off, b_off;
  std::vector<int32_t> la, lb;
  std::vector<int64_t> a_bound, b_bound;
};
// TA_E_INVALID for an unknown handle or one of another alphabet than p's.
int resident_resolve(ta_ctx *ctx, const ta_params *p, const ta_seq_pair *pairs, int32_t n, ResidentBatch &out);

// Finds the device address of ta_align_pair_static's scratch array on
// `device` (current), once per device; later calls return at once.
void pair_scratch_setup(int device);

// A global pair the packed int16 lanes hold without saturating (the batch
// runtime's rule, shared with the device-queue planner).
bool global_int16_safe(const ta_params *p, int maxS, int la, int lb);

// Device-queue serving (devq.hip): requests planned straight into the
// running persistent kernel.  devq_create returns nullptr when the device
// cannot host it (serving then builds a batch per request group).
struct DevQ;
struct DevqJob;
DevQ *devq_create(ta_ctx *ctx, int64_t tb_budget);
void devq_destroy(DevQ *q);  // waits for the kernel to leave
// Whether this request can run in the queue (parameters, sizes).
bool devq_eligible(DevQ *q, const ta_params *p, const std::vector<std::string> &a, const std::vector<std::string> &b);
// Admit: TA_OK with *job, TA_E_PENDING when memory or slots are short now,
// another error when the request cannot run (a letter outside the alphabet).
int devq_admit(DevQ *q, const ta_params *p, const std::vector<std::string> &a, const std::vector<std::string> &b,
               DevqJob **job);
bool devq_done(DevQ *q, DevqJob *job);
// Results of a finished job (cigar_off into text), then its memory is freed.
int devq_finish(DevQ *q, DevqJob *job, std::vector<ta_result> &res, std::string &text);
// Frees a job's memory once it has finished (cancelled requests).
void devq_drop(DevQ *q, DevqJob *job);
// Relaunches the kernel when jobs are waiting and the last launch has ended.
void devq_tick(DevQ *q, bool have_jobs);

// TA_OK, or TA_E_INVALID for parameters outside what the kernels take.
int params_check(const ta_params *p);
// CIGAR text of binary runs ((len << 2) | op, op 0 M / 1 D / 2 I); with out
// == nullptr only the length.  Returns the text length.
int64_t cigar_text(const uint32_t *runs, int n, char *out);

// Key of a kind of batch in ta_ctx::tune.
inline int tune_key(const ta_params *p, int bucket) {
  return ((p->mode * 2 + (p->traceback ? 1 : 0)) * 2 + p->alphabet) * 8 + bucket;
}

// ta_batch_create on a given stream (the serving loop gives each in-flight
// batch its own).
// `force`: use this tuning instead of the context's table (the autotuner).
// `res`: the pairs are resident sequences (resident_resolve) and `pairs` is
// unused.
int batch_create_on(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n, hipStream_t stream,
                    ta_batch **out, const ta_tuning *force = nullptr, const ResidentBatch *res = nullptr);
// Direction-plane bytes a batch of these pairs would allocate (admission).
int64_t batch_tb_estimate(const ta_params *p, const ta_pair *pairs, int32_t n);
// Results of a finished batch with the CIGAR text in one string
// (ta_result::cigar_off indexes `text`): ta_batch_results in one pass.
int batch_results_text(ta_ctx *ctx, ta_batch *B, std::vector<ta_result> &out, std::string &text);
// Device memory from the library's cache (blocks are reused, never handed to
// hipFree while work runs; device_free returns them to the cache).
void *device_alloc(size_t bytes);
void device_free(void *p);
// Host worker threads kept for the process's life (up to 15, plus the
// caller): the parallel sections of batch creation and results run on them
// instead of spawning threads per section.  One job at a time; a caller
// that finds the pool busy gets `false` and runs its section on threads of
// its own.
class HostPool {
 public:
  static HostPool &get() {
    static HostPool *pool = new HostPool();  // never destroyed: its workers outlive static destruction
    return *pool;
  }
  int workers() const { return nw_; }
  // f(t) for t in [0, T), T <= workers() + 1; f(0) on the caller.
  bool run(int T, const std::function<void(int)> &f) {
    std::unique_lock<std::mutex> own(job_mu_, std::try_to_lock);
    if (!own.owns_lock() || T > nw_ + 1) return false;
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = &f;
      tasks_ = T;
      pending_ = T - 1;
      ++gen_;
    }
    cv_.notify_all();
    f(0);
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [&] { return pending_ == 0; });
    job_ = nullptr;
    return true;
  }

 private:
  HostPool() {
    nw_ = (int)std::min(15u, std::max(1u, std::thread::hardware_concurrency()) - 1);
    for (int w = 1; w <= nw_; ++w) std::thread([this, w] { loop(w); }).detach();
  }
  void loop(int w) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> g(mu_);
    for (;;) {
      cv_.wait(g, [&] { return gen_ != seen; });
      seen = gen_;
      if (w >= tasks_) continue;
      const std::function<void(int)> *f = job_;
      g.unlock();
      (*f)(w);
      g.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }
  int nw_ = 0;
  std::mutex job_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *job_ = nullptr;
  int tasks_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
};

// Runs f(t) for t in [0, T) on T host threads (f(0) inline when T is 1).
template <class F>
void parallel_tasks(int T, F f) {
  if (T <= 1) {
    f(0);
    return;
  }
  if (HostPool::get().run(T, std::function<void(int)>([&f](int t) { f(t); }))) return;
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 0; t < T; ++t) th.emplace_back([&f, t] { f(t); });
  for (auto &x : th) x.join();
}

// Runs f(i) for i in [0, n) on up to 16 host threads in contiguous chunks
// (f must not depend on the order).
template <class F>
void parallel_for(size_t n, F f) {
  const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t nt = std::min<size_t>(hw, (n + 255) / 256);
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  parallel_tasks((int)nt, [&](int t) {
    for (size_t i = n * (size_t)t / nt; i < n * ((size_t)t + 1) / nt; ++i) f(i);
  });
}

// Host-side phase timings on stderr when TWEAKALIGN_TRACE=1: each mark()
// prints the time since the previous one under the trace's name.
struct PhaseTrace {
  const char *name;
  bool on;
  std::chrono::steady_clock::time_point t;
  explicit PhaseTrace(const char *n) : name(n) {
    const char *e = getenv("TWEAKALIGN_TRACE");
    on = e && e[0] == '1';
    if (on) t = std::chrono::steady_clock::now();
  }
  void mark(const char *what) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    fprintf(stderr, "[tweakalign] %s %-14s %8.3f ms\n", name, what,
            std::chrono::duration<double>(now - t).count() * 1e3);
    t = now;
  }
};

// Stops the serving thread (called by ta_destroy).
void serve_shutdown(ta_ctx *ctx);
