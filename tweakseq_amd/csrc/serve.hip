// Serving loop: continuous batching of alignment requests.
//
// Replaces the one-job-at-a-time QProcess lifecycle of the reference
// (tweakseq/UI/SeqEditMainWin.cpp:1616-1663 start, :836-861 finish, :803-812
// cancel) with an in-process queue.  ta_submit copies a request and returns a
// ticket at once.  A scheduler thread per context admits pending requests in
// arrival order while the traceback memory of everything in flight stays
// within the budget and fewer than `max_inflight` batches run, and groups
// admitted requests with equal parameters into a batch.  Each batch goes to
// one of `max_inflight` worker threads, which builds it, runs it on the
// worker's own stream, waits for it and splits the results back per request:
// the host work of one batch (encoding, planning, result formatting)
// overlaps the GPU work and the host work of the others, and requests that
// arrive while others run are picked up as soon as memory and a slot free
// up.  Results are per pair and do not depend on what else shares the batch
// (pairs never interact).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime_internal.h"

namespace {

enum class RState { Pending, InFlight, Done, Cancelled, Failed };

struct Request {
  uint64_t ticket = 0;
  ta_params p{};
  std::vector<std::string> a, b;
  int64_t tb_bytes = 0;
  RState state = RState::Pending;
  int err = 0;
  bool walk_failed = false;    // a pair's walk gave up: delivered with TA_E_WALK
  std::vector<ta_result> res;  // cigar_off relative to `cigar`
  std::string cigar;
  int dq = -1;                 // runs in the device queue: -1 not known yet, 0 no, 1 yes
  DevqJob *job = nullptr;      // its job there while in flight
  bool resident = false;       // pairs of resident sequences (ta_submit_resident), resolved in rb
  ResidentBatch rb;
  size_t npairs() const { return resident ? rb.la.size() : a.size(); }
};

struct Group {
  std::vector<Request *> reqs;
  int64_t tb = 0;
};

bool same_params(const ta_params &x, const ta_params &y) { return memcmp(&x, &y, sizeof(ta_params)) == 0; }

}  // namespace

struct Serve {
  ta_ctx *ctx = nullptr;
  std::mutex mu;
  std::condition_variable cv;   // scheduler: requests, slots, memory
  std::condition_variable wcv;  // workers: admitted groups
  std::deque<Request *> pending;
  std::unordered_map<uint64_t, Request *> all;
  std::deque<Group> groups;     // admitted, waiting for a worker
  int64_t budget = 64ll << 30, used = 0;
  int max_inflight = 4, inflight = 0;
  uint64_t next_ticket = 1;
  uint64_t changes = 0;         // bumped whenever admission may change its mind
  bool stop = false, stop_workers = false;
  std::thread th;
  std::vector<std::thread> workers;
  // device queue (devq.hip): requests admitted into the running persistent kernel
  DevQ *dq = nullptr;
  bool dq_tried = false, dq_stop = false;
  std::vector<Request *> dq_jobs;  // in flight there
  std::thread dq_th;
  std::condition_variable dq_cv;

  void loop();
  void worker();
  void run_group(Group &g, hipStream_t st);
  void dq_loop();
  void admit_dq();
};

// Requests the device queue can take, in arrival order, while it has room
// (lock held).  The queue is created on first use; without it (no room on the
// device, or TWEAKALIGN_DEVICE_QUEUE=0) every request goes to a batch.
void Serve::admit_dq() {
  if (!dq_tried) {
    dq_tried = true;
    dq = devq_create(ctx, budget);
    if (dq) dq_th = std::thread([this] { dq_loop(); });
  }
  if (!dq) return;
  bool any = false;
  for (auto it = pending.begin(); it != pending.end();) {
    Request *r = *it;
    if (r->state == RState::Cancelled) { it = pending.erase(it); continue; }
    if (r->dq < 0) r->dq = devq_eligible(dq, &r->p, r->a, r->b) ? 1 : 0;
    if (!r->dq) { ++it; continue; }
    if (r->tb_bytes > budget) {  // the budget was lowered below it: it can never run
      r->state = RState::Failed;
      r->err = TA_E_BUDGET;
      it = pending.erase(it);
      continue;
    }
    if (used + r->tb_bytes > budget) break;
    DevqJob *job = nullptr;
    const int rc = devq_admit(dq, &r->p, r->a, r->b, &job);
    if (rc == TA_E_PENDING) break;  // device memory or slots: wait for completions
    it = pending.erase(it);
    if (rc != TA_OK) {
      r->state = RState::Failed;
      r->err = rc;
      continue;
    }
    r->job = job;
    r->state = RState::InFlight;
    used += r->tb_bytes;
    dq_jobs.push_back(r);
    any = true;
  }
  if (any) {
    devq_tick(dq, true);
    dq_cv.notify_all();
  }
}

// Completions of device-queue jobs: results formatted from host memory, the
// job's memory freed, waiting requests admitted; the kernel relaunched while
// jobs wait and stopped after a quiet spell.
void Serve::dq_loop() {
  std::unique_lock<std::mutex> lk(mu);
  while (!dq_stop) {
    std::vector<Request *> fin;
    for (auto it = dq_jobs.begin(); it != dq_jobs.end();) {
      if (devq_done(dq, (*it)->job)) {
        fin.push_back(*it);
        it = dq_jobs.erase(it);
      } else {
        ++it;
      }
    }
    const bool have = !dq_jobs.empty();
    std::vector<std::vector<ta_result>> res(fin.size());
    std::vector<std::string> text(fin.size());
    std::vector<int> rcs(fin.size(), TA_OK);
    std::vector<bool> cancelled(fin.size());
    for (size_t k = 0; k < fin.size(); ++k) cancelled[k] = fin[k]->state == RState::Cancelled;
    lk.unlock();
    devq_tick(dq, have);
    for (size_t k = 0; k < fin.size(); ++k) {
      if (cancelled[k]) devq_drop(dq, fin[k]->job);
      else rcs[k] = devq_finish(dq, fin[k]->job, res[k], text[k]);
      fin[k]->job = nullptr;
    }
    lk.lock();
    for (size_t k = 0; k < fin.size(); ++k) {
      Request *r = fin[k];
      used -= r->tb_bytes;
      if (r->state == RState::Cancelled) {
        all.erase(r->ticket);
        delete r;
        continue;
      }
      r->res.swap(res[k]);
      r->cigar.swap(text[k]);
      r->walk_failed = rcs[k] == TA_E_WALK;
      r->state = RState::Done;
      r->a.clear(); r->a.shrink_to_fit();
      r->b.clear(); r->b.shrink_to_fit();
    }
    if (!fin.empty()) {
      ++changes;
      cv.notify_all();
    }
    if (dq_jobs.empty()) {
      dq_cv.wait_for(lk, std::chrono::milliseconds(10), [&] { return dq_stop || !dq_jobs.empty(); });
    } else {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    }
  }
}

// One batch from build to per-request results (a worker, without the lock).
// A group is all letters or all resident sequences of one arena.
void Serve::run_group(Group &g, hipStream_t st) {
  std::vector<ta_pair> pairs;
  ResidentBatch rb;
  const bool resident = g.reqs[0]->resident;
  if (resident) {
    rb.arena = g.reqs[0]->rb.arena;
    for (Request *r : g.reqs) {
      const ResidentBatch &x = r->rb;
      rb.a_off.insert(rb.a_off.end(), x.a_off.begin(), x.a_off.end());
      rb.b_off.insert(rb.b_off.end(), x.b_off.begin(), x.b_off.end());
      rb.la.insert(rb.la.end(), x.la.begin(), x.la.end());
      rb.lb.insert(rb.lb.end(), x.lb.begin(), x.lb.end());
      rb.a_bound.insert(rb.a_bound.end(), x.a_bound.begin(), x.a_bound.end());
      rb.b_bound.insert(rb.b_bound.end(), x.b_bound.begin(), x.b_bound.end());
    }
  } else {
    for (Request *r : g.reqs)
      for (size_t k = 0; k < r->a.size(); ++k)
        pairs.push_back({r->a[k].data(), (int32_t)r->a[k].size(), r->b[k].data(), (int32_t)r->b[k].size()});
  }
  const int32_t np = resident ? (int32_t)rb.la.size() : (int32_t)pairs.size();
  ta_batch *B = nullptr;
  std::vector<ta_result> res;
  std::string text;
  int rc = resident ? batch_create_on(ctx, &g.reqs[0]->p, nullptr, np, st, &B, nullptr, &rb)
                    : batch_create_on(ctx, &g.reqs[0]->p, pairs.data(), np, st, &B);
  if (!rc) rc = ta_batch_run(ctx, B);
  if (!rc) rc = ta_batch_wait(ctx, B);
  if (!rc) rc = batch_results_text(ctx, B, res, text);
  if (B) ta_batch_destroy(B);
  std::lock_guard<std::mutex> lk(mu);
  int32_t k = 0;
  for (Request *r : g.reqs) {
    const int32_t n = (int32_t)r->npairs();
    r->rb = ResidentBatch();  // the arena is no longer read by this request
    if (r->state == RState::Cancelled) {  // cancelled while in flight: retire now
      k += n;
      all.erase(r->ticket);
      delete r;
      continue;
    }
    if (rc != TA_OK) {
      r->state = RState::Failed;
      r->err = rc;
    } else {
      r->res.assign(res.begin() + k, res.begin() + k + n);
      const int64_t base = n ? r->res[0].cigar_off : 0;
      int64_t len = 0;
      for (auto &x : r->res) len += x.cigar_len;
      r->cigar.assign(text.data() + base, (size_t)len);
      for (auto &x : r->res) {
        x.cigar_off -= base;
        if (x.flags & TA_F_WALK_FAILED) r->walk_failed = true;
      }
      r->state = RState::Done;
      r->a.clear(); r->a.shrink_to_fit();
      r->b.clear(); r->b.shrink_to_fit();
    }
    k += n;
  }
}

void Serve::worker() {
  (void)hipSetDevice(ctx->device);
  hipStream_t st = nullptr;
  const bool have_stream = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    wcv.wait(lk, [&] { return stop_workers || !groups.empty(); });
    if (groups.empty()) break;  // stopping, nothing left
    Group g = std::move(groups.front());
    groups.pop_front();
    lk.unlock();
    if (have_stream) {
      run_group(g, st);
    } else {
      std::lock_guard<std::mutex> g2(mu);
      for (Request *r : g.reqs) {
        if (r->state == RState::InFlight) {
          r->state = RState::Failed;
          r->err = TA_E_DEVICE;
        } else if (r->state == RState::Cancelled) {  // cancelled while in flight: retire now
          all.erase(r->ticket);
          delete r;
        }
      }
    }
    lk.lock();
    --inflight;
    used -= g.tb;
    ++changes;
    cv.notify_all();
  }
  lk.unlock();
  if (have_stream) (void)hipStreamDestroy(st);
}

void Serve::loop() {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    if (stop) break;
    const uint64_t seen = changes;
    if (!pending.empty()) admit_dq();
    // Admission in arrival order while memory and in-flight slots allow;
    // requests with the parameters of the first admitted one form the batch.
    if (!pending.empty() && inflight < max_inflight) {
      Group g;
      for (auto it = pending.begin(); it != pending.end();) {
        Request *r = *it;
        if (r->state == RState::Cancelled) { it = pending.erase(it); continue; }
        if (r->dq == 1) { ++it; continue; }  // waits for room in the device queue
        if (r->tb_bytes > budget) {  // the budget was lowered below it: it can never run
          r->state = RState::Failed;
          r->err = TA_E_BUDGET;
          it = pending.erase(it);
          continue;
        }
        if (!g.reqs.empty() && (!same_params(r->p, g.reqs[0]->p) || r->resident != g.reqs[0]->resident ||
                                (r->resident && r->rb.arena != g.reqs[0]->rb.arena))) {
          ++it;
          continue;
        }
        if (used + g.tb + r->tb_bytes > budget) break;
        g.tb += r->tb_bytes;
        r->state = RState::InFlight;
        g.reqs.push_back(r);
        it = pending.erase(it);
      }
      if (!g.reqs.empty()) {
        used += g.tb;
        ++inflight;
        while ((int)workers.size() < max_inflight) workers.emplace_back([this] { worker(); });
        groups.push_back(std::move(g));
        wcv.notify_one();
        continue;
      }
    }
    cv.wait(lk, [&] { return stop || changes != seen; });
  }
}

void serve_shutdown(ta_ctx *ctx) {
  if (!ctx || !ctx->serve) return;
  Serve *s = ctx->serve;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    for (Request *r : s->pending) r->state = RState::Cancelled;
    s->pending.clear();
  }
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  {  // the scheduler admits nothing more; workers finish what they hold
    std::lock_guard<std::mutex> g(s->mu);
    s->stop_workers = true;
  }
  s->wcv.notify_all();
  for (auto &w : s->workers) w.join();
  if (s->dq) {  // device-queue jobs in flight finish (the kernel's waits are bounded)
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait_for(lk, std::chrono::seconds(60), [&] { return s->dq_jobs.empty(); });
    s->dq_stop = true;
    lk.unlock();
    s->dq_cv.notify_all();
    s->dq_th.join();
    for (Request *r : s->dq_jobs)  // only when the wait above ran out
      if (r->job) devq_drop(s->dq, r->job);
    devq_destroy(s->dq);
  }
  for (auto &kv : s->all) delete kv.second;
  delete s;
  ctx->serve = nullptr;
}

namespace {
Serve *serve_of(ta_ctx *ctx) {
  if (!ctx->serve) {
    Serve *s = new Serve();
    s->ctx = ctx;
    ctx->serve = s;
    s->th = std::thread([s] { s->loop(); });
  }
  return ctx->serve;
}
}  // namespace

int ta_serve_config(ta_ctx *ctx, int64_t tb_budget_bytes, int32_t max_inflight) {
  if (!ctx || tb_budget_bytes <= 0 || max_inflight < 0) return TA_E_INVALID;
  Serve *s = serve_of(ctx);
  std::lock_guard<std::mutex> g(s->mu);
  s->budget = tb_budget_bytes;
  if (max_inflight > 0) s->max_inflight = max_inflight;
  ++s->changes;
  s->cv.notify_all();
  return TA_OK;
}

int ta_submit(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n, uint64_t *ticket) {
  if (!ctx || !p || !ticket || n < 0 || (n > 0 && !pairs)) return TA_E_INVALID;
  for (int32_t k = 0; k < n; ++k)
    if (pairs[k].la < 0 || pairs[k].lb < 0 || (pairs[k].la && !pairs[k].a) || (pairs[k].lb && !pairs[k].b))
      return TA_E_INVALID;
  Serve *s = serve_of(ctx);
  Request *r = new Request();
  r->p = *p;
  r->tb_bytes = batch_tb_estimate(p, pairs, n);
  r->a.reserve(n);
  r->b.reserve(n);
  for (int32_t k = 0; k < n; ++k) {
    r->a.emplace_back(pairs[k].a, (size_t)pairs[k].la);
    r->b.emplace_back(pairs[k].b, (size_t)pairs[k].lb);
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (r->tb_bytes > s->budget) {
    delete r;
    return TA_E_BUDGET;
  }
  r->ticket = s->next_ticket++;
  s->all[r->ticket] = r;
  s->pending.push_back(r);
  *ticket = r->ticket;
  ++s->changes;
  s->cv.notify_all();
  return TA_OK;
}

int ta_submit_resident(ta_ctx *ctx, const ta_params *p, const ta_seq_pair *pairs, int32_t n, uint64_t *ticket) {
  if (!ctx || !p || !ticket || n < 0 || (n > 0 && !pairs)) return TA_E_INVALID;
  Request *r = new Request();
  const int rc = resident_resolve(ctx, p, pairs, n, r->rb);
  if (rc) {
    delete r;
    return rc;
  }
  r->p = *p;
  r->resident = true;
  r->dq = 0;  // batches only: the device queue stages letters from host memory
  std::vector<ta_pair> lens((size_t)n);
  for (int32_t k = 0; k < n; ++k) lens[k] = ta_pair{nullptr, r->rb.la[k], nullptr, r->rb.lb[k]};
  r->tb_bytes = batch_tb_estimate(p, lens.data(), n);
  Serve *s = serve_of(ctx);
  std::lock_guard<std::mutex> g(s->mu);
  if (r->tb_bytes > s->budget) {
    delete r;
    return TA_E_BUDGET;
  }
  r->ticket = s->next_ticket++;
  s->all[r->ticket] = r;
  s->pending.push_back(r);
  *ticket = r->ticket;
  ++s->changes;
  s->cv.notify_all();
  return TA_OK;
}

int ta_poll(ta_ctx *ctx, uint64_t ticket, ta_result *out, char *cigar_buf, int64_t cap, int64_t *used) {
  if (!ctx) return TA_E_INVALID;
  if (!ctx->serve) return TA_E_TICKET;
  Serve *s = ctx->serve;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return TA_E_TICKET;
  Request *r = it->second;
  switch (r->state) {
    case RState::Pending:
    case RState::InFlight:
      return TA_E_PENDING;
    case RState::Cancelled:
      return TA_E_CANCELLED;
    case RState::Failed:
      return r->err ? r->err : TA_E_DEVICE;
    case RState::Done:
      break;
  }
  if (used) *used = (int64_t)r->cigar.size();
  if (!out) return TA_E_INVALID;
  memcpy(out, r->res.data(), r->res.size() * sizeof(ta_result));
  if ((int64_t)r->cigar.size() > cap || (!cigar_buf && !r->cigar.empty())) return TA_E_SPACE;
  if (!r->cigar.empty()) memcpy(cigar_buf, r->cigar.data(), r->cigar.size());
  const int rc = r->walk_failed ? TA_E_WALK : TA_OK;
  s->all.erase(it);  // delivered: the ticket is retired
  delete r;
  return rc;
}

int ta_cancel(ta_ctx *ctx, uint64_t ticket) {
  if (!ctx) return TA_E_INVALID;
  if (!ctx->serve) return TA_E_TICKET;
  Serve *s = ctx->serve;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return TA_E_TICKET;
  Request *r = it->second;
  const RState was = r->state;
  if (was == RState::Pending || was == RState::InFlight) r->state = RState::Cancelled;
  if (was != RState::InFlight) {
    // Pending requests leave the queue at the next admission pass; done,
    // failed or cancelled ones are retired now.
    for (auto q = s->pending.begin(); q != s->pending.end(); ++q)
      if (*q == r) { s->pending.erase(q); break; }
    s->all.erase(it);
    delete r;
  }
  return TA_OK;
}
