// Serving loop: continuous batching of alignment requests.
//
// Replaces the one-job-at-a-time QProcess lifecycle of the reference
// (tweakseq/UI/SeqEditMainWin.cpp:1616-1663 start, :836-861 finish, :803-812
// cancel) with an in-process queue.  ta_submit copies a request and returns a
// ticket at once.  A scheduler thread per context admits pending requests in
// arrival order while the traceback memory of everything in flight stays
// within the budget, groups admitted requests with equal parameters into a
// batch, and launches it on its own stream, so requests that arrive while
// others run are picked up as soon as memory and an in-flight slot free up
// and their kernels overlap on the device.  Results are per pair and do not
// depend on what else shares the batch (pairs never interact).
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
  std::vector<ta_result> res;  // cigar_off relative to `cigar`
  std::string cigar;
};

struct Flight {
  ta_batch *B = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  std::vector<Request *> reqs;
  int64_t tb = 0;
  int32_t npairs = 0;
};

bool same_params(const ta_params &x, const ta_params &y) { return memcmp(&x, &y, sizeof(ta_params)) == 0; }

}  // namespace

struct Serve {
  ta_ctx *ctx = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request *> pending;
  std::unordered_map<uint64_t, Request *> all;
  std::vector<Flight> flights;
  std::vector<hipStream_t> free_streams;
  int64_t budget = 64ll << 30, used = 0;
  int max_inflight = 4;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread th;

  void loop();
  void finish(Flight &f);
  bool launch(std::vector<Request *> &group);
};

void Serve::finish(Flight &f) {
  // Results of the whole batch, split back into the requests.
  std::vector<ta_result> res(f.npairs);
  int64_t need = 0;
  int rc = ta_batch_results(ctx, f.B, res.data(), nullptr, 0, &need);
  std::vector<char> text;
  if (rc == TA_OK || rc == TA_E_SPACE) {
    text.resize((size_t)need + 1);
    rc = ta_batch_results(ctx, f.B, res.data(), text.data(), need + 1, &need);
  }
  std::lock_guard<std::mutex> g(mu);
  int32_t k = 0;
  for (Request *r : f.reqs) {
    const int32_t n = (int32_t)r->a.size();
    if (r->state == RState::Cancelled) {  // cancelled while in flight: retire now
      k += n;
      all.erase(r->ticket);
      delete r;
      continue;
    }
    if (r->state == RState::InFlight) {
      if (rc != TA_OK) {
        r->state = RState::Failed;
        r->err = rc;
      } else {
        r->res.assign(res.begin() + k, res.begin() + k + n);
        int64_t base = n ? r->res[0].cigar_off : 0;
        int64_t len = 0;
        for (auto &x : r->res) len += x.cigar_len;
        r->cigar.assign(text.data() + base, (size_t)len);
        for (auto &x : r->res) x.cigar_off -= base;
        r->state = RState::Done;
      }
    }
    k += n;
  }
  used -= f.tb;
  ta_batch_destroy(f.B);
  (void)hipEventDestroy(f.done);
  free_streams.push_back(f.st);
  f.B = nullptr;
}

bool Serve::launch(std::vector<Request *> &group) {
  Flight f;
  std::vector<ta_pair> pairs;
  for (Request *r : group)
    for (size_t k = 0; k < r->a.size(); ++k)
      pairs.push_back({r->a[k].data(), (int32_t)r->a[k].size(), r->b[k].data(), (int32_t)r->b[k].size()});
  {
    std::lock_guard<std::mutex> g(mu);
    if (free_streams.empty()) {
      hipStream_t s;
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return false;
      free_streams.push_back(s);
    }
    f.st = free_streams.back();
    free_streams.pop_back();
  }
  f.reqs = group;
  f.npairs = (int32_t)pairs.size();
  for (Request *r : group) f.tb += r->tb_bytes;
  int rc = batch_create_on(ctx, &group[0]->p, pairs.data(), f.npairs, f.st, &f.B);
  if (!rc) rc = ta_batch_run(ctx, f.B);
  if (!rc && hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess) rc = TA_E_DEVICE;
  if (!rc && hipEventRecord(f.done, f.st) != hipSuccess) rc = TA_E_DEVICE;
  std::lock_guard<std::mutex> g(mu);
  if (rc) {
    for (Request *r : group)
      if (r->state == RState::InFlight) { r->state = RState::Failed; r->err = rc; }
    used -= f.tb;
    if (f.B) ta_batch_destroy(f.B);
    free_streams.push_back(f.st);
    return false;
  }
  flights.push_back(f);
  return true;
}

void Serve::loop() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    if (stop && pending.empty() && flights.empty()) break;
    // Completed batches first: they return memory to the budget.
    for (size_t x = 0; x < flights.size();) {
      if (hipEventQuery(flights[x].done) == hipSuccess) {
        Flight f = flights[x];
        flights.erase(flights.begin() + x);
        lk.unlock();
        finish(f);
        lk.lock();
        cv.notify_all();
      } else {
        ++x;
      }
    }
    // Admission in arrival order while memory and in-flight slots allow;
    // requests with the parameters of the first admitted one form the batch.
    if (!pending.empty() && (int)flights.size() < max_inflight) {
      std::vector<Request *> group;
      int64_t add = 0;
      for (auto it = pending.begin(); it != pending.end();) {
        Request *r = *it;
        if (r->state == RState::Cancelled) { it = pending.erase(it); continue; }
        if (!group.empty() && !same_params(r->p, group[0]->p)) { ++it; continue; }
        if (used + add + r->tb_bytes > budget) break;
        add += r->tb_bytes;
        r->state = RState::InFlight;
        group.push_back(r);
        it = pending.erase(it);
      }
      if (!group.empty()) {
        used += add;
        lk.unlock();
        launch(group);
        lk.lock();
        continue;
      }
    }
    if (flights.empty()) cv.wait(lk, [&] { return stop || !pending.empty(); });
    else cv.wait_for(lk, std::chrono::microseconds(50));
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
  for (auto &kv : s->all) delete kv.second;
  for (hipStream_t st : s->free_streams) (void)hipStreamDestroy(st);
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
  s->all.erase(it);  // delivered: the ticket is retired
  delete r;
  return TA_OK;
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
