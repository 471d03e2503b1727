// Open-loop serving benchmark: submits the seeded request workload
// (Poisson arrivals, include/tweakalign_workload.h configuration 4) at its
// scheduled times, polls completions, and prints one JSON line with p50/p99
// latency (completion - scheduled arrival, so queueing delay counts) and
// GCUPS.  Usage: tweakalign-serve-bench [requests] [max_len] [budget_GB] [inflight]
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tweakalign.h"
#include "../../include/tweakalign_workload.h"

using Clock = std::chrono::steady_clock;

int main(int argc, char **argv) {
  const int nreq = argc > 1 ? atoi(argv[1]) : 4000;
  const int max_len = argc > 2 ? atoi(argv[2]) : 0;
  const double budget_gb = argc > 3 ? atof(argv[3]) : 64.0;
  const int inflight = argc > 4 ? atoi(argv[4]) : 4;

  ta_workload *w = nullptr;
  if (ta_workload_create(4, 4, nreq, max_len, &w)) return 1;
  const int np = ta_workload_npairs(w);
  std::vector<ta_pair> pairs(np);
  ta_workload_pairs(w, pairs.data());
  ta_params base;
  ta_workload_params(w, &base);
  const int R = ta_workload_nrequests(w);
  struct Req { double at; int first, count, mode; uint64_t ticket; double done, sub; int64_t cells; };
  std::vector<Req> rq(R);
  int64_t total_cells = 0;
  for (int q = 0; q < R; ++q) {
    ta_workload_request(w, q, &rq[q].at, &rq[q].first, &rq[q].count, &rq[q].mode);
    rq[q].cells = 0;
    for (int k = 0; k < rq[q].count; ++k)
      rq[q].cells += (int64_t)pairs[rq[q].first + k].la * pairs[rq[q].first + k].lb;
    total_cells += rq[q].cells;
    rq[q].done = -1;
  }

  ta_ctx *ctx = nullptr;
  if (ta_create(&ctx, 0)) { fprintf(stderr, "no device\n"); return 1; }
  ta_serve_config(ctx, (int64_t)(budget_gb * (1ll << 30)), inflight);

  // Warm the device and the memory cache with one small request.
  {
    ta_params p = base;
    uint64_t t;
    ta_submit(ctx, &p, &pairs[0], 1, &t);
    std::vector<ta_result> res(1);
    std::vector<char> cig(1 << 16);
    int64_t used;
    while (ta_poll(ctx, t, res.data(), cig.data(), (int64_t)cig.size(), &used) == TA_E_PENDING)
      std::this_thread::sleep_for(std::chrono::microseconds(100));
  }

  // Every delivered CIGAR must consume exactly the aligned span of both
  // sequences (query: M and I; target: M and D), and an empty local result
  // has no CIGAR: a check of every served pair at full size.
  auto cigar_ok = [&](const ta_result &r, const ta_pair &q, const char *text) {
    int64_t qa = 0, qb = 0, run = 0;
    for (int64_t c = 0; c < r.cigar_len; ++c) {
      const char ch = text[r.cigar_off + c];
      if (ch >= '0' && ch <= '9') { run = 10 * run + (ch - '0'); continue; }
      if (ch != 'M' && ch != 'I' && ch != 'D') return false;
      if (ch != 'D') qa += run;
      if (ch != 'I') qb += run;
      run = 0;
    }
    if (run) return false;
    if (r.flags & TA_F_EMPTY) return r.cigar_len == 0;
    return r.a_begin >= 0 && r.a_end < q.la && r.b_begin >= 0 && r.b_end < q.lb &&
           qa == r.a_end - r.a_begin + 1 && qb == r.b_end - r.b_begin + 1;
  };
  std::atomic<int64_t> bad_cigars{0};

  std::atomic<int> submitted{0};
  std::atomic<bool> failed{false};
  const Clock::time_point t0 = Clock::now();
  auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };

  std::thread poller([&] {
    std::vector<ta_result> res(300);
    std::vector<char> cig(1 << 24);
    int finished = 0, lo = 0;
    while (finished < R && !failed) {
      const int s = submitted.load();
      bool any = false;
      while (lo < s && rq[lo].done >= 0) ++lo;
      for (int q = lo; q < s; ++q) {
        if (rq[q].done >= 0) continue;
        int64_t used = 0;
        const int rc = ta_poll(ctx, rq[q].ticket, res.data(), cig.data(), (int64_t)cig.size(), &used);
        if (rc == TA_E_PENDING) continue;
        if (rc != TA_OK) { fprintf(stderr, "request %d failed: %s\n", q, ta_strerror(rc)); failed = true; break; }
        rq[q].done = now_s();
        ++finished;
        any = true;
        for (int k = 0; k < rq[q].count; ++k)
          if (!cigar_ok(res[k], pairs[rq[q].first + k], cig.data())) ++bad_cigars;
      }
      if (!any) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
  });

  for (int q = 0; q < R && !failed; ++q) {
    const double wait = rq[q].at - now_s();
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    ta_params p = base;
    p.mode = rq[q].mode;
    rq[q].sub = now_s();
    if (ta_submit(ctx, &p, &pairs[rq[q].first], rq[q].count, &rq[q].ticket)) { failed = true; break; }
    submitted.store(q + 1);
  }
  poller.join();
  const double end = now_s();
  ta_destroy(ctx);
  if (failed) return 2;

  std::vector<double> lat(R);
  double last = 0;
  for (int q = 0; q < R; ++q) {
    lat[q] = rq[q].done - rq[q].at;
    last = std::max(last, rq[q].done);
  }
  // the five slowest requests: arrival, submission lag, latency, pairs, mode, longest side
  std::vector<int> order(R);
  for (int q = 0; q < R; ++q) order[q] = q;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return lat[x] > lat[y]; });
  std::string slow;
  for (int k = 0; k < std::min(R, 5); ++k) {
    const Req &r = rq[order[k]];
    int longest = 0;
    for (int i = 0; i < r.count; ++i) longest = std::max({longest, pairs[r.first + i].la, pairs[r.first + i].lb});
    char b[160];
    snprintf(b, sizeof b, "%s[%d, %.3f, %.2f, %.2f, %d, %d, %d]", k ? ", " : "", order[k], r.at, (r.sub - r.at) * 1e3,
             lat[order[k]] * 1e3, r.count, r.mode, longest);
    slow += b;
  }
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double p) { return lat[std::min(R - 1, (int)(p * (R - 1) + 0.5))]; };
  const double span = rq[R - 1].at;
  printf("{\"requests\": %d, \"pairs\": %d, \"cells\": %lld, \"arrival_span_s\": %.3f, \"makespan_s\": %.3f, "
         "\"offered_gcups\": %.2f, \"served_gcups\": %.2f, \"p50_ms\": %.2f, \"p99_ms\": %.2f, \"max_ms\": %.2f, "
         "\"budget_gb\": %.1f, \"inflight\": %d, \"max_len\": %d, \"p90_ms\": %.2f, \"bad_cigars\": %lld, "
         "\"slowest\": [%s], \"slowest_fields\": \"request, arrival_s, submit_lag_ms, latency_ms, pairs, mode, longest\"}\n",
         R, np, (long long)total_cells, span, last, total_cells / span / 1e9, total_cells / last / 1e9,
         pct(0.5) * 1e3, pct(0.99) * 1e3, lat[R - 1] * 1e3, budget_gb, inflight, max_len ? max_len : 5000,
         pct(0.9) * 1e3, (long long)bad_cigars.load(), slow.c_str());
  (void)end;
  ta_workload_destroy(w);
  return 0;
}
