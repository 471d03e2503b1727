// Call -> CIGAR latency of one pair, timed in C around the C ABI: the
// 300 x 320 global protein pair of the seeded workloads (configuration 1),
// through ta_align_pair (the single-pair path) and through ta_align_batch
// with one pair.  Prints one JSON line with medians and 99th percentiles.
// Usage: tweakalign-pair-bench [repeats]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/tweakalign.h"
#include "../../include/tweakalign_workload.h"

using Clock = std::chrono::steady_clock;

int main(int argc, char **argv) {
  const int reps = argc > 1 ? std::max(10, atoi(argv[1])) : 2000;
  ta_workload *w = nullptr;
  if (ta_workload_create(1, 1, 0, 0, &w)) return 1;
  ta_pair pr;
  ta_params p;
  ta_workload_pairs(w, &pr);
  ta_workload_params(w, &p);
  ta_ctx *ctx = nullptr;
  if (ta_create(&ctx, 0)) { fprintf(stderr, "no device\n"); return 1; }
  std::vector<char> cig(2 * (pr.la + pr.lb) + 1);
  ta_result res;
  int64_t used = 0;
  auto time_it = [&](bool pair_path, std::vector<double> &us) {
    for (int r = 0; r < reps + 50; ++r) {
      const Clock::time_point t0 = Clock::now();
      const int rc = pair_path ? ta_align_pair(ctx, &p, pr.a, pr.la, pr.b, pr.lb, &res, cig.data(),
                                               (int64_t)cig.size(), &used)
                               : ta_align_batch(ctx, &p, &pr, 1, &res, cig.data(), (int64_t)cig.size(), &used);
      const double dt = std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
      if (rc) { fprintf(stderr, "rc %d: %s\n", rc, ta_strerror(rc)); return false; }
      if (r >= 50) us.push_back(dt);  // the first calls warm up
    }
    std::sort(us.begin(), us.end());
    return true;
  };
  std::vector<double> a, b;
  if (!time_it(true, a) || !time_it(false, b)) return 2;
  auto q = [](const std::vector<double> &v, double f) { return v[std::min(v.size() - 1, (size_t)(f * (v.size() - 1)))]; };
  const double cells = (double)pr.la * pr.lb;
  printf("{\"pair\": \"%dx%d global protein, traceback\", \"repeats\": %d, \"align_pair_us_p50\": %.1f, "
         "\"align_pair_us_p99\": %.1f, \"align_pair_gcups\": %.3f, \"align_batch_us_p50\": %.1f, "
         "\"align_batch_us_p99\": %.1f, \"score\": %d}\n",
         pr.la, pr.lb, reps, q(a, 0.5), q(a, 0.99), cells / q(a, 0.5) / 1e3, q(b, 0.5), q(b, 0.99), res.score);
  ta_destroy(ctx);
  ta_workload_destroy(w);
  return 0;
}
