// Seeded synthetic workloads (include/tweakalign_workload.h).
#include "../../include/tweakalign_workload.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Rng {  // xoshiro256** (Blackman & Vigna 2018), seeded by splitmix64
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9e3779b97f4a7c15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
      s[i] = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
  int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1)); }
  double normal() {  // Box-Muller
    double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int geometric(double mean) {  // support 1, 2, ...; P(k) = (1-q)^(k-1) q, q = 1/mean
    const double q = 1.0 / mean;
    if (q >= 1.0) return 1;
    const double u = 1.0 - uniform();  // (0, 1]
    return 1 + (int)std::floor(std::log(u) / std::log(1.0 - q));
  }
};

// BLOSUM62 background frequencies (Henikoff & Henikoff 1992), order ARNDCQEGHILKMFPSTWYV.
const char kAA[] = "ARNDCQEGHILKMFPSTWYV";
const double kFreq[20] = {.074, .052, .045, .053, .025, .034, .054, .074, .026, .068,
                          .099, .058, .025, .047, .039, .057, .051, .013, .032, .073};

struct Alph {
  std::string letters;
  std::vector<uint64_t> cdf;  // cumulative thresholds on 32-bit draws
  uint8_t first[256];         // per top byte of a draw: the first threshold it can fall under
  explicit Alph(bool dna) {
    if (dna) {
      letters = "ACGT";
      for (int i = 1; i <= 4; ++i) cdf.push_back((uint64_t)i << 30);
    } else {
      letters = kAA;
      double sum = 0, acc = 0;
      for (double f : kFreq) sum += f;
      for (double f : kFreq) {
        acc += f / sum;
        cdf.push_back((uint64_t)std::llround(acc * 4294967296.0));
      }
      cdf.back() = 1ull << 32;
    }
    // Draws with top byte t are >= t << 24, so no threshold at or below that
    // can hold them: the scan starts past those (same letter as a full scan).
    for (int t = 0; t < 256; ++t) {
      size_t i = 0;
      while (i + 1 < cdf.size() && cdf[i] <= ((uint64_t)t << 24)) ++i;
      first[t] = (uint8_t)i;
    }
  }
  char draw(Rng &r) const {
    const uint64_t u = r.next() >> 32;
    for (size_t i = first[u >> 24]; i < cdf.size(); ++i)
      if (u < cdf[i]) return letters[i];
    return letters.back();
  }
  char draw_other(Rng &r, char old) const {
    for (;;) {
      char c = draw(r);
      if (c != old) return c;
    }
  }
};

std::string random_seq(Rng &r, const Alph &A, int len) {
  std::string s(len, 'A');
  for (int i = 0; i < len; ++i) s[i] = A.draw(r);
  return s;
}

std::string mutate(Rng &r, const Alph &A, const std::string &p, double psub, double pindel, double mean) {
  std::string out;
  out.reserve(p.size() + 16);
  size_t i = 0;
  while (i < p.size()) {
    if (r.uniform() < pindel) {
      const int k = r.geometric(mean);
      if (r.uniform() < 0.5) {
        for (int t = 0; t < k; ++t) out.push_back(A.draw(r));
      } else {
        i += (size_t)k;
        continue;
      }
    }
    char c = p[i];
    if (r.uniform() < psub) c = A.draw_other(r, c);
    out.push_back(c);
    ++i;
  }
  if (out.empty()) out.push_back(A.draw(r));
  return out;
}

struct Request {
  double arrival;
  int32_t first, count, mode;
};

}  // namespace

struct ta_workload {
  int config = 0;
  ta_params params{};
  std::vector<std::string> seqs;
  std::vector<std::pair<int, int>> pairs;  // indices into seqs
  std::vector<Request> reqs;
};

extern "C" {

int ta_workload_create(int32_t config, uint64_t seed, int32_t size, int32_t max_len, ta_workload **out) {
  if (!out || config < 1 || config > 5 || size < 0 || max_len < 0) return TA_E_INVALID;
  ta_workload *w = new ta_workload();
  w->config = config;
  Rng r(seed);
  const Alph prot(false), dna(true);
  ta_params &p = w->params;
  p.alphabet = TA_PROTEIN; p.gap_open = -10; p.gap_extend = -1; p.match = 2; p.mismatch = -3;
  p.traceback = 1; p.lanes = TA_LANES_AUTO; p.mode = TA_GLOBAL;
  auto add_pair = [&](std::string a, std::string b) {
    w->seqs.push_back(std::move(a));
    w->seqs.push_back(std::move(b));
    w->pairs.push_back({(int)w->seqs.size() - 2, (int)w->seqs.size() - 1});
  };
  switch (config) {
    case 1: {
      add_pair(random_seq(r, prot, 300), random_seq(r, prot, 320));
      break;
    }
    case 2: {
      const int nseq = size ? size : 512;
      p.traceback = 0;
      for (int s = 0; s < nseq; ++s) {
        if (s > 0 && r.uniform() < 0.3) {
          const int parent = (int)r.range(0, s - 1);
          w->seqs.push_back(mutate(r, prot, w->seqs[parent], 0.20, 0.05, 3.0));
        } else {
          w->seqs.push_back(random_seq(r, prot, (int)r.range(100, 600)));
        }
      }
      for (int i = 0; i < nseq; ++i)
        for (int j = i + 1; j < nseq; ++j) w->pairs.push_back({i, j});
      break;
    }
    case 3: {
      const int np = size ? size : 65536;
      p.mode = TA_LOCAL; p.gap_open = -11;
      for (int k = 0; k < np; ++k) {
        double x = std::exp(std::log(350.0) + 0.6 * r.normal());
        int la = (int)std::llround(x);
        la = la < 30 ? 30 : (la > 2000 ? 2000 : la);
        std::string a = random_seq(r, prot, la);
        std::string b = mutate(r, prot, a, 0.25, 0.04, 3.0);
        add_pair(std::move(a), std::move(b));
      }
      break;
    }
    case 4: {
      // Arrivals, pair counts and modes come from the workload's stream; the
      // sequences of request q from a stream of their own (seeded by the
      // workload seed and q), so that requests are generated on all host
      // threads (1.92 M pairs, ~9.7 GB of residues at full size).
      const int nreq = size ? size : 120000;
      const int lmax = max_len ? max_len : 5000;
      const int lmin = lmax < 50 ? 1 : 50;
      double t = 0;
      int32_t npairs = 0;
      for (int q = 0; q < nreq; ++q) {
        t += -std::log(1.0 - r.uniform()) / 2000.0;
        int cnt = r.geometric(16.0);
        if (cnt > 256) cnt = 256;
        const int mode = r.uniform() < 0.5 ? TA_GLOBAL : TA_LOCAL;
        w->reqs.push_back(Request{t, npairs, cnt, mode});
        npairs += cnt;
      }
      w->seqs.resize(2 * (size_t)npairs);
      w->pairs.resize(npairs);
      for (int32_t k = 0; k < npairs; ++k) w->pairs[k] = {2 * k, 2 * k + 1};
      const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      std::vector<std::thread> th;
      for (unsigned x = 0; x < nth; ++x)
        th.emplace_back([&, x] {
          for (int q = (int)x; q < nreq; q += (int)nth) {
            Rng rq(seed ^ (0x9e3779b97f4a7c15ull * (uint64_t)(q + 1)));
            const Request &R = w->reqs[q];
            for (int k = R.first; k < R.first + R.count; ++k) {
              w->seqs[2 * (size_t)k] = random_seq(rq, prot, (int)rq.range(lmin, lmax));
              w->seqs[2 * (size_t)k + 1] = random_seq(rq, prot, (int)rq.range(lmin, lmax));
            }
          }
        });
      for (auto &x : th) x.join();
      break;
    }
    case 5: {
      const int np = size ? size : 256;
      const int lmax = max_len ? max_len : 100000;
      const int lmin = lmax >= 10000 ? 10000 : lmax / 10 + 1;
      p.alphabet = TA_DNA; p.gap_open = -5; p.gap_extend = -2; p.match = 2; p.mismatch = -3;
      for (int k = 0; k < np; ++k) {
        std::string a = random_seq(r, dna, (int)r.range(lmin, lmax));
        std::string b = mutate(r, dna, a, 0.10, 0.02, 5.0);
        add_pair(std::move(a), std::move(b));
      }
      break;
    }
  }
  if (w->reqs.empty()) w->reqs.push_back(Request{0.0, 0, (int32_t)w->pairs.size(), p.mode});
  *out = w;
  return TA_OK;
}

void ta_workload_destroy(ta_workload *w) { delete w; }

int32_t ta_workload_npairs(const ta_workload *w) { return w ? (int32_t)w->pairs.size() : 0; }

int ta_workload_pairs(const ta_workload *w, ta_pair *out) {
  if (!w || !out) return TA_E_INVALID;
  for (size_t k = 0; k < w->pairs.size(); ++k) {
    const std::string &a = w->seqs[w->pairs[k].first], &b = w->seqs[w->pairs[k].second];
    out[k].a = a.data(); out[k].la = (int32_t)a.size();
    out[k].b = b.data(); out[k].lb = (int32_t)b.size();
  }
  return TA_OK;
}

int ta_workload_params(const ta_workload *w, ta_params *out) {
  if (!w || !out) return TA_E_INVALID;
  *out = w->params;
  return TA_OK;
}

int32_t ta_workload_nrequests(const ta_workload *w) { return w ? (int32_t)w->reqs.size() : 0; }

int ta_workload_request(const ta_workload *w, int32_t q, double *arrival_s, int32_t *first, int32_t *count,
                        int32_t *mode) {
  if (!w || q < 0 || q >= (int32_t)w->reqs.size()) return TA_E_INVALID;
  const Request &rq = w->reqs[q];
  if (arrival_s) *arrival_s = rq.arrival;
  if (first) *first = rq.first;
  if (count) *count = rq.count;
  if (mode) *mode = rq.mode;
  return TA_OK;
}

}  // extern "C"
