// Host-code sanitizer driver (built with -fsanitize=address,undefined by
// tests/test_sanitizers.py; no GPU).  Exercises the host code that runs
// without a device: the CLI's FASTA reader / writer and CIGAR helpers
// (tweakseq_amd/csrc/fasta.h), the seeded workload generator
// (tweakseq_amd/csrc/workload.cpp) and the CPU oracle (oracle/*.c), on
// random and edge-case inputs.  Prints "sanitize ok" and exits 0 when every
// check holds; a sanitizer report aborts it.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tweakalign_workload.h"
#include "../../oracle/oracle.h"
#include "../../tweakseq_amd/csrc/fasta.h"

#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      exit(1);                                                        \
    }                                                                 \
  } while (0)

static void fasta_checks(std::mt19937 &rng) {
  const char *alpha = "ACDEFGHIKLMNPQRSTVWY-.";
  for (int t = 0; t < 200; ++t) {
    std::vector<Record> recs;
    const int n = 1 + (int)(rng() % 5);
    for (int r = 0; r < n; ++r) {
      Record x;
      x.header = (rng() & 1 ? ">" : ";") + std::string("seq") + std::to_string(r) + " descr";
      const int len = (int)(rng() % 300);
      for (int i = 0; i < len; ++i) x.seq.push_back(alpha[rng() % 22]);
      if (x.seq.empty()) x.seq = "A";
      recs.push_back(x);
    }
    std::ostringstream o;
    write_fasta(o, recs);
    std::istringstream in(o.str());
    std::vector<Record> back;
    read_fasta(in, back);
    CHECK(back.size() == recs.size());
    for (size_t r = 0; r < recs.size(); ++r) {
      CHECK(back[r].header == recs[r].header);
      CHECK(back[r].seq == recs[r].seq);
      const std::string d = dealign(back[r].seq);
      CHECK(d.find('-') == std::string::npos && d.find('.') == std::string::npos);
    }
  }
  // odd inputs: blank lines, comment-only records, no trailing newline, '\r'
  const char *odd[] = {"", "\n\n", ">only header", ";c1\n;c2\nACD\n", ">a\r\nAC\r\nDE\r\n>b\n\nWW", "garbage\n>x\nA"};
  for (const char *s : odd) {
    std::istringstream in(s);
    std::vector<Record> v;
    read_fasta(in, v);
    std::ostringstream o;
    write_fasta(o, v);
  }
  // CIGAR text -> ops -> rows
  CHECK(expand("3M2I1D", 6) == "MMMIID");
  CHECK(consumes("MMMIID", 5, 4) && !consumes("MMMIID", 5, 5) && !consumes("MX", 1, 1));
  std::string ra, rb;
  render("ACDEF", "ACW", "MMIID", ra, rb);
  CHECK(ra == "ACDE-" && rb == "AC--W");
}

static void workload_checks() {
  const int cfg[][3] = {{1, 0, 0}, {2, 20, 0}, {3, 60, 0}, {4, 8, 300}, {5, 2, 3000}};
  for (auto &c : cfg) {
    ta_workload *w = nullptr;
    CHECK(ta_workload_create(c[0], (uint64_t)c[0], c[1], c[2], &w) == 0);
    const int n = ta_workload_npairs(w);
    CHECK(n > 0);
    std::vector<ta_pair> pairs(n);
    CHECK(ta_workload_pairs(w, pairs.data()) == 0);
    for (auto &p : pairs) CHECK(p.la > 0 && p.lb > 0 && p.a && p.b);
    const int R = ta_workload_nrequests(w);
    for (int q = 0; q < R; ++q) {
      double at;
      int32_t first, count, mode;
      CHECK(ta_workload_request(w, q, &at, &first, &count, &mode) == 0);
      CHECK(first >= 0 && count >= 1 && first + count <= n);
    }
    ta_workload_destroy(w);
  }
  ta_workload *w = nullptr;
  CHECK(ta_workload_create(9, 1, 0, 0, &w) != 0);
}

static void oracle_checks(std::mt19937 &rng) {
  std::vector<int8_t> S(23 * 23);
  for (int r = 0; r < 23; ++r)
    for (int c = 0; c < 23; ++c) S[r * 23 + c] = (int8_t)(r == c ? 5 : ((r + c) % 3) - 2);
  for (int t = 0; t < 300; ++t) {
    or_params p{(int32_t)(t & 1), 0, -10, -1, 2, -3, 1, 0};
    const int m = (int)(rng() % 120), n = (int)(rng() % 120);
    std::vector<uint8_t> a(m + 1), b(n + 1);
    for (int i = 0; i < m; ++i) a[i] = (uint8_t)(rng() % 23);
    for (int j = 0; j < n; ++j) b[j] = (uint8_t)(rng() % 23);
    or_result res;
    std::vector<char> cig(2 * (m + n) + 16);
    CHECK(or_align(&p, S.data(), 23, a.data(), m, b.data(), n, &res, cig.data(), (int64_t)cig.size()) == 0);
    int32_t sc = 0;
    if (res.cigar_len > 0 || p.mode == 0) {
      CHECK(or_rescore(&p, S.data(), 23, a.data(), b.data(), &res, cig.data(), &sc) == 0);
      CHECK(sc == res.score);
    }
    // too small a CIGAR buffer reports the size needed
    if (res.cigar_len > 1) {
      or_result r2;
      CHECK(or_align(&p, S.data(), 23, a.data(), m, b.data(), n, &r2, cig.data(), 1) == -3);
    }
  }
  // batch driver, consensus and progressive alignment on small inputs
  std::vector<uint8_t> codes;
  std::vector<int64_t> off;
  std::vector<int32_t> len;
  for (int s = 0; s < 12; ++s) {
    off.push_back((int64_t)codes.size());
    const int L = 5 + (int)(rng() % 60);
    len.push_back(L);
    for (int i = 0; i < L; ++i) codes.push_back((uint8_t)(rng() % 20));
  }
  codes.push_back(0);
  std::vector<int32_t> joins(22), col(codes.size());
  int32_t alen = 0;
  CHECK(or_msa(S.data(), 23, -10, -1, codes.data(), off.data(), len.data(), 12, joins.data(), &alen, col.data()) == 0);
  CHECK(alen > 0);
  std::vector<int64_t> ao, bo;
  std::vector<int32_t> la, lb;
  for (int k = 0; k + 1 < 12; ++k) {
    ao.push_back(off[k]); la.push_back(len[k]);
    bo.push_back(off[k + 1]); lb.push_back(len[k + 1]);
  }
  std::vector<or_result> rr(11);
  std::vector<int64_t> co(11), cc(11);
  int64_t tot = 0;
  for (int k = 0; k < 11; ++k) { co[k] = tot; cc[k] = 2 * (la[k] + lb[k]) + 16; tot += cc[k]; }
  std::vector<char> cb(tot);
  or_params p{1, 0, -11, -1, 2, -3, 1, 0};
  CHECK(or_align_batch(&p, S.data(), 23, codes.data(), ao.data(), la.data(), bo.data(), lb.data(), 11, rr.data(),
                       cb.data(), co.data(), cc.data(), 2) == 0);
  std::vector<int32_t> map26(26);
  for (int c = 0; c < 26; ++c) map26[c] = c % 23;
  const int nseq = 6, L = 40;
  std::string msa;
  for (int s = 0; s < nseq; ++s)
    for (int i = 0; i < L; ++i) msa.push_back("ACDEFGH-"[rng() % 8]);
  std::vector<char> cons(L);
  CHECK(or_consensus(S.data(), map26.data(), msa.data(), nseq, L, -1.0, cons.data()) == 0);
}

int main() {
  std::mt19937 rng(12345);
  fasta_checks(rng);
  workload_checks();
  oracle_checks(rng);
  printf("sanitize ok\n");
  return 0;
}
