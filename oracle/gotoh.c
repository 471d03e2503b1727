/*
 * CPU oracle: scalar Gotoh global / local alignment with full traceback.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h for who may call it and what pins it.
 *
 * Deliberately the plainest possible restatement: full (m+1) x (n+1) matrix
 * walk in row-major order, one byte of direction code per cell, explicit stop
 * codes for the local zero floor.  Nothing here is shared with the HIP path.
 *
 * Recurrence (rows i = 1..m over a, columns j = 1..n over b):
 *   E(i,j) = max(H(i,j-1) + open, E(i,j-1) + ext)      gap consuming b -> 'D'
 *   F(i,j) = max(H(i-1,j) + open, F(i-1,j) + ext)      gap consuming a -> 'I'
 *   H(i,j) = max(H(i-1,j-1) + S(a_i,b_j), E, F [, 0 local])
 * Ties: E and F prefer the opening term; H prefers diag > E > F (first wins,
 * as tweakseq/Core/Utility.cpp:44-50 find_largest_int); local H == 0 is a stop.
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEG_INF (INT32_MIN / 4)

/* Protein: tweakseq/Core/Annotations/Consensus.cpp:99-104 maps 'A'..'Z'
 * through BLOSUM62map and anything else to a sentinel; here anything else is
 * an error.  Lowercase is folded to uppercase first. */
int or_encode(const int32_t *letter_map26, int32_t alphabet, const char *s,
              int32_t len, uint8_t *out) {
  for (int32_t i = 0; i < len; ++i) {
    int c = (unsigned char)s[i];
    if (c >= 'a' && c <= 'z') c -= 'a' - 'A';
    if (c < 'A' || c > 'Z') return -2;
    if (alphabet == 0) {
      out[i] = (uint8_t)letter_map26[c - 'A'];
    } else {
      switch (c) {
        case 'A': out[i] = 0; break;
        case 'C': out[i] = 1; break;
        case 'G': out[i] = 2; break;
        case 'T': out[i] = 3; break;
        default: return -2;
      }
    }
  }
  return 0;
}

/* Emit one run into a reversed run list. */
typedef struct { char op; int32_t len; } run_t;

static int push_op(run_t **runs, int32_t *nruns, int32_t *capruns, char op) {
  if (*nruns > 0 && (*runs)[*nruns - 1].op == op) {
    (*runs)[*nruns - 1].len++;
    return 0;
  }
  if (*nruns == *capruns) {
    int32_t nc = *capruns ? *capruns * 2 : 64;
    run_t *r = (run_t *)realloc(*runs, (size_t)nc * sizeof(run_t));
    if (!r) return -5;
    *runs = r;
    *capruns = nc;
  }
  (*runs)[*nruns].op = op;
  (*runs)[*nruns].len = 1;
  (*nruns)++;
  return 0;
}

static int push_n(run_t **runs, int32_t *nruns, int32_t *capruns, char op, int32_t k) {
  for (int32_t t = 0; t < k; ++t) {
    int rc = push_op(runs, nruns, capruns, op);
    if (rc) return rc;
  }
  return 0;
}

static int ndigits(int32_t v) {
  int d = 1;
  while (v >= 10) { v /= 10; ++d; }
  return d;
}

int or_align(const or_params *p, const int8_t *smat, int32_t K,
             const uint8_t *a, int32_t m, const uint8_t *b, int32_t n,
             or_result *res, char *cigar, int64_t cap) {
  if (!p || !smat || !res || m < 0 || n < 0 || K <= 0) return -1;
  if (p->gap_open > 0 || p->gap_extend > 0) return -1;
  const int local = p->mode == 1;
  const int32_t open = p->gap_open, ext = p->gap_extend;
  const int want_tb = p->traceback != 0;

  memset(res, 0, sizeof(*res));
  for (int32_t i = 0; i < m; ++i) if (a[i] >= K) return -2;
  for (int32_t j = 0; j < n; ++j) if (b[j] >= K) return -2;

  /* Rolling H row and F row; full direction-code matrix when tracing. */
  int32_t *Hprev = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  int32_t *Hcur = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  int32_t *Frow = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  uint8_t *tb = NULL;
  if (want_tb) tb = (uint8_t *)malloc((size_t)(m + 1) * (size_t)(n + 1));
  if (!Hprev || !Hcur || !Frow || (want_tb && !tb)) {
    free(Hprev); free(Hcur); free(Frow); free(tb);
    return -5;
  }
#define TB(i, j) tb[(size_t)(i) * (size_t)(n + 1) + (size_t)(j)]

  /* Row 0. */
  Hprev[0] = 0;
  for (int32_t j = 1; j <= n; ++j) {
    Hprev[j] = local ? 0 : open + (j - 1) * ext;
    Frow[j] = NEG_INF;
  }
  Frow[0] = NEG_INF;

  int32_t best = 0, bi = 0, bj = 0;
  for (int32_t i = 1; i <= m; ++i) {
    const int8_t *srow = smat + (size_t)a[i - 1] * (size_t)K;
    int32_t E = NEG_INF;
    Hcur[0] = local ? 0 : open + (i - 1) * ext;
    int32_t Hleft = Hcur[0];
    int32_t Hdiag = Hprev[0];
    for (int32_t j = 1; j <= n; ++j) {
      int32_t eo = Hleft + open, ee = E + ext;
      int ebit = ee > eo;
      E = ebit ? ee : eo;
      int32_t fo = Hprev[j] + open, fe = Frow[j] + ext;
      int fbit = fe > fo;
      int32_t F = fbit ? fe : fo;
      Frow[j] = F;
      int32_t h = Hdiag + srow[b[j - 1]];
      int src = 0;
      if (E > h) { h = E; src = 1; }
      if (F > h) { h = F; src = 2; }
      if (local && h <= 0) { h = 0; src = 3; }
      Hdiag = Hprev[j];
      Hcur[j] = h;
      Hleft = h;
      if (want_tb) TB(i, j) = (uint8_t)(src | (ebit << 2) | (fbit << 3));
      if (local && h > best) { best = h; bi = i; bj = j; }
    }
    int32_t *t = Hprev; Hprev = Hcur; Hcur = t;
  }

  if (local) {
    res->score = best;
    if (best == 0) {
      res->a_begin = res->a_end = res->b_begin = res->b_end = -1;
    } else {
      res->a_end = bi - 1;
      res->b_end = bj - 1;
    }
  } else {
    res->score = Hprev[n]; /* row m after the swap */
    if (m == 0) res->score = n == 0 ? 0 : open + (n - 1) * ext;
    res->a_begin = 0; res->a_end = m - 1;
    res->b_begin = 0; res->b_end = n - 1;
  }

  int rc = 0;
  if (want_tb && !(local && best == 0)) {
    run_t *runs = NULL;
    int32_t nruns = 0, capruns = 0;
    int32_t i = local ? bi : m, j = local ? bj : n;
    int state = 0; /* 0 H, 1 E, 2 F */
    for (;;) {
      if (local) {
        if (state == 0 && (i == 0 || j == 0 || (TB(i, j) & 3) == 3)) break;
      } else {
        if (i == 0) { rc = push_n(&runs, &nruns, &capruns, 'D', j); break; }
        if (j == 0) { rc = push_n(&runs, &nruns, &capruns, 'I', i); break; }
      }
      uint8_t code = TB(i, j);
      if (state == 0) {
        int src = code & 3;
        if (src == 0) {
          rc = push_op(&runs, &nruns, &capruns, 'M');
          --i; --j;
        } else {
          state = src; /* 1 E, 2 F: same cell, gap state */
        }
      } else if (state == 1) {
        rc = push_op(&runs, &nruns, &capruns, 'D');
        if (!(code & 4)) state = 0;
        --j;
      } else {
        rc = push_op(&runs, &nruns, &capruns, 'I');
        if (!(code & 8)) state = 0;
        --i;
      }
      if (rc) break;
    }
    if (local) { res->a_begin = i; res->b_begin = j; }
    if (!rc) {
      int64_t need = 0;
      for (int32_t k = 0; k < nruns; ++k) need += ndigits(runs[k].len) + 1;
      res->cigar_len = (int32_t)need;
      if (cigar) {
        if (need > cap) {
          rc = -3;
        } else {
          int64_t pos = 0;
          for (int32_t k = nruns - 1; k >= 0; --k) {
            int d = ndigits(runs[k].len);
            int32_t v = runs[k].len;
            for (int t = d - 1; t >= 0; --t) { cigar[pos + t] = (char)('0' + v % 10); v /= 10; }
            pos += d;
            cigar[pos++] = runs[k].op;
          }
        }
      }
    }
    free(runs);
  }
#undef TB
  free(Hprev); free(Hcur); free(Frow); free(tb);
  return rc;
}

int or_align_batch(const or_params *p, const int8_t *smat, int32_t K,
                   const uint8_t *codes, const int64_t *a_off, const int32_t *la,
                   const int64_t *b_off, const int32_t *lb, int32_t npairs,
                   or_result *res, char *cigar, const int64_t *cig_off,
                   const int64_t *cig_cap, int32_t nthreads) {
  int first_err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int32_t k = 0; k < npairs; ++k) {
    char *c = cigar ? cigar + cig_off[k] : NULL;
    int64_t cap = cigar ? cig_cap[k] : 0;
    int rc = or_align(p, smat, K, codes + a_off[k], la[k], codes + b_off[k], lb[k],
                      &res[k], c, cap);
    if (cigar) res[k].cigar_off = cig_off[k];
    if (rc) {
#ifdef _OPENMP
#pragma omp critical
#endif
      if (!first_err) first_err = rc;
    }
  }
  (void)nthreads;
  return first_err;
}

int or_rescore(const or_params *p, const int8_t *smat, int32_t K,
               const uint8_t *a, const uint8_t *b, const or_result *res,
               const char *cigar, int32_t *score) {
  int32_t i = res->a_begin, j = res->b_begin;
  int64_t s = 0;
  int32_t pos = 0;
  if (res->cigar_len == 0) {
    *score = 0;
    if (p->mode == 1) return res->a_begin == -1 ? 0 : -1;
    /* Global with nothing to align. */
    return (res->a_end - res->a_begin + 1 == 0 && res->b_end - res->b_begin + 1 == 0) ? 0 : -1;
  }
  while (pos < res->cigar_len) {
    int32_t k = 0;
    while (pos < res->cigar_len && cigar[pos] >= '0' && cigar[pos] <= '9')
      k = k * 10 + (cigar[pos++] - '0');
    if (pos >= res->cigar_len || k <= 0) return -1;
    char op = cigar[pos++];
    if (op == 'M') {
      for (int32_t t = 0; t < k; ++t, ++i, ++j) s += smat[(size_t)a[i] * (size_t)K + b[j]];
    } else if (op == 'I' || op == 'D') {
      /* Best the recurrence can give one maximal run of k gap positions:
       * open + (k-1) ext when extending pays, else k separate openings. */
      int64_t ext_run = p->gap_open + (int64_t)(k - 1) * p->gap_extend;
      int64_t reopen = (int64_t)k * p->gap_open;
      s += ext_run > reopen ? ext_run : reopen;
      if (op == 'I') i += k; else j += k;
    } else {
      return -1;
    }
  }
  if (i != res->a_end + 1 || j != res->b_end + 1) return -1;
  *score = (int32_t)s;
  return 0;
}
