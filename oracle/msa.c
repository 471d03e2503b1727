/*
 * Progressive multiple alignment -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * No counterpart in the reference: tweakseq hands N > 2 sequences to an
 * external MSA program (tweakseq/UI/SeqEditMainWin.cpp:1647-1659) and reads
 * the aligned FASTA back (tweakseq/Core/Project.cpp:908-1032).  This restates
 * the specification this project implements for that step, plainly:
 *
 *  1. s_ij: global Gotoh score of every pair i < j (or_align);
 *  2. sim_ij = s_ij * 1000000 / max(1, min(self_i, self_j)), truncated toward
 *     zero, self_i = sum over sequence i of S(c, c);
 *  3. UPGMA on sim: n - 1 times join the two live clusters of the largest
 *     mean similarity (cluster sums compared by cross-multiplication, so the
 *     choice is exact); ties go to the smallest (lower id, higher id); the
 *     t-th join creates cluster n + t, whose sums are the sums of its parts;
 *  4. a join aligns the lower id's profile (DP rows i, its columns) with the
 *     higher id's (DP columns j) by global Gotoh over columns: match score
 *     M(i,j) = sum_x sum_y cntA[i][x] cntB[j][y] S(x,y) (cnt: residues of
 *     each code in a column; gaps count for nothing), gap open / extend
 *     times nA * nB, recurrence, boundaries, ties and traceback exactly those
 *     of or_align (diag > E > F; E and F prefer opening).  'M' takes a column
 *     of each, 'D' a column of B against a gap in A's rows, 'I' a column of A
 *     against a gap in B's rows;
 *  5. the final profile's columns; every residue's column is reported.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NEG64 (INT64_MIN / 4)

typedef struct {
  int32_t nrows, len;
  int32_t *rows;  /* sequence ids */
  int32_t *cell;  /* nrows x len: residue code, or -1 for a gap */
} profile_t;

static void prof_free(profile_t *p) {
  free(p->rows);
  free(p->cell);
  p->rows = NULL;
  p->cell = NULL;
}

/* Global Gotoh over the columns of A (rows) and B (columns); ops written in
 * alignment order into ops[0..*nops).  Returns 0 or -5. */
static int align_profiles(const int8_t *smat, int32_t K, int64_t open, int64_t ext, const profile_t *A,
                          const profile_t *B, char *ops, int32_t *nops, int64_t *score) {
  const int32_t m = A->len, n = B->len;
  int64_t *cntA = (int64_t *)calloc((size_t)(m ? m : 1) * (size_t)K, sizeof(int64_t));
  int64_t *PB = (int64_t *)calloc((size_t)(n ? n : 1) * (size_t)K, sizeof(int64_t));
  int64_t *Hprev = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  int64_t *Hcur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  int64_t *Frow = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  uint8_t *tb = (uint8_t *)malloc((size_t)(m + 1) * (size_t)(n + 1));
  char *rev = (char *)malloc((size_t)(m + n + 1));
  int rc = 0;
  if (!cntA || !PB || !Hprev || !Hcur || !Frow || !tb || !rev) { rc = -5; goto out; }
  for (int32_t r = 0; r < A->nrows; ++r)
    for (int32_t i = 0; i < m; ++i) {
      const int32_t c = A->cell[(size_t)r * m + i];
      if (c >= 0) cntA[(size_t)i * K + c] += 1;
    }
  /* PB[j][x] = sum_y S(x, y) cntB[j][y] */
  for (int32_t r = 0; r < B->nrows; ++r)
    for (int32_t j = 0; j < n; ++j) {
      const int32_t c = B->cell[(size_t)r * n + j];
      if (c < 0) continue;
      for (int32_t x = 0; x < K; ++x) PB[(size_t)j * K + x] += smat[(size_t)x * K + c];
    }
#define TB(i, j) tb[(size_t)(i) * (size_t)(n + 1) + (size_t)(j)]
  Hprev[0] = 0;
  for (int32_t j = 1; j <= n; ++j) {
    Hprev[j] = open + (j - 1) * ext;
    Frow[j] = NEG64;
  }
  Frow[0] = NEG64;
  for (int32_t i = 1; i <= m; ++i) {
    const int64_t *ca = cntA + (size_t)(i - 1) * K;
    int64_t E = NEG64;
    Hcur[0] = open + (i - 1) * ext;
    int64_t Hleft = Hcur[0], Hdiag = Hprev[0];
    for (int32_t j = 1; j <= n; ++j) {
      const int64_t *pb = PB + (size_t)(j - 1) * K;
      int64_t mij = 0;
      for (int32_t x = 0; x < K; ++x) mij += ca[x] * pb[x];
      const int64_t eo = Hleft + open, ee = E + ext;
      const int ebit = ee > eo;
      E = ebit ? ee : eo;
      const int64_t fo = Hprev[j] + open, fe = Frow[j] + ext;
      const int fbit = fe > fo;
      const int64_t F = fbit ? fe : fo;
      Frow[j] = F;
      int64_t h = Hdiag + mij;
      int src = 0;
      if (E > h) { h = E; src = 1; }
      if (F > h) { h = F; src = 2; }
      Hdiag = Hprev[j];
      Hcur[j] = h;
      Hleft = h;
      TB(i, j) = (uint8_t)(src | (ebit << 2) | (fbit << 3));
    }
    int64_t *t = Hprev; Hprev = Hcur; Hcur = t;
  }
  *score = m == 0 ? (n == 0 ? 0 : open + (n - 1) * ext) : Hprev[n];
  {
    int32_t i = m, j = n, k = 0;
    int state = 0;
    for (;;) {
      if (i == 0) { while (j-- > 0) rev[k++] = 'D'; break; }
      if (j == 0) { while (i-- > 0) rev[k++] = 'I'; break; }
      const uint8_t code = TB(i, j);
      if (state == 0) {
        const int src = code & 3;
        if (src == 0) { rev[k++] = 'M'; --i; --j; }
        else state = src;
      } else if (state == 1) {
        rev[k++] = 'D';
        if (!(code & 4)) state = 0;
        --j;
      } else {
        rev[k++] = 'I';
        if (!(code & 8)) state = 0;
        --i;
      }
    }
    for (int32_t t = 0; t < k; ++t) ops[t] = rev[k - 1 - t];
    *nops = k;
  }
#undef TB
out:
  free(cntA); free(PB); free(Hprev); free(Hcur); free(Frow); free(tb); free(rev);
  return rc;
}

int or_msa(const int8_t *smat, int32_t K, int32_t gap_open, int32_t gap_extend, const uint8_t *codes,
           const int64_t *off, const int32_t *len, int32_t nseq, int32_t *joins, int32_t *alen,
           int32_t *col) {
  if (!smat || K <= 0 || !codes || !off || !len || nseq <= 0 || !alen || !col) return -1;
  if (gap_open > 0 || gap_extend > 0) return -1;
  for (int32_t s = 0; s < nseq; ++s) {
    if (len[s] < 0) return -1;
    for (int32_t k = 0; k < len[s]; ++k) if (codes[off[s] + k] >= K) return -2;
  }
  const int32_t nc = 2 * nseq - 1;
  int64_t *sum = (int64_t *)calloc((size_t)nc * (size_t)nc, sizeof(int64_t));
  int32_t *size = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
  char *live = (char *)calloc((size_t)nc, 1);
  int64_t *self = (int64_t *)calloc((size_t)nseq, sizeof(int64_t));
  profile_t *prof = (profile_t *)calloc((size_t)nc, sizeof(profile_t));
  int rc = 0;
  if (!sum || !size || !live || !self || !prof) { rc = -5; goto out; }

  /* 1-2: pair scores and similarities */
  for (int32_t s = 0; s < nseq; ++s)
    for (int32_t k = 0; k < len[s]; ++k) {
      const int c = codes[off[s] + k];
      self[s] += smat[(size_t)c * K + c];
    }
  {
    or_params p;
    memset(&p, 0, sizeof(p));
    p.mode = 0;
    p.gap_open = gap_open;
    p.gap_extend = gap_extend;
    for (int32_t a = 0; a < nseq; ++a)
      for (int32_t b = a + 1; b < nseq; ++b) {
        or_result r;
        rc = or_align(&p, smat, K, codes + off[a], len[a], codes + off[b], len[b], &r, NULL, 0);
        if (rc) goto out;
        int64_t d = self[a] < self[b] ? self[a] : self[b];
        if (d < 1) d = 1;
        const int64_t sim = (int64_t)r.score * 1000000 / d;
        sum[(size_t)a * nc + b] = sum[(size_t)b * nc + a] = sim;
      }
  }
  /* leaves */
  for (int32_t s = 0; s < nseq; ++s) {
    size[s] = 1;
    live[s] = 1;
    prof[s].nrows = 1;
    prof[s].len = len[s];
    prof[s].rows = (int32_t *)malloc(sizeof(int32_t));
    prof[s].cell = (int32_t *)malloc(sizeof(int32_t) * (size_t)(len[s] ? len[s] : 1));
    if (!prof[s].rows || !prof[s].cell) { rc = -5; goto out; }
    prof[s].rows[0] = s;
    for (int32_t k = 0; k < len[s]; ++k) prof[s].cell[k] = codes[off[s] + k];
  }
  /* 3-4: joins */
  for (int32_t t = 0; t < nseq - 1; ++t) {
    int32_t ba = -1, bb = -1;
    for (int32_t a = 0; a < nseq + t; ++a) {
      if (!live[a]) continue;
      for (int32_t b = a + 1; b < nseq + t; ++b) {
        if (!live[b]) continue;
        if (ba < 0) { ba = a; bb = b; continue; }
        /* mean(a,b) > mean(ba,bb)  <=>  sum(a,b) * |ba||bb| > sum(ba,bb) * |a||b| */
        const __int128 lhs = (__int128)sum[(size_t)a * nc + b] * size[ba] * size[bb];
        const __int128 rhs = (__int128)sum[(size_t)ba * nc + bb] * size[a] * size[b];
        if (lhs > rhs) { ba = a; bb = b; }
      }
    }
    const int32_t c = nseq + t;
    if (joins) { joins[2 * t] = ba; joins[2 * t + 1] = bb; }
    const profile_t *A = &prof[ba], *B = &prof[bb];
    char *ops = (char *)malloc((size_t)(A->len + B->len + 1));
    if (!ops) { rc = -5; goto out; }
    int32_t nops = 0;
    int64_t score = 0;
    const int64_t w = (int64_t)A->nrows * B->nrows;
    rc = align_profiles(smat, K, gap_open * w, gap_extend * w, A, B, ops, &nops, &score);
    if (rc) { free(ops); goto out; }
    profile_t C;
    C.nrows = A->nrows + B->nrows;
    C.len = nops;
    C.rows = (int32_t *)malloc(sizeof(int32_t) * (size_t)C.nrows);
    C.cell = (int32_t *)malloc(sizeof(int32_t) * (size_t)C.nrows * (size_t)(nops ? nops : 1));
    if (!C.rows || !C.cell) { free(ops); free(C.rows); free(C.cell); rc = -5; goto out; }
    for (int32_t r = 0; r < A->nrows; ++r) C.rows[r] = A->rows[r];
    for (int32_t r = 0; r < B->nrows; ++r) C.rows[A->nrows + r] = B->rows[r];
    int32_t ia = 0, jb = 0;
    for (int32_t q = 0; q < nops; ++q) {
      const char o = ops[q];
      for (int32_t r = 0; r < A->nrows; ++r)
        C.cell[(size_t)r * nops + q] = o == 'D' ? -1 : A->cell[(size_t)r * A->len + ia];
      for (int32_t r = 0; r < B->nrows; ++r)
        C.cell[(size_t)(A->nrows + r) * nops + q] = o == 'I' ? -1 : B->cell[(size_t)r * B->len + jb];
      if (o != 'D') ++ia;
      if (o != 'I') ++jb;
    }
    free(ops);
    prof_free(&prof[ba]);
    prof_free(&prof[bb]);
    prof[c] = C;
    size[c] = size[ba] + size[bb];
    live[ba] = live[bb] = 0;
    live[c] = 1;
    for (int32_t x = 0; x < c; ++x) {
      if (!live[x] || x == c) continue;
      const int64_t v = sum[(size_t)ba * nc + x] + sum[(size_t)bb * nc + x];
      sum[(size_t)c * nc + x] = sum[(size_t)x * nc + c] = v;
    }
  }
  /* 5: columns of every residue */
  {
    const profile_t *F = &prof[nc - 1];
    *alen = F->len;
    for (int32_t r = 0; r < F->nrows; ++r) {
      const int32_t s = F->rows[r];
      int32_t k = 0;
      for (int32_t q = 0; q < F->len; ++q)
        if (F->cell[(size_t)r * F->len + q] >= 0) col[off[s] + k++] = q;
    }
  }
out:
  if (prof)
    for (int32_t x = 0; x < nc; ++x) prof_free(&prof[x]);
  free(prof); free(sum); free(size); free(live); free(self);
  return rc;
}
