/*
 * Column consensus of an alignment -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates Consensus::calculate, tweakseq/Core/Annotations/Consensus.cpp:80-161,
 * loop for loop: every sequence's residue at a column scores against every
 * other sequence's residue there (BLOSUM62 for two letters, +1 for two
 * non-letters, -4 for one of each, all weights 1.0, in double as the
 * reference accumulates), the first sequence with the highest score wins
 * (strict >, sequence 0 first), and its character is the consensus if the
 * number of positive pairings ("matches") reaches the plurality, else '?'.
 * A character is a letter when it is 'A'..'Z' (the reference maps
 * (latin1 - 65) through BLOSUM62map and anything else to a sentinel; here
 * anything else is the same sentinel); the consensus keeps the winner's
 * original byte.  The default plurality (setSequences, :163-173) is
 * nseq / 2.
 */
#include <stdlib.h>

#include "oracle.h"

#define NONLETTER 99

int or_consensus(const int8_t *smat23, const int32_t *map26, const char *msa, int32_t nseq, int32_t len,
                 double plurality, char *out) {
  if (!smat23 || !map26 || nseq <= 0 || len < 0 || (len > 0 && (!msa || !out))) return -1;
  if (plurality < 0) plurality = nseq / 2.0;
  int *idx = (int *)malloc(sizeof(int) * (size_t)nseq * (size_t)(len ? len : 1));
  double *scores = (double *)malloc(sizeof(double) * (size_t)nseq);
  if (!idx || !scores) {
    free(idx);
    free(scores);
    return -5;
  }
  for (int s = 0; s < nseq; s++)
    for (int c = 0; c < len; c++) {
      const int x = (unsigned char)msa[(size_t)s * len + c] - 65;
      idx[(size_t)s * len + c] = (x < 0 || x > 25) ? NONLETTER : map26[x];
    }
  for (int c = 0; c < len; c++) {
    double hi = 0.0, hi_matches = 0.0;
    int hi_ri = 0;
    for (int ri = 0; ri < nseq; ri++) {
      double matches = 0.0;
      scores[ri] = 0.0;
      for (int rj = 0; rj < nseq; rj++) {
        if (ri == rj) continue;
        const double weight = 1.0;
        const int resi = idx[(size_t)ri * len + c], resj = idx[(size_t)rj * len + c];
        if (resi == NONLETTER && resj == NONLETTER) {
          scores[ri] += weight;
          if (weight > 0) matches += weight;
        } else if (resi == NONLETTER || resj == NONLETTER) {
          scores[ri] += -4 * weight;
        } else {
          const double t = smat23[resi * 23 + resj] * weight;
          scores[ri] += t;
          if (t > 0) matches += weight;
        }
      }
      if (ri == 0 || scores[ri] > hi) {
        hi = scores[ri];
        hi_ri = ri;
        hi_matches = matches;
      }
    }
    out[c] = hi_matches >= plurality ? msa[(size_t)hi_ri * len + c] : '?';
  }
  free(idx);
  free(scores);
  return 0;
}
