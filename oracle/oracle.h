/*
 * CPU oracle for the pairwise-alignment hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / the timed CPU baseline.  The
 * product path (libtweakalign.so) never calls into it.
 *
 * What it restates.  The reference (groundstate/tweakseq) delegates alignment
 * to external programs (tweakseq/Core/AlignmentTool.h:36-71,
 * tweakseq/UI/SeqEditMainWin.cpp:1616-1663); it holds no DP code.  The only
 * numeric data it holds for this path is the BLOSUM62 table and letter map
 * (tweakseq/Core/Annotations/Consensus.cpp:34-69), which callers pass in here
 * (the tests read them from tests/golden/blosum62.json, extracted from that
 * file as text).  The recurrence, boundaries, tie-breaks and traceback follow
 * the written specification kept with this project (Gotoh 1982, Durbin et al.
 * 1998 affine convention); ties are first-wins in the order
 * diag > E > F, the idiom of find_largest_int (tweakseq/Core/Utility.cpp:44-50).
 *
 * Pinning: tests/test_oracle.py checks every function here against brute-force
 * enumeration of all alignments (tiny inputs), closed forms, symmetry, CIGAR
 * re-scoring, and the reference table fixture.
 */
#ifndef TWEAKALIGN_ORACLE_H
#define TWEAKALIGN_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Same field layout as ta_params / ta_result in include/tweakalign.h. */
typedef struct {
  int32_t mode;       /* 0 global (Needleman-Wunsch), 1 local (Smith-Waterman) */
  int32_t alphabet;   /* 0 protein, 1 DNA */
  int32_t gap_open;   /* score of a gap of length 1 (<= 0) */
  int32_t gap_extend; /* added per further gap position (<= 0) */
  int32_t match, mismatch; /* DNA only */
  int32_t traceback;  /* 0 score + coordinates, 1 also CIGAR */
  int32_t lanes;      /* ignored by the oracle */
} or_params;

typedef struct {
  int32_t score, a_begin, a_end, b_begin, b_end;
  int64_t cigar_off;
  int32_t cigar_len, flags;
} or_result;

/* Letter -> code.  Protein: uppercase, then letter_map['A'..'Z'] (the
 * reference's BLOSUM62map, Consensus.cpp:61-69); DNA: A,C,G,T -> 0..3.
 * Returns 0, or -2 when a character is outside the alphabet. */
int or_encode(const int32_t *letter_map26, int32_t alphabet, const char *s,
              int32_t len, uint8_t *out);

/* Align codes a[0..m) (rows, query) against b[0..n) (columns, reference)
 * with a K x K substitution matrix smat (row-major, int8).  Writes the
 * result; when p->traceback and cigar != NULL writes the run-length CIGAR
 * (M/I/D, no terminator) into cigar[0..cap).  Returns 0, -3 if cap is too
 * small (res->cigar_len then holds the size needed), -1 on bad arguments,
 * -5 on allocation failure. */
int or_align(const or_params *p, const int8_t *smat, int32_t K,
             const uint8_t *a, int32_t m, const uint8_t *b, int32_t n,
             or_result *res, char *cigar, int64_t cap);

/* Many pairs, OpenMP over nthreads (<= 0: all).  codes holds all residues;
 * pair k is a = codes + a_off[k] (la[k]), b = codes + b_off[k] (lb[k]).
 * CIGAR of pair k goes to cigar + cig_off[k] with room cig_cap[k] (cigar may
 * be NULL).  Returns 0 or the first negative code met. */
int or_align_batch(const or_params *p, const int8_t *smat, int32_t K,
                   const uint8_t *codes, const int64_t *a_off, const int32_t *la,
                   const int64_t *b_off, const int32_t *lb, int32_t npairs,
                   or_result *res, char *cigar, const int64_t *cig_off,
                   const int64_t *cig_cap, int32_t nthreads);

/* Score a given CIGAR under the same cost model (affine runs as in the
 * recurrence).  Returns 0 and *score, or -1 if the CIGAR does not consume
 * exactly a[a_begin..a_end] and b[b_begin..b_end]. */
int or_rescore(const or_params *p, const int8_t *smat, int32_t K,
               const uint8_t *a, const uint8_t *b, const or_result *res,
               const char *cigar, int32_t *score);

/* Median wall time (seconds) of `repeats` or_align calls on one pair, timed
 * in C (timing.c): the single-core CPU baseline without call overhead. */
int or_time_align(const or_params *p, const int8_t *smat, int32_t K, const uint8_t *a, int32_t m,
                  const uint8_t *b, int32_t n, int32_t repeats, char *cigar, int64_t cap, double *median_s);

/* Column consensus of an alignment of nseq rows of len characters (msa is
 * row-major, row s at msa + s * len), restating Consensus::calculate
 * (tweakseq/Core/Annotations/Consensus.cpp:80-161) with the 23 x 23 BLOSUM62
 * table and 26-letter map passed in; plurality < 0 means nseq / 2.  Writes
 * out[0..len).  Returns 0, -1 on bad arguments, -5 on allocation failure. */
int or_consensus(const int8_t *smat23, const int32_t *map26, const char *msa, int32_t nseq, int32_t len,
                 double plurality, char *out);
/* Progressive multiple alignment (msa.c states the specification): global
 * pair scores, UPGMA on normalised similarities, profile joins by global
 * Gotoh over columns.  Sequence s is codes[off[s] .. off[s] + len[s]).
 * Writes the alignment length to *alen, the column of every residue to
 * col[off[s] + k], and (joins != NULL) the two clusters of join t to
 * joins[2t], joins[2t + 1] (leaves 0..nseq-1, join t creates nseq + t).
 * Returns 0, -1 on bad arguments, -2 for a code >= K, -5 on allocation
 * failure. */
int or_msa(const int8_t *smat, int32_t K, int32_t gap_open, int32_t gap_extend, const uint8_t *codes,
           const int64_t *off, const int32_t *len, int32_t nseq, int32_t *joins, int32_t *alen,
           int32_t *col);

#ifdef __cplusplus
}
#endif
#endif
