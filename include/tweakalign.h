/*
 * tweakalign -- MI355X (gfx950) pairwise alignment engine, plain C ABI.
 *
 * Replaces the external aligner process behind tweakseq's AlignmentTool
 * boundary.  In the reference an alignment is: export FASTA
 * (tweakseq/Core/Project.cpp:870-894), AlignmentTool::makeCommand
 * (tweakseq/Core/AlignmentTool.h:54, Muscle.cpp:49-53), QProcess::start
 * (tweakseq/UI/SeqEditMainWin.cpp:1660), success test on exit code + output
 * file (SeqEditMainWin.cpp:836-861), ingest (Project.cpp:908-1032).  The
 * calls below are what a host binds instead of spawning that process; the
 * tweakalign CLI (tweakseq_amd/csrc/cli.cpp) wraps them behind the same
 * "-in FILE -out FILE" contract for an unmodified host.
 *
 * Conventions: negative TA_E_* return codes, never abort, no C++ types or
 * exceptions across the boundary.  Inputs are caller-owned and only read
 * during the call (synchronous calls) or until ta_poll reports completion.
 * A ta_ctx is used by one host thread at a time; several may share a GPU.
 */
#ifndef TWEAKALIGN_H
#define TWEAKALIGN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define TA_OK 0
#define TA_E_INVALID -1       /* bad argument */
#define TA_E_ALPHABET -2      /* residue outside the alphabet */
#define TA_E_SPACE -3         /* caller buffer too small; *used = bytes needed */
#define TA_E_DEVICE -4        /* HIP runtime error */
#define TA_E_NOMEM -5         /* allocation failed */
#define TA_E_BUDGET -6        /* request exceeds the traceback-memory budget */
#define TA_E_CANCELLED -7     /* request was cancelled */
#define TA_E_TICKET -8        /* unknown ticket */
#define TA_E_NODEVICE -9      /* no usable GPU */
#define TA_E_WALK -10         /* a traceback walk did not complete: the pairs concerned carry
                                 TA_F_WALK_FAILED; their coordinates and CIGAR are not valid */
#define TA_E_LIMIT -11        /* input beyond the fixed size of ta_align_pair_static's scratch */
#define TA_E_PENDING 1        /* ta_poll: not finished yet (not an error) */

#define TA_GLOBAL 0           /* Needleman-Wunsch (Gotoh affine gaps) */
#define TA_LOCAL 1            /* Smith-Waterman (Gotoh affine gaps) */
#define TA_PROTEIN 0          /* BLOSUM62, tweakseq/Core/Annotations/Consensus.cpp:34-69 */
#define TA_DNA 1              /* A,C,G,T; match / mismatch scores */

/* ta_params.lanes */
#define TA_LANES_AUTO 0       /* packed int16 where safe, int32 otherwise; global pairs
                                 of >= 2^25 cells take the long-pair path */
#define TA_LANES_I16 16       /* force packed int16 (saturating pairs rerun in int32) */
#define TA_LANES_I32 32       /* force int32 */
#define TA_LANES_LONG 64      /* global only: force the long-pair path (int32 strips
                                 pipelined across CUs, checkpointed traceback) */

/* ta_result.flags */
#define TA_F_I32 0x1          /* result computed by the int32 kernels */
#define TA_F_RERUN 0x2        /* int16 pass saturated; rerun in int32 */
#define TA_F_LONG 0x4         /* long-pair (multi-strip pipeline) path */
#define TA_F_EMPTY 0x8        /* local: no cell scored above zero */
#define TA_F_WALK_FAILED 0x100 /* the traceback walk gave up (step budget spent, CIGAR region
                                  overflowed or the walk left the planes): begin coordinates and
                                  CIGAR are invalid; the call that returns it fails with TA_E_WALK */

typedef struct ta_ctx ta_ctx;

typedef struct {
  int32_t mode;       /* TA_GLOBAL | TA_LOCAL */
  int32_t alphabet;   /* TA_PROTEIN | TA_DNA */
  int32_t gap_open;   /* score of a 1-long gap, <= 0 (gap of k = open + (k-1)*extend) */
  int32_t gap_extend; /* <= 0 */
  int32_t match, mismatch; /* TA_DNA only */
  int32_t traceback;  /* 0: score + coordinates; 1: also CIGAR */
  int32_t lanes;      /* TA_LANES_* */
} ta_params;

typedef struct {
  const char *a; int32_t la;   /* query (rows, CIGAR 'I' consumes it), ASCII letters */
  const char *b; int32_t lb;   /* reference (columns, CIGAR 'D' consumes it) */
} ta_pair;

typedef struct {
  int32_t score;
  int32_t a_begin, a_end;      /* 0-based inclusive; -1 when empty (local) */
  int32_t b_begin, b_end;
  int64_t cigar_off;           /* offset of this pair's CIGAR in the caller buffer */
  int32_t cigar_len;           /* bytes, no terminator */
  int32_t flags;               /* TA_F_* */
} ta_result;

/* Message of a TA_E_* code (the reference routes a failed aligner's stderr to
 * its message window, tweakseq/UI/SeqEditMainWin.cpp:829-834). */
const char *ta_strerror(int code);

/* One context per device per host thread.  Replaces the per-project tool
 * objects (tweakseq/Core/Project.cpp:1142-1152, deleted at :84-86). */
int ta_create(ta_ctx **out, int device);
void ta_destroy(ta_ctx *ctx);

/* Synchronous batch: encode, upload, align on the GPU, return results and
 * CIGARs (packed back to back in cigar_buf).  Replaces one aligner run:
 * makeCommand + QProcess start + the success test on exit status
 * (tweakseq/Core/AlignmentTool.h:54, tweakseq/UI/SeqEditMainWin.cpp:1616-1663,
 * :836-861).  TA_E_SPACE with *used set when
 * cap is too small (results are still written).  cigar_buf may be NULL when
 * p->traceback == 0.  TA_E_WALK when any pair's traceback walk gave up
 * (results and CIGARs are still written; those pairs carry TA_F_WALK_FAILED):
 * the counterpart of an aligner that exits non-zero, which the reference
 * reports as "Alignment not completed" (SeqEditMainWin.cpp:852-856). */
int ta_align_batch(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n,
                   ta_result *out, char *cigar_buf, int64_t cap, int64_t *used);

/* One pair at low latency: the editor's "realign selection" of two rows
 * (tweakseq/UI/SeqEditMainWin.cpp:1616-1663 runs an external aligner on the
 * selection; tweakseq/Core/Project.cpp:964-1026 splices its output back).
 * One kernel launch with the letters read from mapped host staging, the
 * direction codes kept on chip, and the result written back to staging; no
 * device allocation.  Same result and conventions as ta_align_batch with
 * n = 1 (which pairs too large for the on-chip form, or with an empty side,
 * go through).  Results carry TA_F_I32 (the pair kernel computes in int32). */
int ta_align_pair(ta_ctx *ctx, const ta_params *p, const char *a, int32_t la, const char *b, int32_t lb,
                  ta_result *out, char *cigar_buf, int64_t cap, int64_t *used);

/* ta_align_pair without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The letters go to, and the result
 * comes back from, a fixed device array of TA_PAIR_STATIC_SCRATCH bytes that
 * each GPU holds until the process ends (set up by ta_create); host staging
 * is the calling thread's stack.  Limits: la + lb <= TA_PAIR_STATIC_MAX and
 * the pair's direction codes must fit on chip (the form ta_align_pair uses
 * for such pairs); larger pairs get TA_E_LIMIT, never a fallback.  Calls
 * that overlap on one GPU (several contexts or threads) race on its array. */
#define TA_PAIR_STATIC_MAX 4096
#define TA_PAIR_STATIC_SCRATCH 32768
int ta_align_pair_static(ta_ctx *ctx, const ta_params *p, const char *a, int32_t la, const char *b, int32_t lb,
                         ta_result *out, char *cigar_buf, int64_t cap, int64_t *used);

/* Device-resident batch: inputs encoded and uploaded once, the alignment
 * run as often as wanted with inputs already in HBM.  No counterpart in the
 * reference (it starts one process per job). */
typedef struct ta_batch ta_batch;
int ta_batch_create(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n,
                    ta_batch **out);
int ta_batch_run(ta_ctx *ctx, ta_batch *b);         /* enqueue on the context stream */
int ta_batch_run_plain(ta_ctx *ctx, ta_batch *b, int32_t first, int32_t count);
                                                     /* one launch per pair (baseline) */
int ta_batch_wait(ta_ctx *ctx, ta_batch *b);
int ta_batch_results(ta_ctx *ctx, ta_batch *b, ta_result *out, char *cigar_buf,
                     int64_t cap, int64_t *used);   /* TA_E_SPACE / TA_E_WALK as ta_align_batch */
/* Replay the batch's runs as one captured hipGraph (on != 0): the first run
 * captures init, fill and traceback kernels, later runs launch the graph.
 * Batches that may need the int32 rerun keep launching eagerly.  No
 * counterpart in the reference. */
int ta_batch_graph(ta_batch *b, int32_t on);
/* Kernel timings of the last ta_batch_run, milliseconds (HIP events on the
 * batch stream): [0] fill kernel(s), [1] int32 rerun, [2] traceback,
 * [3] whole run; [0]-[2] are -1 after a captured (graph) run.  Returns the
 * number of values written. */
int ta_batch_timings(ta_batch *b, double *ms, int32_t cap);
/* Cells (sum of la*lb) and direction-plane bytes of the batch. */
int ta_batch_info(ta_batch *b, int64_t *cells, int64_t *tb_bytes, int32_t *tasks);
/* Waits for the batch's last run, then frees it. */
void ta_batch_destroy(ta_batch *b);

/* Resident sequences: each encoded once into device memory of the
 * context's GPU and aligned by any number of later calls without being sent
 * again (the rows an editor realigns over and over, the sequences of a
 * guide tree's all-pairs scores).  The reference writes every sequence to
 * the external aligner's input file for each run
 * (tweakseq/UI/SeqEditMainWin.cpp:1616-1663).  A handle stays valid, for the
 * alphabet it was added with, until ta_seq_reset; the store grows as needed,
 * and ta_seq_reset and ta_destroy free it once no batch or served request
 * still reads it. */
typedef int64_t ta_seq;
typedef struct {
  ta_seq a, b;
} ta_seq_pair;
int ta_seq_add(ta_ctx *ctx, int32_t alphabet, const char *letters, int32_t len, ta_seq *out);
                                      /* TA_E_ALPHABET: a letter outside the alphabet;
                                         TA_E_INVALID: len outside [0, 1,000,000] */
int ta_seq_reset(ta_ctx *ctx);        /* forgets every resident sequence */
/* ta_align_batch on resident sequences: same results and return codes;
 * TA_E_INVALID for an unknown handle or one of another alphabet than
 * p->alphabet. */
int ta_align_resident(ta_ctx *ctx, const ta_params *p, const ta_seq_pair *pairs, int32_t n,
                      ta_result *out, char *cigar_buf, int64_t cap, int64_t *used);

/* Serving: continuous batching.  ta_submit copies the pairs and returns at
 * once; requests that arrive while others run are admitted as traceback
 * memory (budget set by ta_serve_config) frees up, and run concurrently.
 * ta_submit / ta_poll replace the asynchronous QProcess start and its
 * finished signal (tweakseq/UI/SeqEditMainWin.cpp:1660, :836-861), which
 * allow one job per window (:1618-1624); ta_cancel replaces the kill of
 * alignmentStop (:803-812).  ta_serve_config has no counterpart. */
int ta_serve_config(ta_ctx *ctx, int64_t tb_budget_bytes, int32_t max_inflight);
int ta_submit(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n,
              uint64_t *ticket);
int ta_poll(ta_ctx *ctx, uint64_t ticket, ta_result *out, char *cigar_buf, int64_t cap,
            int64_t *used);                          /* TA_OK done, TA_E_PENDING, <0 error;
                                                        TA_E_WALK: delivered (results and CIGARs
                                                        written, ticket retired) but a walk of the
                                                        request gave up (TA_F_WALK_FAILED) */
int ta_cancel(ta_ctx *ctx, uint64_t ticket);
/* ta_submit on resident sequences (ta_seq_add): the handles are resolved
 * when the call returns, so the store may grow or be reset while the request
 * waits; results through ta_poll.  Such requests run in batches of their own
 * (no letters to copy); no counterpart in the reference. */
int ta_submit_resident(ta_ctx *ctx, const ta_params *p, const ta_seq_pair *pairs, int32_t n,
                       uint64_t *ticket);

/* Autotuning (no counterpart in the reference).  A tuning is the launch
 * shape the batch runtime uses for one kind of batch: mode, traceback,
 * alphabet and length bucket (ta_tune_bucket). */
typedef struct {
  int32_t waves_per_cu;  /* resident fill waves per CU (4 = one per SIMD); 0 = default */
  int32_t lanes;         /* TA_LANES_AUTO | TA_LANES_I16 | TA_LANES_I32 (an explicit
                            ta_params.lanes takes precedence) */
  int32_t wave_tasks;    /* 0: whole-wave groups by the cost rule; 1: for every task
                            taller than one 16-lane strip */
} ta_tuning;
/* Length bucket of a batch by its median max(la, lb): 0 (< 256), 1 (< 1024),
 * 2 (< 4096), 3. */
int ta_tune_bucket(const ta_pair *pairs, int32_t n);
/* Times candidate tunings on these pairs (best of `repeats` runs each),
 * stores the fastest in the context for this kind of batch and returns it;
 * ta_batch_create / ta_align_batch / serving on this context then use it. */
int ta_autotune(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n, int32_t repeats,
                ta_tuning *best, double *best_ms);
/* Set / read the context's tuning for (p's mode, traceback, alphabet, bucket)
 * directly (ta_tune_get: TA_E_INVALID when none is stored). */
int ta_tune_set(ta_ctx *ctx, const ta_params *p, int32_t bucket, const ta_tuning *t);
int ta_tune_get(ta_ctx *ctx, const ta_params *p, int32_t bucket, ta_tuning *t);

/* Substitution table in use for an alphabet (row-major K x K, int8): BLOSUM62
 * of tweakseq/Core/Annotations/Consensus.cpp:34-59 for TA_PROTEIN. */
int ta_matrix(int32_t alphabet, int32_t match, int32_t mismatch, int8_t *out, int32_t *K);
/* Letter -> code, as the engine encodes input (TA_E_ALPHABET on others): the
 * letter map of Consensus.cpp:61-69; gaps must be removed first (the
 * residues tweakseq exports keep them, tweakseq/Core/Sequence.cpp:57-69). */
int ta_encode(int32_t alphabet, const char *s, int32_t len, uint8_t *out);

/* Column consensus of an alignment on the GPU: replaces
 * Consensus::calculate (tweakseq/Core/Annotations/Consensus.cpp:80-161).
 * rows[0..nseq) are the aligned sequences as tweakseq holds them (gaps
 * kept, latin1), each at least len characters; out[c] gets the character of
 * the first sequence whose residue scores best against the column
 * (BLOSUM62 for two letters 'A'..'Z', +1 for two non-letters, -4 for one of
 * each, summed over the other sequences), or '?' when that residue's
 * positive pairings number fewer than `plurality` (< 0: nseq / 2, the
 * reference's default with unit weights, Consensus.cpp:163-173). */
int ta_consensus(ta_ctx *ctx, const char *const *rows, int32_t nseq, int32_t len, double plurality, char *out);

/* Progressive multiple alignment of nseq sequences (seqs[s][0..lens[s]),
 * lens[s] >= 1, letters as ta_encode takes them; p->mode TA_GLOBAL, gaps and
 * alphabet as for pairs): global pair scores on the GPU, UPGMA on normalised
 * scores, then the profile joins level by level on the GPU.  Replaces the
 * external MSA run behind "Align all" (tweakseq/UI/SeqEditMainWin.cpp
 * :1647-1659, result read by Project::readNewAlignment,
 * tweakseq/Core/Project.cpp:908-1032).  Writes the alignment length to
 * *alen and the column of residue k of sequence s to cols[coff_s + k]
 * (coff_s = lens[0] + ... + lens[s-1]); joins (may be NULL) gets the two
 * clusters of each of the nseq - 1 joins (leaves 0..nseq-1, join t makes
 * cluster nseq + t).  A join's nA * nB is limited to 8,000,000
 * (TA_E_INVALID beyond). */
int ta_msa(ta_ctx *ctx, const ta_params *p, const char *const *seqs, const int32_t *lens, int32_t nseq,
           int32_t *alen, int32_t *cols, int32_t *joins);

#ifdef __cplusplus
}
#endif
#endif
