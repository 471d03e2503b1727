/*
 * Seeded synthetic workloads for the alignment engine's benchmarks and tests.
 * No counterpart in the reference (it ships no benchmarks or test data).
 *
 * Generator: xoshiro256** seeded through splitmix64; residues drawn by
 * inverse CDF over BLOSUM62 background frequencies (protein) or uniformly
 * (DNA); mutation = per-position substitution (new residue != old) and indel
 * events (insertion / deletion 50:50, geometric length).  Pairs are drawn in
 * sequence, so a workload built with a smaller size is a prefix of the full
 * one.
 *
 * Configurations (size = 0 -> the full size):
 *  1  one protein pair 300 x 320, global -10/-1, CIGAR
 *  2  all-vs-all of `size` (512) proteins, L ~ U[100,600], 30% derived from a
 *     parent (20% substitutions, 5% indels, mean 3); global -10/-1, score only
 *  3  `size` (65,536) pairs, la = clip(round(exp(N(ln 350, 0.6))), 30, 2000),
 *     b = mutate(a; 25%, 4%, mean 3); local -11/-1, CIGAR
 *  4  `size` (120,000) requests, Poisson 2,000/s, pairs per request
 *     Geometric(mean 16) truncated at 256, a and b independent U[50,5000],
 *     global / local 50:50 per request, -10/-1, CIGAR
 *  5  `size` (256) DNA pairs, L ~ U[10k,100k] (max_len caps it),
 *     b = mutate(a; 10%, 2%, mean 5); global +2/-3/-5/-2, CIGAR
 */
#ifndef TWEAKALIGN_WORKLOAD_H
#define TWEAKALIGN_WORKLOAD_H
#include <stdint.h>

#include "tweakalign.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ta_workload ta_workload;

int ta_workload_create(int32_t config, uint64_t seed, int32_t size, int32_t max_len,
                       ta_workload **out);
void ta_workload_destroy(ta_workload *w);
int32_t ta_workload_npairs(const ta_workload *w);
/* Pointers into storage owned by the workload. */
int ta_workload_pairs(const ta_workload *w, ta_pair *out);
int ta_workload_params(const ta_workload *w, ta_params *out);
/* Requests (config 4; other configs report one request holding all pairs). */
int32_t ta_workload_nrequests(const ta_workload *w);
int ta_workload_request(const ta_workload *w, int32_t r, double *arrival_s, int32_t *first,
                        int32_t *count, int32_t *mode);

#ifdef __cplusplus
}
#endif
#endif
