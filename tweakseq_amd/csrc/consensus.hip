// Column consensus of an alignment on the GPU.
//
// Replaces Consensus::calculate (tweakseq/Core/Annotations/Consensus.cpp:80-161),
// which scores, for every column, each sequence's residue against every
// other's (BLOSUM62 for two letters, +1 for two non-letters, -4 for one of
// each) in O(nseq^2) per column, keeps the first sequence with the highest
// score, and takes its character when the number of its positive pairings
// reaches the plurality (else '?').
//
// A sequence's score depends only on the class of its residue (23 BLOSUM62
// rows or "not a letter"), so one thread per column counts the classes in
// one pass over the rows (coalesced: consecutive threads read consecutive
// columns of a row) and scores each class present against the counts,
// O(nseq + 24^2) per column: score(k) = sum_j cnt[j] * S(k, j) - S(k, k),
// the self pairing removed.  The winner is the class of best score whose
// first sequence comes first -- the reference's strict ">" over sequences in
// order.  Scores are integers, as the reference's double sums of integer
// weights are.  HBM-bound: nseq * len bytes read once.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <cstring>
#include <vector>

#include "runtime_internal.h"

namespace {

constexpr int CK = 24;       // 23 BLOSUM62 rows, then "not a letter"
constexpr int CT = 256;      // threads (columns) per workgroup

struct ConsArgs {
  const uint8_t *msa;        // row s at msa + s * len
  int32_t nseq, len;
  double plurality;
  char *out;
  const int8_t *cls;         // byte -> class (256)
  const int8_t *score;       // CK x CK pairing scores
  const int8_t *match;       // CK x CK: 1 when a pairing counts as a match
};

__global__ __launch_bounds__(CT) void consensus_kernel(ConsArgs A) {
  __shared__ int8_t cls[256];
  __shared__ int32_t sc[CK * CK], mt[CK * CK];
  __shared__ int32_t cnt[CK][CT], first[CK][CT];
  const int tid = threadIdx.x;
  cls[tid] = A.cls[tid];
  for (int x = tid; x < CK * CK; x += CT) {
    sc[x] = A.score[x];
    mt[x] = A.match[x];
  }
  __syncthreads();
  const int c = blockIdx.x * CT + tid;
  if (c >= A.len) return;
#pragma unroll
  for (int k = 0; k < CK; ++k) {
    cnt[k][tid] = 0;
    first[k][tid] = INT_MAX;
  }
  const uint8_t *col = A.msa + c;
  for (int s = 0; s < A.nseq; ++s) {
    const int k = cls[col[(size_t)s * A.len]];
    if (cnt[k][tid]++ == 0) first[k][tid] = s;
  }
  int best = INT_MIN, best_first = INT_MAX, best_match = 0;
  for (int k = 0; k < CK; ++k) {
    const int ck = cnt[k][tid];
    if (ck == 0) continue;
    int s = -sc[k * CK + k], m = -mt[k * CK + k];  // no pairing with itself
#pragma unroll
    for (int j = 0; j < CK; ++j) {
      const int cj = cnt[j][tid];
      s += cj * sc[k * CK + j];
      m += cj * mt[k * CK + j];
    }
    const int f = first[k][tid];
    if (s > best || (s == best && f < best_first)) {
      best = s;
      best_first = f;
      best_match = m;
    }
  }
  A.out[c] = (double)best_match >= A.plurality ? (char)col[(size_t)best_first * A.len] : '?';
}

// Class of every byte, pairing scores and match flags (from the engine's own
// BLOSUM62 table and letter map, ta_matrix / ta_encode).
struct ConsTables {
  int8_t cls[256], score[CK * CK], match[CK * CK];
};

ConsTables make_tables() {
  ConsTables t;
  int8_t S[32 * 32];
  int32_t K = 0;
  ta_matrix(TA_PROTEIN, 0, 0, S, &K);
  for (int b = 0; b < 256; ++b) {
    t.cls[b] = CK - 1;
    if (b >= 'A' && b <= 'Z') {  // the reference maps only 'A'..'Z' (latin1 - 65)
      const char ch = (char)b;
      uint8_t code = 0;
      if (ta_encode(TA_PROTEIN, &ch, 1, &code) == TA_OK) t.cls[b] = (int8_t)code;
    }
  }
  for (int a = 0; a < CK; ++a)
    for (int b = 0; b < CK; ++b) {
      int s, m;
      if (a == CK - 1 && b == CK - 1) { s = 1; m = 1; }
      else if (a == CK - 1 || b == CK - 1) { s = -4; m = 0; }
      else { s = S[a * K + b]; m = s > 0; }
      t.score[a * CK + b] = (int8_t)s;
      t.match[a * CK + b] = (int8_t)m;
    }
  return t;
}

}  // namespace

int ta_consensus(ta_ctx *ctx, const char *const *rows, int32_t nseq, int32_t len, double plurality, char *out) {
  if (!ctx || !rows || nseq <= 0 || len < 0 || (len > 0 && !out)) return TA_E_INVALID;
  for (int32_t s = 0; s < nseq; ++s)
    if (!rows[s] && len > 0) return TA_E_INVALID;
  if (len == 0) return TA_OK;
  if (plurality < 0) plurality = nseq / 2.0;  // Consensus::setSequences, weights 1.0
  if (hipSetDevice(ctx->device) != hipSuccess) return TA_E_DEVICE;
  static const ConsTables tables = make_tables();
  const size_t bytes = (size_t)nseq * (size_t)len;
  std::vector<uint8_t> host(bytes);
  for (int32_t s = 0; s < nseq; ++s) memcpy(host.data() + (size_t)s * len, rows[s], (size_t)len);
  uint8_t *d_msa = (uint8_t *)device_alloc(bytes);
  char *d_out = (char *)device_alloc((size_t)len);
  ConsTables *d_tab = (ConsTables *)device_alloc(sizeof(ConsTables));
  int rc = (d_msa && d_out && d_tab) ? TA_OK : TA_E_NOMEM;
  hipStream_t st = ctx->stream;
  if (!rc && (hipMemcpyAsync(d_msa, host.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemcpyAsync(d_tab, &tables, sizeof(ConsTables), hipMemcpyHostToDevice, st) != hipSuccess))
    rc = TA_E_DEVICE;
  if (!rc) {
    ConsArgs a;
    a.msa = d_msa; a.nseq = nseq; a.len = len; a.plurality = plurality; a.out = d_out;
    const int8_t *tb = (const int8_t *)d_tab;
    a.cls = tb + offsetof(ConsTables, cls);
    a.score = tb + offsetof(ConsTables, score);
    a.match = tb + offsetof(ConsTables, match);
    (void)hipGetLastError();
    hipLaunchKernelGGL(consensus_kernel, dim3((len + CT - 1) / CT), dim3(CT), 0, st, a);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, d_out, (size_t)len, hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = TA_E_DEVICE;
  }
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = TA_E_DEVICE;
  device_free(d_msa);
  device_free(d_out);
  device_free(d_tab);
  return rc;
}
