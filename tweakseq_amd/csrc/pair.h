// The single-pair (low-latency) kernel's interface to its host code.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace ta {

constexpr int PAIR_WAVES = 4;                // waves of the workgroup (one per SIMD)
constexpr int PAIR_R = 2;                    // query rows per lane
constexpr int PAIR_STRIP = 64 * PAIR_R;      // rows per strip
constexpr int PAIR_PAD = 31;                 // code of a row / column outside the sequences
constexpr int PAIR_BAD_LETTER = 1;           // PairOut::status bits
constexpr int PAIR_GAVE_UP = 2;

// LDS byte offsets of the kernel's arrays for an m x n pair.
struct PairLayout {
  int32_t T, S;             // steps per strip (n + 63, rounded to 8), strips
  int32_t ring_len;         // int2 entries per ring slot
  int32_t runs_cap;         // CIGAR run slots
  int32_t acode, bcode, sync, ring, planes, runs;
  int32_t bytes;            // total LDS
};

// Written by the kernel: the result, then nruns CIGAR runs ((len << 2) | op,
// op 0 M / 1 D / 2 I) in alignment order right after this record.
struct PairOut {
  ResultDev r;
  int32_t nruns;
  int32_t status;           // PAIR_BAD_LETTER, PAIR_GAVE_UP
  int32_t pad_[2];
};

struct PairArgs {
  const char *a, *b;        // letters, device-readable (mapped host staging or device scratch)
  int32_t m, n;
  int32_t alphabet, match, mismatch;
  int32_t open, ext, local, traceback;
  int32_t spin_limit;
  int32_t walk_cap;         // most steps the walk may take before it ends flagged (above any walk normally)
  PairLayout lay;
  PairOut *out;             // device-writable; m + n + 2 run slots follow it
};

PairLayout pair_layout(int m, int n);
// One workgroup on stream st; TA_E_DEVICE when the launch fails.
int pair_launch(const PairArgs &A, hipStream_t st);

}  // namespace ta
