// Shared host/device definitions of the alignment engine.
#pragma once
#include <stdint.h>

namespace ta {

// A DP group is 16 lanes (one DPP row; a wave64 runs 4 groups), 8 lanes (two
// per row, interleaved; a wave runs 8: the smaller packed couples of a large
// batch, whose 64-row strips waste fewer pad rows) or the whole 64-lane wave
// (long tasks).  A group owns one task (a pair, or a couple of pairs packed
// into int16 halves) and sweeps it strip by strip.
constexpr int G = 16;
constexpr int GWAVE = 64;
constexpr int NGW = GWAVE / G;  // groups per wave
// Rows per lane held in registers (a strip is GG * R rows of the query).
constexpr int R = 8;
constexpr int STRIP = G * R;
// Ring index of column j is j + GG (keeps the last lane's block stores and
// lane 0's block loads in bounds and lane 0's loads 16-byte aligned).

constexpr int16_t NEG16 = -16384;          // -inf of the int16 lanes
constexpr int32_t NEG32 = INT32_MIN / 4;   // -inf of the int32 lanes
constexpr int8_t PAD_SCORE = -64;          // score of a pad row or pad column

// Task: the unit a DP group pulls off the work queue.
struct Task {
  int32_t pair[2];     // pair ids of the lo / hi half (hi == lo for a single)
  int32_t m[2];        // query rows of each half
  int32_t off[2];      // leading pad columns of each half (local couples)
  int32_t n;           // columns of the task
  int32_t S;           // strips = ceil(max m / (GG * R))
  int32_t NB;          // 8-step blocks per strip = ceil((n + GG) / 8)
  int32_t single;      // 1: hi half duplicates lo (results of hi discarded);
                       // 2: filler of a wave unit (computes, writes no results)
  uint64_t row_off;    // rowcodes element offset (S * STRIP uint16)
  uint64_t col_off;    // colcodes element offset (8 * NB + 8 uint16)
  uint64_t ring_off;   // ring element offset (2 * ring_len uint32)
  uint64_t tb_off;     // direction-plane byte offset (0 when not tracing)
  int32_t ring_len;    // elements per ring array
  int32_t flags_or;    // OR-ed into the FillOut flags of the task's pairs
};

// Per-pair output of the fill.
struct FillOut {
  int32_t score;
  int32_t end_i, end_j;  // 1-based end cell (global: m, n); 0 when empty
  int32_t flags;
};

// Where a pair's direction planes live.
struct PairTB {
  uint64_t tb_off;
  int32_t P;      // 2: packed couple layout, 1: int32 layout, 0: none
  int32_t half;   // 0 lo, 1 hi
  int32_t NB;
  int32_t off;    // leading pad columns
  int32_t GG;     // lanes per group of the task (16 or 64)
  int32_t pad_;
};

// Pair as the runtime keeps it on the device.
struct PairDev {
  uint64_t a_off, b_off;  // into the residue arena (codes)
  int32_t la, lb;
  uint64_t cig_off;       // byte offset of this pair's CIGAR scratch
  int32_t cig_cap;
  int32_t pad_;
};

// Final per-pair result (same layout as ta_result).
struct ResultDev {
  int32_t score, a_begin, a_end, b_begin, b_end;
  int64_t cigar_off;
  int32_t cigar_len, flags;
};

// Bytes of direction planes per strip of a task.
inline uint64_t tb_strip_bytes(int P, int NB, int GG) {
  return (uint64_t)NB * GG * 8 * (P == 2 ? 8 : 4);
}

}  // namespace ta
