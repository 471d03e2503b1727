// Types of the device-queue serving kernel (devq.hip), shared by its device
// and host halves.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "fill.h"

namespace ta {

constexpr int SQ_MAXLEN = 8192;                                   // longest side of a queued pair
constexpr int SQ_MAXS = (SQ_MAXLEN + GWAVE * R - 1) / (GWAVE * R);  // strips of the tallest task
constexpr uint32_t SQ_CAP = 1u << 16;                             // item slots in the ring
constexpr int SQ_FREE_CUS = 8;                                    // CUs left to other streams

// One queued whole-wave task.  row_off / col_off index the host code arena
// (uint16 units); the claiming wave copies those codes into its scratch and
// rewrites the offsets (and ring_off) for its own scratch.  tb_off is in the
// device plane arena.
struct SQItem {
  Task task;
  int32_t kind;    // bit 0: local; bit 1: int32 lanes (P = 1)
  int32_t pad_[3];
};

struct SQArgs {
  FillArgs fa[2];                 // [0] global, [1] local (the walks' TbArgs inside)
  const SQItem *items;            // ring in host memory
  const uint32_t *tail;           // host: items published
  uint32_t *head;                 // device: items claimed
  unsigned long long *mirror;     // device: (tail << 32) | time of the last host read
  const uint16_t *hcodes;         // host code arena
  uint8_t *dcodes;                // device copy of its residue codes (same offsets), for the walkers
  const PairDev *hpairs;          // host pair records ...
  const PairTB *hptb;
  PairDev *dpairs;                // ... and their device copies (the walkers read these)
  PairTB *dptb;
  const uint32_t *stop;           // host: nonzero asks the kernel to leave (new gap scores)
  uint32_t *halt;                 // device: the stop request as seen by the waves
  uint32_t *flag;                 // host: per pair record, set (plain store) once its result is out
  int32_t idle_ticks;             // 100 MHz ticks without an item before a wave leaves
};

int devq_lds_bytes(int Kp);
size_t devq_scratch_rows();       // uint16 row codes per workgroup
size_t devq_scratch_cols();       // uint16 column codes per workgroup
size_t devq_scratch_ring();       // uint32 ring entries per workgroup
int devq_launch(const SQArgs *S, int grid, int lds, hipStream_t st);  // S: device memory

}  // namespace ta
