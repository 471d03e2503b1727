// Serving through a device queue: admission into a running persistent kernel.
//
// The serving loop (serve.hip) replaces the reference's one-process-per-job
// lifecycle (tweakseq/UI/SeqEditMainWin.cpp:1616-1663, :836-861).  Requests
// it can take here never build a batch of their own: the host plans each
// admitted request into whole-wave fill tasks, writes them as items into a
// ring in pinned host memory and bumps a producer index there; a persistent
// kernel (devq_kernel) that is already running claims items as they appear,
// fills them and walks their tracebacks beside the fill, and writes each
// pair's result and CIGAR runs straight back to pinned host memory, then
// sets the pair's completion flag there.  No copy command, allocation or launch
// happens per request while the kernel runs.
//
//  * Communication is through coherent (fine-grained) pinned host memory
//    only: the kernel's loads of items, codes and pair records and its
//    stores of results bypass the GPU caches, so nothing written by the host
//    can be stale on the device and nothing written by the device waits in a
//    cache.  A workgroup copies its task's row and column codes into its
//    private device scratch before sweeping it (the sweep reads them every 8
//    steps), and the task's pair records and residues into device copies
//    its walkers read.  Direction planes live in a device arena the host hands out per
//    request (the traceback-memory admission budget).  Completion is a store of
//    a per-pair flag with system-scope release order after the result's and
//    runs' system-scope stores: no read-modify-write ever targets host memory
//    (PCIe atomics are not assumed).
//  * Claims: a device counter of claimed items; the producer index is read
//    from host memory by one wave at a time, at most every SQ_REFRESH ticks,
//    into a device mirror the others poll.  A wave claims only items below
//    the mirrored index, so an item is never claimed before it is whole.
//  * Lifetime: waves that find no item for idle_ticks exit; the host
//    relaunches the kernel when it has work and the last launch has ended.
//    Every wait in the kernel is bounded.  The grid leaves SQ_FREE_CUS CUs
//    free, so other streams' kernels (and runtime copy kernels) still run.
//  * Global and local requests share the running kernel (each item carries
//    its mode; each pair record carries it for the walker).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "../../include/tweakalign.h"
#include "common.h"
#include "devq.h"
#include "fill.h"
#include "runtime_internal.h"

namespace ta {

// ---------------------------------------------------------------------------
// Device side.
namespace {

constexpr int SQ_ROWS = SQ_MAXS * GWAVE * R;             // row codes of the tallest task
constexpr int SQ_NB = (SQ_MAXLEN + GWAVE + 7) / 8;        // 8-step blocks of the widest task
constexpr int SQ_COLS = 8 * SQ_NB + 8;                   // its column codes
constexpr int SQ_RING = 2 * (8 * SQ_NB + 2 * GWAVE + 16);  // its ring (H and F rows)
constexpr uint32_t SQ_REFRESH = 200;                      // ticks (100 MHz: 2 us) between host reads
constexpr int SQ_TASK_WORDS = (int)(sizeof(SQItem) / 4);

// The next item for this wave (uniform), or -1 after idle_ticks without one.
__device__ int64_t devq_claim(const SQArgs &S) {
  int64_t got = -1;
  if ((threadIdx.x & 63) == 0) {
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
      if (__hip_atomic_load(S.halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;  // the host wants a relaunch
      const uint32_t h = __hip_atomic_load(S.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long mv = __hip_atomic_load(S.mirror, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t tl = (uint32_t)(mv >> 32);
      if ((int32_t)(tl - h) > 0) {
        uint32_t expect = h;
        if (__hip_atomic_compare_exchange_strong(S.head, &expect, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
          got = h;
          break;
        }
        continue;
      }
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      if ((uint32_t)now - (uint32_t)mv > SQ_REFRESH) {
        // refresh the mirror from host memory; the timestamp CAS elects one reader
        unsigned long long expect = mv;
        const unsigned long long mine = (mv & 0xffffffff00000000ull) | (uint32_t)now;
        if (__hip_atomic_compare_exchange_strong(S.mirror, &expect, mine, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
          const uint32_t ht = __hip_atomic_load(S.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          if (__hip_atomic_load(S.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
            __hip_atomic_store(S.halt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(S.mirror, ((unsigned long long)ht << 32) | (uint32_t)now, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        continue;
      }
      if (now - t0 > (uint64_t)S.idle_ticks) break;
      __builtin_amdgcn_s_sleep(32);
    }
  }
  const int lo = __shfl((int)(uint32_t)got, 0, 64), hi = __shfl((int)((uint64_t)got >> 32), 0, 64);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Copy of `count` 16-byte words from host memory to device memory with
// system-coherent loads (never a stale cached line of host memory).
__device__ __forceinline__ uint4 load16_sys(const uint4 *p) {
  const uint32_t *w = (const uint32_t *)p;
  return make_uint4(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM),
                    __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM),
                    __hip_atomic_load(w + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM),
                    __hip_atomic_load(w + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
}
__device__ void stage_codes(const uint4 *src, uint4 *dst, int count) {
  const int lane = threadIdx.x & 63;
  for (int base = lane; base < count; base += 4 * 64) {
    // four unconditional (clamped) loads in flight, then the stores
    const int q1 = base + 64, q2 = base + 128, q3 = base + 192, last = count - 1;
    const uint4 v0 = load16_sys(src + base), v1 = load16_sys(src + (q1 < count ? q1 : last));
    const uint4 v2 = load16_sys(src + (q2 < count ? q2 : last)), v3 = load16_sys(src + (q3 < count ? q3 : last));
    dst[base] = v0;
    if (q1 < count) dst[q1] = v1;
    if (q2 < count) dst[q2] = v2;
    if (q3 < count) dst[q3] = v3;
  }
}

// A claimed task's pair records and residue codes, from host memory to the
// device copies its walker reads (the walker runs on this CU, so it sees
// these stores).
__device__ void stage_pairs(const SQArgs &S, const Task &T) {
  const int lane = threadIdx.x & 63;
  constexpr int PDW = (int)(sizeof(PairDev) / 4), TDW = (int)(sizeof(PairTB) / 4);
  const int h = lane >> 5, k = lane & 31;  // half h: lanes 32 h .. 32 h + PDW + TDW - 1
  const int np = T.single ? 1 : 2;
  const int p = T.pair[h];
  uint32_t v = 0;
  if (h < np && k < PDW + TDW) {
    const uint32_t *src = k < PDW ? (const uint32_t *)(S.hpairs + p) + k : (const uint32_t *)(S.hptb + p) + (k - PDW);
    uint32_t *dst = k < PDW ? (uint32_t *)(S.dpairs + p) + k : (uint32_t *)(S.dptb + p) + (k - PDW);
    v = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    *dst = v;
  }
  // residues: a at dwords 0-1 (offset) and 4 (length), b at 2-3 and 5
  for (int hh = 0; hh < np; ++hh) {
    const uint32_t aoff_lo = __shfl(v, 32 * hh + 0, 64), aoff_hi = __shfl(v, 32 * hh + 1, 64);
    const uint32_t boff_lo = __shfl(v, 32 * hh + 2, 64), boff_hi = __shfl(v, 32 * hh + 3, 64);
    const int la = (int)__shfl(v, 32 * hh + 4, 64), lb = (int)__shfl(v, 32 * hh + 5, 64);
    const uint64_t aoff = ((uint64_t)aoff_hi << 32) | aoff_lo, boff = ((uint64_t)boff_hi << 32) | boff_lo;
    const uint8_t *hc = (const uint8_t *)S.hcodes;
    stage_codes((const uint4 *)(hc + aoff), (uint4 *)(S.dcodes + aoff), (la + 15) / 16);
    stage_codes((const uint4 *)(hc + boff), (uint4 *)(S.dcodes + boff), (lb + 15) / 16);
  }
}

// Walker side of the serving kernel: as walk_loop, for pairs of either mode,
// and each finished pair's flag set (host memory) once its result and runs
// are stored.  A run walks pairs of one lane width PW (its own register
// allocation: the two walkers never hold state at once); when the oldest
// waiting pair has the other width, it drains its lanes and returns so the
// caller switches.  Returns (head << 1) | 1 for a switch, head << 1 when the
// fill side has finished and every lane is idle.
template <int PW>
__device__ __noinline__ int devq_walk_run(const SQArgs &S, WalkQ *q, const int8_t *tab, uint32_t *wlds, int head) {
  const int wl = threadIdx.x & 63;
  uint32_t *slice = wlds + wl * WSTRIDE;
  Walker<PW> w;
  bool busy = false;
  int pair = 0;
  for (;;) {
    const uint64_t all = __ballot(true);
    const uint64_t idle = __ballot(!busy);
    if (idle) {
      const int done = __hip_atomic_load(&q->done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
      const int tail = __hip_atomic_load(&q->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
      const int avail = tail - head;
      if (avail > 0) {
        const int nidle = __popcll(idle);
        int take = avail < nidle ? avail : nidle;
        const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0));
        const bool cand = !busy && rank < take;
        const int p = cand ? q->ids[(head + rank) & (WQ_CAP - 1)] : 0;
        const bool other = cand && ((S.fa[0].wk.ptb[p].P == 1 ? 1 : 2) != PW);
        const uint64_t oth = __ballot(other);
        if (oth) take = __shfl(rank, __builtin_ctzll(oth), 64);  // the leading run of this width only
        if (take == 0) {
          if (idle == all) return (head << 1) | 1;  // the oldest pair needs the other walker
        } else {
          if (!busy && rank < take) {
            const int local = (S.fa[0].wk.pairs[p].pad_ >> 30) & 1;
            const TbArgs &A = S.fa[local].wk;
            PairDev pd;
            FillOut f;
            ResultDev r;
            if (walk_prologue(A, p, pd, f, r)) {
              w.begin(A, p, pd, f, r);  // keeps the mode and gap scores of A
              pair = p;
              busy = true;
            } else {
              __hip_atomic_store(S.flag + p, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
          }
          head += take;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // ids read before the slots are freed
          if (wl == 0) __hip_atomic_store(&q->head, head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      } else if (done && idle == all) {
        return head << 1;
      }
    }
    if (__any(busy)) {
      if (busy && w.epoch(S.fa[0].wk, tab, slice)) {
        busy = false;
        // release at system scope: the result and runs (system-scope stores)
        // are visible to the host before the flag is
        __hip_atomic_store(S.flag + pair, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    } else {
      __builtin_amdgcn_s_sleep(64);
    }
  }
}

__device__ void devq_walk(const SQArgs &S, WalkQ *q, const int8_t *tab, uint32_t *wlds) {
  int head = 0, pw = 2;
  for (;;) {
    const int r = pw == 2 ? devq_walk_run<2>(S, q, tab, wlds, head) : devq_walk_run<1>(S, q, tab, wlds, head);
    head = r >> 1;
    if (!(r & 1)) return;
    pw = 3 - pw;
  }
}

// A cooperative sweep of one item (own register allocation per lane width
// and mode).
template <int P, bool LOCAL>
__device__ __noinline__ void devq_fill(const FillArgs &A, const Task *tk, int32_t *ctr, const int8_t *tab, int8_t *slot,
                                       WalkQ *to, CoopQ *cq) {
  int wtail = __hip_atomic_load(&to->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  fill_phase<GWAVE, P, LOCAL, true, true>(A, tk, 1, ctr, tab, slot, to, wtail, cq);
}

// Barrier of a workgroup's WALK_FW fill waves (the walker waves never join
// it): LDS arrival count bar[0] and generation bar[1].  A wait gives up after
// 10 s (the leader's claim returns within its 2 s idle bound), returning
// false; the caller then leaves.
__device__ bool fill_barrier(int32_t *bar) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  int ok = 1;
  if ((threadIdx.x & 63) == 0) {
    const int g = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (__hip_atomic_fetch_add(bar, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == WALK_FW - 1) {
      __hip_atomic_store(bar, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_store(bar + 1, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
      while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == g) {
        if (__builtin_amdgcn_s_memrealtime() - t0 > 1000000000ull) { ok = 0; break; }
        __builtin_amdgcn_s_sleep(2);
      }
    }
  }
  ok = __shfl(ok, 0, 64);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return ok != 0;
}

// One workgroup per CU: WALK_FW fill waves and WALK_FW walker waves (the
// walking fill kernel's layout).  The workgroup takes one item at a time:
// wave 0 claims it and stages its codes into the workgroup's scratch, then
// the fill waves sweep it together as a cooperative task (wave w takes
// strips w, w + WALK_FW, ...), so a long pair's fill time is its cells over
// four waves, not one; wave 0 hands the pairs to the walker waves in turn.
__global__ __launch_bounds__(128 * WALK_FW) void devq_kernel(const SQArgs *__restrict__ Sp) {
  extern __shared__ __align__(16) uint8_t lds[];
  const SQArgs &S = *Sp;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int8_t *tab = (int8_t *)lds;
  const int Kp = S.fa[0].Kp;
  for (int x = threadIdx.x; x < 32 * 32; x += blockDim.x) {
    const int r = x >> 5, c = x & 31;
    tab[x] = (r < Kp && c < Kp) ? S.fa[0].table[r * Kp + c] : (int8_t)0;
  }
  const int stride = fill_slot_stride(2, Kp);
  const int base = 1024 + WALK_FW * 64 * stride;
  const int f = wave & (WALK_FW - 1);
  uint32_t *wlds = (uint32_t *)(lds + base) + f * WALK_LANES * WSTRIDE;
  WalkQ *wq0 = (WalkQ *)(lds + base + WALK_FW * WALK_LANES * WSTRIDE * 4);
  WalkQ *wq = wq0 + f;
  CoopQ *cq = (CoopQ *)(wq0 + WALK_FW);
  SQItem *it = (SQItem *)((uint8_t *)(wq0 + WALK_FW) + sizeof(CoopQ));  // the round's item
  int32_t *ctrs = (int32_t *)(it + WALK_FW);  // [0] task counter, [1..2] barrier, [3] round state
  if (wave < WALK_FW && lane == 0) { wq->tail = 0; wq->head = 0; wq->done = 0; }
  if (threadIdx.x == 0) { ctrs[0] = 0; ctrs[1] = 0; ctrs[2] = 0; ctrs[3] = 0; }
  __syncthreads();
  if (wave >= WALK_FW) {
    if (lane < WALK_LANES) devq_walk(S, wq, tab, wlds);
    return;
  }
  int8_t *slot = (int8_t *)lds + 1024 + threadIdx.x * stride;
  const size_t gw = (size_t)blockIdx.x;  // this workgroup's scratch
  int round = 0;
  for (;;) {
    if (wave == 0) {
      const int64_t c = devq_claim(S);
      if (c >= 0) {
        const uint32_t *src = (const uint32_t *)(S.items + (c & (SQ_CAP - 1)));
        if (lane < SQ_TASK_WORDS)
          ((uint32_t *)it)[lane] = __hip_atomic_load(src + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const Task T = it->task;
        // this task's codes into the workgroup's scratch
        stage_codes((const uint4 *)(S.hcodes + T.row_off), (uint4 *)const_cast<uint16_t *>(S.fa[0].rowcodes + gw * SQ_ROWS),
                    T.S * GWAVE * R / 8);
        stage_codes((const uint4 *)(S.hcodes + T.col_off), (uint4 *)const_cast<uint16_t *>(S.fa[0].colcodes + gw * SQ_COLS),
                    (8 * T.NB + 8) / 8);
        stage_pairs(S, T);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        if (lane == 0) {
          it->task.row_off = gw * SQ_ROWS;
          it->task.col_off = gw * SQ_COLS;
          it->task.ring_off = gw * SQ_RING;
          ctrs[0] = 0;
          cq->epoch = 0;
          cq->arrived = 0;
        }
      }
      if (lane == 0) ctrs[3] = c >= 0 ? 1 : 0;
    }
    if (!fill_barrier(ctrs + 1)) break;
    if (!__hip_atomic_load(ctrs + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
    const Task *tk = &it->task;
    WalkQ *to = wq0 + (round & (WALK_FW - 1));  // the walker that takes this item's pairs
    switch (it->kind & 3) {
      case 0: devq_fill<2, false>(S.fa[0], tk, ctrs, tab, slot, to, cq); break;
      case 1: devq_fill<2, true>(S.fa[1], tk, ctrs, tab, slot, to, cq); break;
      case 2: devq_fill<1, false>(S.fa[0], tk, ctrs, tab, slot, to, cq); break;
      default: devq_fill<1, true>(S.fa[1], tk, ctrs, tab, slot, to, cq); break;
    }
    ++round;
    if (!fill_barrier(ctrs + 1)) break;  // every wave is out of the task before wave 0 resets its state
  }
  if (wave == 0 && lane < WALK_FW) __hip_atomic_store(&wq0[lane].done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

}  // namespace

int devq_lds_bytes(int Kp) {
  return walk_lds_bytes(2, Kp) + WALK_FW * ((int)sizeof(SQItem) + 4);
}
size_t devq_scratch_rows() { return SQ_ROWS; }
size_t devq_scratch_cols() { return SQ_COLS; }
size_t devq_scratch_ring() { return SQ_RING; }

int devq_launch(const SQArgs *S, int grid, int lds, hipStream_t st) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(devq_kernel, dim3(grid), dim3(128 * WALK_FW), lds, st, S);
  return hipGetLastError() == hipSuccess ? TA_OK : TA_E_DEVICE;
}

}  // namespace ta

// ---------------------------------------------------------------------------
// Host side.
using namespace ta;

namespace {

constexpr int32_t NSLOT = 1 << 16;   // pair records
constexpr size_t HDATA = (size_t)256 << 20, CIGB = (size_t)256 << 20;

// First-fit allocator over [0, cap) that merges freed neighbours.
struct RangeAlloc {
  std::map<size_t, size_t> free_;  // offset -> length
  void init(size_t cap) {
    free_.clear();
    if (cap) free_[0] = cap;
  }
  bool take(size_t n, size_t align, size_t &off) {
    n = std::max<size_t>(n, 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      const size_t s = it->first, e = s + it->second, a = (s + align - 1) / align * align;
      if (a + n > e) continue;
      free_.erase(it);
      if (a > s) free_[s] = a - s;
      if (a + n < e) free_[a + n] = e - (a + n);
      off = a;
      return true;
    }
    return false;
  }
  void give(size_t off, size_t n) {
    n = std::max<size_t>(n, 1);
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

template <typename T>
T *host_coherent(size_t n) {
  void *p = nullptr;
  if (hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocCoherent) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  memset(p, 0, std::max<size_t>(n, 1) * sizeof(T));
  return (T *)p;
}

}  // namespace

struct DevqJob {
  ta_params p{};
  size_t slot0 = 0, nslot = 0;           // pair records
  size_t checked = 0;                    // records before this one have their flags set
  size_t h_off = 0, h_len = 0, c_off = 0, c_len = 0, t_off = 0, t_len = 0;
  uint32_t first_item = 0;
  int32_t ngpu = 0;                      // pairs the kernel walks
  std::vector<int32_t> slot;             // request pair -> record, -1: empty side (done on the host)
  std::vector<ta_result> host_res;       // empty-side results
  std::vector<std::string> host_cig;
};

struct DevQ {
  ta_ctx *ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  bool launched = false;
  int grid = 0, lds = 0, Kp = 0, maxS = 0, minS = 0;
  int8_t map[256];
  int best[32];
  // coherent host memory
  SQItem *items = nullptr;
  uint32_t *tail = nullptr, *flag = nullptr;
  PairDev *pairs = nullptr;
  PairTB *ptb = nullptr;
  ResultDev *res = nullptr;
  uint8_t *hdata = nullptr, *cig = nullptr;
  // device memory
  uint32_t *head = nullptr;
  unsigned long long *mirror = nullptr;
  PairDev *dpairs = nullptr;
  PairTB *dptb = nullptr;
  uint8_t *dcodes = nullptr;
  uint16_t *wrow = nullptr, *wcol = nullptr;
  uint32_t *wring = nullptr;
  FillOut *fo = nullptr;
  uint8_t *tb = nullptr;
  int8_t *table = nullptr;
  int32_t *err = nullptr;
  SQArgs *dargs = nullptr;  // the kernel's arguments (device copy of hargs)
  SQArgs hargs{};
  size_t tb_cap = 0;
  RangeAlloc ha, ca, ta, slots;
  uint32_t published = 0;
  std::multiset<uint32_t> firsts;        // first items of the jobs in flight (ring reuse)
  bool lane_set[2] = {false, false};
  int32_t lane_open[2] = {0, 0}, lane_ext[2] = {0, 0};
  bool stopping = false;   // a stop was requested from the running kernel
  double busy_at = 0;      // last time jobs were in flight (seconds)
  std::mutex mu;
};

namespace {

void free_all(DevQ *q) {
  for (void *h : {(void *)q->items, (void *)q->tail, (void *)q->flag, (void *)q->pairs, (void *)q->ptb,
                  (void *)q->res, (void *)q->hdata, (void *)q->cig})
    if (h) (void)hipHostFree(h);
  for (void *d : {(void *)q->head, (void *)q->mirror, (void *)q->dpairs, (void *)q->dptb, (void *)q->dcodes,
                  (void *)q->wrow, (void *)q->wcol, (void *)q->wring,
                  (void *)q->fo, (void *)q->tb, (void *)q->table, (void *)q->err, (void *)q->dargs})
    if (d) (void)hipFree(d);
  if (q->ev) (void)hipEventDestroy(q->ev);
  if (q->stream) (void)hipStreamDestroy(q->stream);
}

FillArgs fill_args(DevQ *q, int local) {
  FillArgs a{};
  a.table = q->table; a.Kp = q->Kp;
  a.rowcodes = q->wrow; a.colcodes = q->wcol; a.ring = q->wring; a.tb = q->tb; a.out = q->fo;
  a.open = q->lane_open[local]; a.ext = q->lane_ext[local];
  a.sat_limit = 32767 - 2 * q->maxS;
  a.bias = -q->minS;
  a.walk = 1;
  a.ncoop = 0;
  a.err = q->err;
  a.spin_limit = 1 << 24;
  TbArgs &w = a.wk;  // pair records and residues: the device copies; results and runs: host memory
  w.pairs = q->dpairs; w.npairs = NSLOT; w.order = nullptr; w.fo = q->fo; w.ptb = q->dptb;
  w.codes = q->dcodes; w.table = q->table; w.Kp = q->Kp; w.tb = q->tb; w.cig = (char *)q->cig; w.res = q->res;
  w.local = local; w.open = a.open; w.ext = a.ext; w.do_cigar = 1;
  w.sys = 1;
  w.guard_cap = INT32_MAX;
  if (const char *e = getenv("TWEAKALIGN_WALK_GUARD")) w.guard_cap = std::max(0, atoi(e));
  return a;
}

void launch(DevQ *q) {
  SQArgs &S = q->hargs;
  S = SQArgs{};
  S.fa[0] = fill_args(q, 0);
  S.fa[1] = fill_args(q, 1);
  S.items = q->items; S.tail = q->tail; S.head = q->head; S.mirror = q->mirror;
  S.stop = q->tail + 32;  // its own 128 bytes of the same block
  S.halt = q->head + 32;
  S.hcodes = (const uint16_t *)q->hdata;
  S.dcodes = q->dcodes;
  S.hpairs = q->pairs; S.hptb = q->ptb; S.dpairs = q->dpairs; S.dptb = q->dptb;
  S.flag = q->flag;
  S.idle_ticks = 200000000;  // 2 s: a safety bound; the host stops an idle kernel itself
  if (hipMemsetAsync(q->head + 32, 0, 4, q->stream) == hipSuccess &&  // halt cleared before the launch
      hipMemcpyAsync(q->dargs, &q->hargs, sizeof(SQArgs), hipMemcpyHostToDevice, q->stream) == hipSuccess &&
      devq_launch(q->dargs, q->grid, q->lds, q->stream) == TA_OK && hipEventRecord(q->ev, q->stream) == hipSuccess)
    q->launched = true;
}

void release(DevQ *q, DevqJob *j) {
  if (j->h_len) q->ha.give(j->h_off, j->h_len);
  if (j->c_len) q->ca.give(j->c_off, j->c_len);
  if (j->t_len) q->ta.give(j->t_off, j->t_len);
  if (j->nslot) q->slots.give(j->slot0, j->nslot);
  if (j->ngpu) {
    auto it = q->firsts.find(j->first_item);
    if (it != q->firsts.end()) q->firsts.erase(it);
  }
}

size_t up16(size_t x) { return (x + 15) & ~(size_t)15; }

}  // namespace

DevQ *devq_create(ta_ctx *ctx, int64_t tb_budget) {
  if (const char *e = getenv("TWEAKALIGN_DEVICE_QUEUE"))
    if (e[0] == '0') return nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return nullptr;
  DevQ *q = new DevQ();
  q->ctx = ctx;
  q->device = ctx->device;
  int8_t S[32 * 32];
  int32_t K = 0;
  ta_matrix(TA_PROTEIN, 0, 0, S, &K);
  q->Kp = K + 1;
  q->maxS = -128; q->minS = 127;
  for (int r = 0; r < K; ++r) {
    q->best[r] = 0;
    for (int c = 0; c < K; ++c) {
      q->maxS = std::max<int>(q->maxS, S[r * K + c]);
      q->minS = std::min<int>(q->minS, S[r * K + c]);
      q->best[r] = std::max<int>(q->best[r], S[r * K + c]);
    }
  }
  for (int b = 0; b < 256; ++b) {
    uint8_t code = 0;
    const char ch = (char)b;
    q->map[b] = ta_encode(TA_PROTEIN, &ch, 1, &code) == TA_OK ? (int8_t)code : (int8_t)-1;
  }
  q->lds = devq_lds_bytes(q->Kp);
  q->grid = std::max(1, ctx->cus - SQ_FREE_CUS);
  bool ok = q->lds <= ctx->lds_max;
  const size_t nw = (size_t)q->grid;  // scratch per workgroup
  ok = ok && (q->items = host_coherent<SQItem>(SQ_CAP)) && (q->tail = host_coherent<uint32_t>(64)) &&
       (q->flag = host_coherent<uint32_t>(NSLOT)) && (q->pairs = host_coherent<PairDev>(NSLOT)) &&
       (q->ptb = host_coherent<PairTB>(NSLOT)) && (q->res = host_coherent<ResultDev>(NSLOT)) &&
       (q->hdata = host_coherent<uint8_t>(HDATA)) && (q->cig = host_coherent<uint8_t>(CIGB));
  auto dev = [&](void **p, size_t bytes) { return ok && hipMalloc(p, bytes) == hipSuccess; };
  ok = dev((void **)&q->head, 256) && dev((void **)&q->mirror, 256) &&
       dev((void **)&q->dpairs, NSLOT * sizeof(PairDev)) && dev((void **)&q->dptb, NSLOT * sizeof(PairTB)) &&
       dev((void **)&q->dcodes, HDATA) &&
       dev((void **)&q->wrow, nw * devq_scratch_rows() * 2) && dev((void **)&q->wcol, nw * devq_scratch_cols() * 2) &&
       dev((void **)&q->wring, nw * devq_scratch_ring() * 4) && dev((void **)&q->fo, NSLOT * sizeof(FillOut)) &&
       dev((void **)&q->table, 1024) && dev((void **)&q->err, 256) && dev((void **)&q->dargs, sizeof(SQArgs));
  // the plane arena: the traceback budget, at most 16 GB, halved until it fits
  q->tb_cap = (size_t)std::min<int64_t>(tb_budget, (int64_t)16 << 30);
  while (ok && q->tb_cap >= ((size_t)256 << 20) && hipMalloc((void **)&q->tb, q->tb_cap) != hipSuccess) {
    (void)hipGetLastError();
    q->tb_cap /= 2;
  }
  ok = ok && q->tb;
  if (ok) {
    std::vector<int8_t> t((size_t)q->Kp * q->Kp, PAD_SCORE);  // pad row / column
    for (int r = 0; r < K; ++r)
      for (int c = 0; c < K; ++c) t[(size_t)r * q->Kp + c] = S[r * K + c];
    ok = hipMemcpy(q->table, t.data(), t.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(q->head, 0, 256) == hipSuccess && hipMemset(q->mirror, 0, 256) == hipSuccess &&
         hipMemset(q->err, 0, 256) == hipSuccess &&
         hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&q->ev, hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    free_all(q);
    delete q;
    return nullptr;
  }
  q->ha.init(HDATA);
  q->ca.init(CIGB);
  q->ta.init(q->tb_cap);
  q->slots.init(NSLOT);
  return q;
}

void devq_destroy(DevQ *q) {
  if (!q) return;
  (void)hipSetDevice(q->device);
  if (q->tail) __atomic_store_n(q->tail + 32, 1u, __ATOMIC_RELEASE);  // the kernel leaves
  if (q->stream) (void)hipStreamSynchronize(q->stream);
  free_all(q);
  delete q;
}

bool devq_eligible(DevQ *q, const ta_params *p, const std::vector<std::string> &a, const std::vector<std::string> &b) {
  if (!q || p->alphabet != TA_PROTEIN || !p->traceback) return false;
  if (p->lanes != TA_LANES_AUTO && p->lanes != TA_LANES_I16) return false;
  if (a.size() > (size_t)NSLOT / 8) return false;
  for (size_t k = 0; k < a.size(); ++k) {
    if (a[k].size() > (size_t)SQ_MAXLEN || b[k].size() > (size_t)SQ_MAXLEN) return false;
    // global pairs the batch runtime sends to the long-pair pipeline stay there
    if (p->mode == TA_GLOBAL && (int64_t)a[k].size() * (int64_t)b[k].size() >= (1ll << 25)) return false;
  }
  std::lock_guard<std::mutex> g(q->mu);
  const int m = p->mode == TA_LOCAL ? 1 : 0;
  if (!q->lane_set[m]) {
    // The first request of a mode sets its lane's gap scores.  A running
    // kernel holds the lanes it was launched with: it is asked to leave (it
    // finishes the items it holds; the rest wait for the next launch).
    if (q->launched) {
      __atomic_store_n(q->tail + 32, 1u, __ATOMIC_RELEASE);
      (void)hipSetDevice(q->device);
      (void)hipStreamSynchronize(q->stream);
      __atomic_store_n(q->tail + 32, 0u, __ATOMIC_RELEASE);
      q->launched = false;
      q->stopping = false;
    }
    q->lane_set[m] = true;
    q->lane_open[m] = p->gap_open;
    q->lane_ext[m] = p->gap_extend;
  }
  return q->lane_open[m] == p->gap_open && q->lane_ext[m] == p->gap_extend;
}

int devq_admit(DevQ *q, const ta_params *p, const std::vector<std::string> &a, const std::vector<std::string> &b,
               DevqJob **out) {
  *out = nullptr;
  const int n = (int)a.size();
  const bool local = p->mode == TA_LOCAL;
  const int pad = q->Kp - 1;
  DevqJob *j = new DevqJob();
  j->p = *p;
  j->slot.assign(n, -1);
  j->host_res.assign(n, ta_result{});
  j->host_cig.assign(n, std::string());
  // empty sides: the batch path's results, on the host
  std::vector<int> gp;
  for (int k = 0; k < n; ++k) {
    const int la = (int)a[k].size(), lb = (int)b[k].size();
    if (la && lb) { gp.push_back(k); continue; }
    ta_result &r = j->host_res[k];
    r.cigar_off = 0;
    if (local) {
      r.a_begin = r.a_end = r.b_begin = r.b_end = -1;
      r.flags = TA_F_EMPTY;
    } else {
      const int len = la + lb;
      r.score = len ? p->gap_open + (len - 1) * p->gap_extend : 0;
      r.a_begin = 0; r.a_end = la - 1; r.b_begin = 0; r.b_end = lb - 1;
      if (len) j->host_cig[k] = std::to_string(len) + (la ? "I" : "D");
      r.cigar_len = (int32_t)j->host_cig[k].size();
    }
  }
  // codes, per-pair lane width
  std::vector<std::vector<uint8_t>> ca(n), cb(n);
  std::vector<int> P(n, 2);
  for (int k : gp) {
    auto enc = [&](const std::string &s, std::vector<uint8_t> &o, int64_t &bound) {
      o.resize(s.size());
      bound = 0;
      for (size_t i = 0; i < s.size(); ++i) {
        const int c = q->map[(uint8_t)s[i]];
        if (c < 0) return false;
        o[i] = (uint8_t)c;
        bound += q->best[c];
      }
      return true;
    };
    int64_t ba = 0, bb = 0;
    if (!enc(a[k], ca[k], ba) || !enc(b[k], cb[k], bb)) {
      delete j;
      return TA_E_ALPHABET;
    }
    const int la = (int)a[k].size(), lb = (int)b[k].size();
    P[k] = local ? (std::min(ba, bb) > 32767 - 2L * q->maxS ? 1 : 2) : (global_int16_safe(p, q->maxS, la, lb) ? 2 : 1);
  }
  // couples (the batch planner's rule): local int16 pairs by (strips, columns)
  // descending, neighbours together; global int16 pairs of equal columns
  struct Tk { int lo, hi; bool single; };
  std::vector<Tk> tasks;
  {
    std::vector<int> i16, i32;
    for (int k : gp) (P[k] == 2 ? i16 : i32).push_back(k);
    auto key = [&](int k) { return std::make_pair(-((int)a[k].size() + STRIP - 1) / STRIP, -(int)b[k].size()); };
    if (local) {
      std::stable_sort(i16.begin(), i16.end(), [&](int x, int y) { return key(x) < key(y); });
      for (size_t t = 0; t < i16.size(); t += 2)
        tasks.push_back(t + 1 < i16.size() ? Tk{i16[t], i16[t + 1], false} : Tk{i16[t], i16[t], true});
    } else {
      std::map<int, std::vector<int>> byn;
      for (int k : i16) byn[(int)b[k].size()].push_back(k);
      for (auto &kv : byn)
        for (size_t t = 0; t < kv.second.size(); t += 2)
          tasks.push_back(t + 1 < kv.second.size() ? Tk{kv.second[t], kv.second[t + 1], false}
                                                   : Tk{kv.second[t], kv.second[t], true});
    }
    for (int k : i32) tasks.push_back(Tk{k, k, true});
  }
  // sizes
  size_t hbytes = 0, cbytes = 0, tbytes = 0;
  for (int k : gp) {
    hbytes += up16(a[k].size()) + up16(b[k].size());
    cbytes += 4 * (a[k].size() + b[k].size() + 1);
  }
  hbytes += 32;  // the walker reads 12-byte windows: slack after the last sequence
  std::vector<Task> T(tasks.size());
  for (size_t t = 0; t < tasks.size(); ++t) {
    const Tk &x = tasks[t];
    Task &u = T[t];
    const int la0 = (int)a[x.lo].size(), la1 = (int)a[x.hi].size();
    const int lb0 = (int)b[x.lo].size(), lb1 = (int)b[x.hi].size();
    u = Task{};
    u.m[0] = la0; u.m[1] = la1;
    u.n = std::max(lb0, lb1);
    u.off[0] = u.n - lb0; u.off[1] = u.n - lb1;
    u.S = (std::max(la0, la1) + GWAVE * R - 1) / (GWAVE * R);
    u.NB = (u.n + GWAVE + 7) / 8;
    u.single = x.single ? 1 : 0;
    u.ring_len = u.S > 1 ? 8 * u.NB + 2 * GWAVE + 16 : 0;
    hbytes += up16(2 * (size_t)u.S * GWAVE * R) + up16(2 * (size_t)(8 * u.NB + 8));
    u.tb_off = tbytes;
    tbytes += (uint64_t)u.S * tb_strip_bytes(P[x.lo], u.NB, GWAVE);
  }
  const int ng = (int)gp.size();
  std::lock_guard<std::mutex> g(q->mu);
  // room: ring slots, records, memory
  const uint32_t oldest = q->firsts.empty() ? q->published : *q->firsts.begin();
  if (q->published + (uint32_t)tasks.size() - oldest > SQ_CAP) {
    delete j;
    return TA_E_PENDING;
  }
  bool got_s = ng == 0 || q->slots.take((size_t)ng, 1, j->slot0);
  bool got_h = got_s && (ng == 0 || q->ha.take(hbytes, 16, j->h_off));
  bool got_c = got_h && (ng == 0 || q->ca.take(cbytes, 16, j->c_off));
  bool got_t = got_c && (tbytes == 0 || q->ta.take(tbytes, 256, j->t_off));
  if (!got_t) {
    if (got_c && ng) q->ca.give(j->c_off, cbytes);
    if (got_h && ng) q->ha.give(j->h_off, hbytes);
    if (got_s && ng) q->slots.give(j->slot0, (size_t)ng);
    delete j;
    return TA_E_PENDING;
  }
  j->nslot = ng ? (size_t)ng : 0;
  j->h_len = ng ? hbytes : 0;
  j->c_len = ng ? cbytes : 0;
  j->t_len = tbytes;
  j->ngpu = ng;

  // residue codes, pair records
  size_t hp = j->h_off, cp = j->c_off;
  std::vector<uint64_t> aoff(n), boff(n);
  for (int t = 0; t < ng; ++t) {
    const int k = gp[t];
    const int s = (int)j->slot0 + t;
    j->slot[k] = s;
    aoff[k] = hp;
    memcpy(q->hdata + hp, ca[k].data(), ca[k].size());
    memset(q->hdata + hp + ca[k].size(), 0, up16(ca[k].size()) - ca[k].size());
    hp += up16(ca[k].size());
    boff[k] = hp;
    memcpy(q->hdata + hp, cb[k].data(), cb[k].size());
    memset(q->hdata + hp + cb[k].size(), 0, up16(cb[k].size()) - cb[k].size());
    hp += up16(cb[k].size());
    PairDev &d = q->pairs[s];
    d.a_off = aoff[k]; d.b_off = boff[k];
    d.la = (int32_t)a[k].size(); d.lb = (int32_t)b[k].size();
    d.cig_cap = 4 * (d.la + d.lb + 1);
    d.cig_off = cp;
    cp += (size_t)d.cig_cap;
    d.pad_ = local ? (1 << 30) : 0;
    __atomic_store_n(q->flag + s, 0u, __ATOMIC_RELAXED);
  }
  memset(q->hdata + hp, 0, 32);
  hp += 32;
  // tasks: row and column codes, planes, items
  for (size_t t = 0; t < tasks.size(); ++t) {
    const Tk &x = tasks[t];
    Task &u = T[t];
    u.pair[0] = j->slot[x.lo]; u.pair[1] = j->slot[x.hi];
    u.tb_off += j->t_off;
    const int rows = u.S * GWAVE * R, cols = 8 * u.NB + 8;
    uint16_t *rc = (uint16_t *)(q->hdata + hp);
    for (int r = 0; r < rows; ++r) {
      const int c0 = r < u.m[0] ? ca[x.lo][r] : pad, c1 = r < u.m[1] ? ca[x.hi][r] : pad;
      rc[r] = (uint16_t)(c0 | (c1 << 8));
    }
    u.row_off = hp / 2;
    hp += up16(2 * (size_t)rows);
    uint16_t *cc = (uint16_t *)(q->hdata + hp);
    const int lb0 = (int)cb[x.lo].size(), lb1 = (int)cb[x.hi].size();
    for (int c = 0; c < cols; ++c) {
      const int j0 = c - u.off[0], j1 = c - u.off[1];
      const int c0 = (j0 >= 0 && j0 < lb0) ? cb[x.lo][j0] : pad, c1 = (j1 >= 0 && j1 < lb1) ? cb[x.hi][j1] : pad;
      cc[c] = (uint16_t)(c0 | (c1 << 8));
    }
    u.col_off = hp / 2;
    hp += up16(2 * (size_t)cols);
    for (int h = 0; h < (x.single ? 1 : 2); ++h) {
      PairTB &d = q->ptb[h ? u.pair[1] : u.pair[0]];
      d.tb_off = u.tb_off; d.P = P[x.lo]; d.half = h; d.NB = u.NB; d.off = u.off[h]; d.GG = GWAVE;
    }
    SQItem &it = q->items[(q->published + (uint32_t)t) & (SQ_CAP - 1)];
    it.task = u;
    it.kind = (local ? 1 : 0) | (P[x.lo] == 1 ? 2 : 0);
  }
  j->first_item = q->published;
  if (!tasks.empty()) {
    q->firsts.insert(j->first_item);
    q->published += (uint32_t)tasks.size();
    __atomic_store_n(q->tail, q->published, __ATOMIC_RELEASE);  // items before the index
  }
  *out = j;
  return TA_OK;
}

bool devq_done(DevQ *q, DevqJob *j) {
  while (j->checked < (size_t)j->ngpu && __atomic_load_n(q->flag + j->slot0 + j->checked, __ATOMIC_ACQUIRE))
    ++j->checked;
  return j->checked == (size_t)j->ngpu;
}

int devq_finish(DevQ *q, DevqJob *j, std::vector<ta_result> &res, std::string &text) {
  const int n = (int)j->slot.size();
  res.assign(n, ta_result{});
  text.clear();
  int rc = TA_OK;
  for (int k = 0; k < n; ++k) {
    ta_result &o = res[k];
    const int s = j->slot[k];
    if (s < 0) {
      o = j->host_res[k];
      o.cigar_off = (int64_t)text.size();
      text += j->host_cig[k];
      continue;
    }
    const ResultDev r = q->res[s];
    o.score = r.score; o.a_begin = r.a_begin; o.a_end = r.a_end; o.b_begin = r.b_begin; o.b_end = r.b_end;
    o.flags = r.flags;
    o.cigar_off = (int64_t)text.size();
    const uint32_t *runs = (const uint32_t *)(q->cig + r.cigar_off);
    const int nr = r.cigar_len / 4;
    const int64_t len = nr > 0 ? cigar_text(runs, nr, nullptr) : 0;
    text.resize(text.size() + (size_t)len);
    if (len) cigar_text(runs, nr, &text[(size_t)o.cigar_off]);
    o.cigar_len = (int32_t)len;
    if (r.flags & TA_F_WALK_FAILED) rc = TA_E_WALK;
  }
  std::lock_guard<std::mutex> g(q->mu);
  release(q, j);
  delete j;
  return rc;
}

void devq_drop(DevQ *q, DevqJob *j) {
  std::lock_guard<std::mutex> g(q->mu);
  release(q, j);
  delete j;
}

void devq_tick(DevQ *q, bool have_jobs) {
  std::lock_guard<std::mutex> g(q->mu);
  const double now = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  if (q->launched && hipEventQuery(q->ev) == hipSuccess) {
    q->launched = false;
    q->stopping = false;
  }
  if (have_jobs) q->busy_at = now;
  if (!q->launched) {
    if (have_jobs) {
      __atomic_store_n(q->tail + 32, 0u, __ATOMIC_RELEASE);
      (void)hipSetDevice(q->device);
      launch(q);
    }
  } else if (!have_jobs && !q->stopping && now - q->busy_at > 0.05) {
    // nothing in flight for 50 ms: let the kernel go (the next job relaunches it)
    __atomic_store_n(q->tail + 32, 1u, __ATOMIC_RELEASE);
    q->stopping = true;
  }
}
