// Gotoh DP fill for short and medium pairs on gfx950.
//
// Layout of the work (no counterpart in the reference, which spawns an
// external aligner; recurrence, boundaries and tie-breaks follow the
// specification the CPU oracle implements, oracle/gotoh.c):
//  * a wave64 = NGW DP groups of G lanes (16: one DPP row each; 8: two per
//    row, interleaved) or, for long tasks, one whole-wave group; a group owns
//    a task and sweeps it strip by strip;
//  * a strip is G lanes x R query rows held in registers (64 lanes for a
//    whole-wave group); lane l works on column j = t - l at step t, so a
//    group covers an anti-diagonal band and H/F of a lane's bottom row reach
//    the next lane one step later through DPP row_shr:1 (row_shr:2 for 8-lane
//    groups, wave_shr:1 for the whole wave; lane 0 takes the strip's top
//    boundary instead);
//  * the bottom row of a strip goes to a ring in global memory and is read
//    back as the top boundary of the next strip;
//  * P == 2 packs two pairs ("a couple") into the int16 halves of every
//    register (saturating v_pk_* arithmetic); P == 1 is one pair in int32;
//  * the substitution scores of a lane's R rows for every column residue
//    are staged in a private LDS slot at the start of each strip, so a step
//    costs one LDS read per half; the score-only packed kernels keep them as
//    4-bit scores (biased to 0..15), which halves the slot so that two or
//    three waves fit per SIMD (they need them to hide the packed ops'
//    latency); the kernels that write direction planes run one wave per SIMD
//    and keep bytes (fewer instructions per cell);
//  * direction bits (4 per cell) leave as sign bits of int differences,
//    packed per step and stored per 8-step block with dwordx4 stores, block
//    major ([strip][block][lane]) so the lanes of a group store one
//    contiguous run per block (lane-major writes scattered 16-byte pieces
//    over 64 cache lines and cost ~10% of the fill).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.h"
#include "traceback.h"

namespace ta {

typedef short s2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t u32(s2v v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s2v s2(uint32_t v) { return __builtin_bit_cast(s2v, v); }

template <int P>
struct Lane;

template <>
struct Lane<2> {
  using V = s2v;
  static __device__ __forceinline__ V add(V a, V b) { return __builtin_elementwise_add_sat(a, b); }
  static __device__ __forceinline__ V sub(V a, V b) { return __builtin_elementwise_sub_sat(a, b); }
  static __device__ __forceinline__ V vmax(V a, V b) { return __builtin_elementwise_max(a, b); }
  static __device__ __forceinline__ V splat(int x) { return V{(short)x, (short)x}; }
  static __device__ __forceinline__ int bits(V v) { return (int)u32(v); }
  static __device__ __forceinline__ V unbits(int x) { return s2((uint32_t)x); }
  static constexpr int NEG = NEG16;
};

template <>
struct Lane<1> {
  using V = int;
  static __device__ __forceinline__ V add(V a, V b) { return a + b; }
  static __device__ __forceinline__ V sub(V a, V b) { return a - b; }
  static __device__ __forceinline__ V vmax(V a, V b) { return a > b ? a : b; }
  static __device__ __forceinline__ V splat(int x) { return x; }
  static __device__ __forceinline__ int bits(V v) { return v; }
  static __device__ __forceinline__ V unbits(int x) { return x; }
  static constexpr int NEG = NEG32;
};

struct FillArgs {
  const Task *tasks[3];  // [0] whole-wave tasks (GG = 64), [1] wave units of NGW G-lane tasks,
                         // [2] wave units of 8 8-lane tasks (packed couples)
  int32_t ntasks[3];     // tasks, wave units, wave units
  int32_t *ctr;          // work-queue heads ctr[0], ctr[1]
  int32_t *ctr8;         // work-queue head of tasks[2]
  const int8_t *table;   // Kp x Kp, row/column Kp-1 = pad
  int32_t Kp;
  const uint16_t *rowcodes;
  const uint16_t *colcodes;
  uint32_t *ring;
  uint8_t *tb;
  FillOut *out;
  int32_t open, ext;
  int32_t sat_limit;     // local int16: best score above this -> int32 rerun
  int32_t bias;          // 4-bit slots: table entries are score + bias (0..15)
  int32_t walk;          // direction-plane kernels: a second wave walks the pairs the first fills
  TbArgs wk;             // the walks' arguments (walk != 0)
  int32_t ncoop;         // walk != 0: the first ncoop whole-wave tasks are swept by a workgroup's
                         // fill waves together, strips pipelined (counter coop_ctr)
  int32_t *coop_ctr;
  int32_t *err;          // | 4: a cooperative wait gave up
  int32_t spin_limit;
};

// Walks inside the fill kernel.  Each fill wave then has a walker wave; the
// fill wave hands each finished task's pairs to its walker through an LDS
// ring, and the walker walks them while the fill goes on.  Both waves sit in
// one workgroup (one CU), so a workgroup-scope release / acquire orders the
// fill's plane and result stores before the walker's loads, and neither wave
// ever waits on a wave outside its workgroup: the walker waits only for its
// fill wave, the fill wave only for ring space the walker frees.  Once its
// queues are empty the fill wave walks too, from the same ring, with the LDS
// of its profile slots as walker slices (the walks of the last tasks then
// have up to three times the lanes).
constexpr int WQ_CAP = 64;  // ring entries (pair ids)
struct WalkQ {
  int32_t ids[WQ_CAP];
  int32_t tail;   // entries pushed (the fill wave)
  int32_t head;   // entries read (every consumer adds the ones it read)
  int32_t done;   // fill wave finished (tail is final)
  int32_t claim;  // entries claimed (consumers, by compare-and-swap)
};
constexpr int WALK_FW = 4;      // fill waves per workgroup when walking (one per SIMD)
constexpr int WALK_LANES = 32;  // walker lanes per fill wave

// Cooperative whole-wave tasks (batches of few tasks, where one long pair on
// one wave would set the run time): the WALK_FW fill waves of a workgroup
// sweep one task together, wave w taking strips w, w + WALK_FW, ...  Strip s
// reads strip s-1's bottom row from the task's ring once strip s-1 has
// published that it stored it (LDS counters, workgroup scope: all strips of
// the task run on one CU).  Wave 0 leads: it starts a task when all waves
// have arrived from the previous one, then reduces their best cells and
// hands the pair to its walker.
struct CoopQ {
  int32_t task, epoch, arrived, pad_;
  int32_t prog[8];                 // strip s publishes in prog[s % 8]: ((s + 1) << 20) | blocks stored
  int32_t best[WALK_FW][2][3];     // per wave and half: H, i, j of its strips' best cell
};
// LDS of a walking workgroup: table, WALK_FW fill waves' slots, walker slices, rings.
__host__ __device__ constexpr int walk_lds_bytes(int P, int Kp) {
  return 1024 + WALK_FW * (64 * (P * Kp * R + 8) + WALK_LANES * WSTRIDE * 4 + (int)sizeof(WalkQ)) +
         (int)sizeof(CoopQ);
}

// Fill side, at a task's end: append its pairs to the ring.  NG groups
// (1 for a whole-wave task, NGW for a wave unit) with their `single` codes and pair
// ids as seen by every lane; a filler (single 2) has no pair, a single
// (1) only its lo pair.
template <int NG>
__device__ __forceinline__ void walk_push(WalkQ *q, int &tail, const int (&sg)[NG], const int (&a0)[NG],
                                          const int (&a1)[NG]) {
  int cnt = 0;
#pragma unroll
  for (int g = 0; g < NG; ++g) cnt += (sg[g] != 2) + (sg[g] == 0);
  // every plane and result store of the task acknowledged, then the ring
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if ((threadIdx.x & 63) == 0) {
    while (tail + cnt - __hip_atomic_load(&q->head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) > WQ_CAP)
      __builtin_amdgcn_s_sleep(2);
    int t = tail;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      if (sg[g] != 2) q->ids[t++ & (WQ_CAP - 1)] = a0[g];
      if (sg[g] == 0) q->ids[t++ & (WQ_CAP - 1)] = a1[g];
    }
    __hip_atomic_store(&q->tail, t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  tail += cnt;
}

// Walker side: each lane walks one pair at a time and takes the next
// handed-over pair as soon as its walk ends, until the fill wave is done and
// the ring is empty.  Lane 0 claims as many entries as the wave has idle
// lanes (a compare-and-swap on `claim`: the fill wave may consume the same
// ring once it is done filling); the lanes (lane 0 among them) run the loop.
template <int P>
__device__ __forceinline__ void walk_loop(const FillArgs &A, WalkQ *q, const int8_t *tab, uint32_t *wlds) {
  const int wl = threadIdx.x & 63;
  uint32_t *slice = wlds + wl * WSTRIDE;
  Walker<P> w;
  bool active = false;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (idle) {
      // done before tail: once done is seen, the tail read after it is final
      const int done = __hip_atomic_load(&q->done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
      int base = 0, take = 0;
      if (wl == 0) {
        const int nidle = __popcll(idle);
        int c = __hip_atomic_load(&q->claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        for (;;) {
          const int t = __hip_atomic_load(&q->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
          const int k = t - c < nidle ? t - c : nidle;
          if (k <= 0) break;
          // on failure c is reloaded with the current claim count
          if (__hip_atomic_compare_exchange_strong(&q->claim, &c, c + k, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP)) {
            base = c;
            take = k;
            break;
          }
        }
      }
      base = __shfl(base, 0, 64);
      take = __shfl(take, 0, 64);
      if (take > 0) {
        // rank of this lane among the idle ones
        const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0));
        if (!active && rank < take) {
          const int p = q->ids[(base + rank) & (WQ_CAP - 1)];
          PairDev pd;
          FillOut f;
          ResultDev r;
          if (walk_prologue(A.wk, p, pd, f, r)) {
            w.begin(A.wk, p, pd, f, r);
            active = true;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // ids read before the slots are freed
        if (wl == 0) __hip_atomic_fetch_add(&q->head, take, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else if (done && idle == __ballot(true)) {
        return;  // done was set after the last push, so tail is final; no walk left
      }
    }
    if (__any(active)) {
      if (active && w.epoch(A.wk, tab, slice)) active = false;
    } else {
      // nothing to walk: a task takes the fill wave tens of microseconds or
      // more, so poll every few microseconds only (polls cost the fill wave
      // LDS and issue slots)
      __builtin_amdgcn_s_sleep(127);
    }
  }
}

// Lanes of a group.  A 16-lane group is a DPP row.  An 8-lane group takes
// every other lane of a row (two groups per row, interleaved: group lane l at
// row lane 2l or 2l + 1), so that one row_shr:2 moves every lane of both
// groups up one and leaves both groups' lane 0 to `old`.  A whole-wave group
// is the wave.
template <int GG>
__device__ __forceinline__ int group_lane(int lane) {
  static_assert(GG == 8 || GG == 16 || GG == GWAVE, "group width");
  return GG == 8 ? (lane & 15) >> 1 : lane & (GG - 1);
}
// Group of a lane, and the wave lane of lane 0 of group g.
template <int GG>
__device__ __forceinline__ int group_of(int lane) {
  return GG == 8 ? ((lane >> 4) << 1) | (lane & 1) : lane / GG;
}
template <int GG>
__device__ __forceinline__ int group_leader(int g) {
  return GG == 8 ? ((g >> 1) << 4) | (g & 1) : g * GG;
}
// Wave-lane distance between group lanes l and l ^ w.
template <int GG>
__device__ __forceinline__ int group_xor(int w) { return GG == 8 ? 2 * w : w; }

// Shift one lane up inside a group (DPP row_shr:2 / row_shr:1 / wave_shr:1);
// the group's lane 0 takes `old`.
template <int GG>
__device__ __forceinline__ int shr1(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, GG == GWAVE ? 0x138 : (GG == 16 ? 0x111 : 0x112), 0xf, 0xf, false);
}

// LDS bytes the fill kernel needs for a table of Kp codes: per lane, R byte
// scores per code and half (padded against bank conflicts) or, with 4-bit
// scores, one dword of R scores per code and half.
__host__ __device__ constexpr int fill_slot_stride(int P, int Kp, bool nib = false) {
  return nib ? 8 * Kp : P * Kp * R + 8;
}
__host__ __device__ constexpr int fill_lds_bytes(int P, int Kp, bool nib = false) {
  return 1024 + 64 * fill_slot_stride(P, Kp, nib);
}

// Packed (sext A[k], sext B[k]) from two 8-byte profile rows.  v_perm can
// sign-extend only odd bytes, so even rows are moved up one byte first.
__device__ __forceinline__ uint32_t prof2(uint32_t a, uint32_t b, uint32_t as, uint32_t bs, int k) {
  switch (k & 3) {
    case 0: return __builtin_amdgcn_perm(bs, as, 0x0A050801u);
    case 1: return __builtin_amdgcn_perm(b, a, 0x0A050801u);
    case 2: return __builtin_amdgcn_perm(bs, as, 0x0B070903u);
    default: return __builtin_amdgcn_perm(b, a, 0x0B070903u);
  }
}

// Signed byte k of the 8-byte profile row (rows 0..3 in x, 4..7 in y).
__device__ __forceinline__ int prof1(uint2 p, int k) {
  uint32_t w = k < 4 ? p.x : p.y;
  return (int)(int8_t)(w >> (8 * (k & 3)));
}

// A lane's profile slot for its R = 8 rows: slot[h][c][k] = S(row code k of
// half h, column code c), built by transposing the rows' table lines with
// v_perm (rows are read as whole 32-byte lines; c runs over 4-byte groups
// with compile-time byte positions so nothing is indexed dynamically).
template <int P>
__device__ __forceinline__ void build_slot(const int8_t *tab, int8_t *slot, int Kp, uint4 rc) {
  const uint32_t rcw[4] = {rc.x, rc.y, rc.z, rc.w};
#pragma unroll
  for (int h = 0; h < P; ++h) {
    uint4 lo[R], hi[R];  // table line of each row: bytes 0..15, 16..31
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int code = (rcw[k >> 1] >> (16 * (k & 1) + 8 * h)) & 0xff;
      lo[k] = *(const uint4 *)(tab + 32 * code);
      hi[k] = *(const uint4 *)(tab + 32 * code + 16);
    }
    int8_t *dst = slot + h * Kp * R;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      if (4 * d >= Kp) break;
      uint32_t w[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint4 &q = d < 4 ? lo[k] : hi[k];
        const int e = d & 3;
        w[k] = e == 0 ? q.x : (e == 1 ? q.y : (e == 2 ? q.z : q.w));
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int c = 4 * d + b;
        if (c >= Kp) break;
        // byte b of rows (0,1), (2,3), ... gathered into one dword per 4 rows
        const uint32_t sel = (uint32_t)b | ((uint32_t)(b + 4) << 8);
        const uint32_t x01 = __builtin_amdgcn_perm(w[1], w[0], sel);
        const uint32_t x23 = __builtin_amdgcn_perm(w[3], w[2], sel);
        const uint32_t x45 = __builtin_amdgcn_perm(w[5], w[4], sel);
        const uint32_t x67 = __builtin_amdgcn_perm(w[7], w[6], sel);
        const uint32_t a = __builtin_amdgcn_perm(x23, x01, 0x05040100u);
        const uint32_t bb = __builtin_amdgcn_perm(x67, x45, 0x05040100u);
        *(uint2 *)(dst + c * R) = make_uint2(a, bb);
      }
    }
  }
}

// Packed-couple slot: slot32[h * Kp + c] holds the biased 4-bit scores of
// the lane's rows k = 0..7 for column code c, row 2q in the low and row
// 2q + 1 in the high nibble of byte q.  Table entries are 0..15.
__device__ __forceinline__ void build_slot_nib(const int8_t *tab, uint32_t *slot32, int Kp, uint4 rc) {
  const uint32_t rcw[4] = {rc.x, rc.y, rc.z, rc.w};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint4 lo[R], hi[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int code = (rcw[k >> 1] >> (16 * (k & 1) + 8 * h)) & 0xff;
      lo[k] = *(const uint4 *)(tab + 32 * code);
      hi[k] = *(const uint4 *)(tab + 32 * code + 16);
    }
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      if (4 * d >= Kp) break;
      uint32_t w[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint4 &q = d < 4 ? lo[k] : hi[k];
        const int e = d & 3;
        w[k] = e == 0 ? q.x : (e == 1 ? q.y : (e == 2 ? q.z : q.w));
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int c = 4 * d + b;
        if (c >= Kp) break;
        const uint32_t sel = (uint32_t)b | ((uint32_t)(b + 4) << 8);
        // byte b of rows (0,2) (4,6) -> even rows; (1,3) (5,7) -> odd rows
        const uint32_t e02 = __builtin_amdgcn_perm(w[2], w[0], sel), e46 = __builtin_amdgcn_perm(w[6], w[4], sel);
        const uint32_t o13 = __builtin_amdgcn_perm(w[3], w[1], sel), o57 = __builtin_amdgcn_perm(w[7], w[5], sel);
        const uint32_t ev = __builtin_amdgcn_perm(e46, e02, 0x05040100u);
        const uint32_t od = __builtin_amdgcn_perm(o57, o13, 0x05040100u);
        slot32[h * Kp + c] = ev | (od << 4);
      }
    }
  }
}

// Progress value of strip s with `blocks` ring blocks stored (grows with s).
__device__ __forceinline__ int coop_prog(int s, int blocks) { return ((s + 1) << 20) | blocks; }

// Waits until an LDS counter reaches `target` (acquire); gives up after the
// spin limit, flagging the error, so a wait can never hang the kernel.
__device__ __forceinline__ void coop_wait(const FillArgs &A, int32_t *c, int target) {
  int spins = 0;
  while (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
    if (++spins > A.spin_limit) {
      if ((threadIdx.x & 63) == 0) atomicOr(A.err, 4);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

template <int GG, int P, bool LOCAL, bool TB, bool COOP = false>
__device__ __forceinline__ void fill_phase(const FillArgs &A, const Task *tasks, int ntasks, int32_t *ctr,
                                           const int8_t *tab, int8_t *slot, WalkQ *wq, int &wtail,
                                           CoopQ *cq = nullptr) {
  static_assert(!COOP || (GG == GWAVE && TB), "cooperative tasks are whole-wave, with walks");
  using L = Lane<P>;
  using V = typename L::V;
  constexpr int ST = GG * R;  // rows per strip
  constexpr bool NIB = P == 2 && !TB;  // 4-bit profile slot
  const int lane = threadIdx.x & 63;
  const int l = group_lane<GG>(lane);
  const int Kp = A.Kp;
  const int wv = threadIdx.x >> 6;            // fill wave of the workgroup (cooperative tasks)
  constexpr int SSTEP = COOP ? WALK_FW : 1;  // strips a wave advances by
  int cepoch = 0;                            // cooperative tasks started (every wave agrees)

  const V vopen = L::splat(A.open);
  const V vext = L::splat(A.ext);
  const V vzero = L::splat(0);
  const V vneg = L::splat(L::NEG);
  // Local fills keep E and F floored at 0: the gap terms are unsigned
  // saturating subtractions of |open| and |ext|, which gives E' = max(E, 0),
  // F' = max(F, 0) for the same H (H = max(HdS, E, F, 0) = max(HdS, E', F')),
  // so H needs no floor of its own.  The direction bits are those of the
  // exact values wherever a walk reads them: every E (F) on a local path is
  // > 0, and a difference of two values differs from the floored one's only
  // where both are <= 0.
  constexpr bool FLOOR = LOCAL;
  const V vgo = FLOOR ? L::splat(-A.open) : vopen;  // gap terms as the gap op takes them
  const V vge = FLOOR ? L::splat(-A.ext) : vext;
  const V vgap0 = FLOOR ? vzero : vneg;             // E and F where no gap can be yet
  auto gap = [](V x, V g) -> V {
    if constexpr (FLOOR) {
      typedef unsigned short u2v __attribute__((ext_vector_type(2)));
      using U = std::conditional_t<P == 2, u2v, unsigned>;
      return __builtin_bit_cast(V, __builtin_elementwise_sub_sat(__builtin_bit_cast(U, x), __builtin_bit_cast(U, g)));
    } else {
      return L::add(x, g);
    }
  };

  // Group state (identical in the lanes of a group).
  bool alive = true;
  int task = -1, s = 0, b = 0;
  int n = 0, S = 0, NB = 0, single = 0;
  int m0 = 0, m1 = 0, off0 = 0, off1 = 0, pid0 = 0, pid1 = 0;
  uint64_t row_off = 0, col_off = 0, ring_off = 0, tb_off = 0;
  int ring_len = 0, flags_or = 0;

  // DP registers.
  V H[R], E[R], HoL[R];
  V hd0 = vzero, Hbot = vzero, Fbot = vzero;
  int ccur = 0;             // residue codes of the current step (lo | hi << 8)
  uint2 pa_n = {0, 0}, pb_n = {0, 0};  // profile rows (lo / hi half) of the current step, read one step early
  uint32_t qa_n = 0, qb_n = 0;         // NIB: 4-bit profile dwords of the current step
  const uint32_t *slot32 = (const uint32_t *)slot;
  const V vbias = L::splat(A.bias);
  // Local end-cell keys (P == 2), one per pair of rows (2p, 2p+1) and half:
  // (H << 16) | (row 2p ? 0x8000 : 0) | (0x7FFF - j); the maximum is the
  // largest H, then the upper row, then the first column.
  uint32_t kp0[R / 2], kp1[R / 2];
  int rmax[R], rcol[R];       // local end-cell tracking (P == 1)
  int bestH0 = 0, besti0 = 0, bestj0 = 0, bestH1 = 0, besti1 = 0, bestj1 = 0;
  uint4 ccodes = {0, 0, 0, 0}, ccodes_n = {0, 0, 0, 0}, ccodes_nn = {0, 0, 0, 0};  // blocks b, b+1, b+2
  uint4 rc_next = {0, 0, 0, 0};  // row codes of the next strip (prefetched)
  uint4 rH0 = {0, 0, 0, 0}, rH1 = {0, 0, 0, 0}, rF0 = {0, 0, 0, 0}, rF1 = {0, 0, 0, 0};
  uint4 nH0 = rH0, nH1 = rH0, nF0 = rH0, nF1 = rH0;
  int carryH = 0, carryF = 0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    H[k] = vzero; E[k] = vgap0; HoL[k] = vzero;
    rmax[k] = 0; rcol[k] = 0;
    if (k < R / 2) { kp0[k] = 0; kp1[k] = 0; }
  }

  // A wave takes one unit of work: a whole-wave task, or NGW G-lane tasks of
  // one shape (same strips and blocks), so the strips of the groups start and
  // end together and a block's form (full / masked) is the same for the
  // whole wave.
  // Cooperative task end, any wave: its strips' best cell per half to LDS,
  // its stores completed, then its arrival.
  auto coop_arrive = [&]() {
    if constexpr (COOP) {
      if constexpr (LOCAL) {
#pragma unroll
        for (int h = 0; h < P; ++h) {
          int bh = h ? bestH1 : bestH0, bi = h ? besti1 : besti0, bj = h ? bestj1 : bestj0;
#pragma unroll
          for (int w = GG / 2; w >= 1; w >>= 1) {
            const int oh = __shfl_xor(bh, w, 64), oi = __shfl_xor(bi, w, 64), oj = __shfl_xor(bj, w, 64);
            const bool take = oh > bh || (oh == bh && (oi < bi || (oi == bi && oj < bj)));
            if (take) { bh = oh; bi = oi; bj = oj; }
          }
          if (lane == 0) {
            cq->best[wv][h][0] = bh; cq->best[wv][h][1] = bi; cq->best[wv][h][2] = bj;
          }
        }
      }
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): planes, ring and results stored
      if (lane == 0) __hip_atomic_fetch_add(&cq->arrived, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  };
  // Cooperative task end, leader (every wave has arrived): the best cell over
  // all waves' strips (H desc, i asc, j asc, as within a wave), the result,
  // and the pair handed to the leader's walker.
  auto coop_finish = [&]() {
    if constexpr (COOP) {
      if constexpr (LOCAL) {
#pragma unroll
        for (int h = 0; h < P; ++h) {
          if (h == 1 && single) continue;
          int bh = -1, bi = 0, bj = 0;
#pragma unroll
          for (int w = 0; w < WALK_FW; ++w) {
            const int oh = cq->best[w][h][0], oi = cq->best[w][h][1], oj = cq->best[w][h][2];
            const bool take = oh > bh || (oh == bh && (oi < bi || (oi == bi && oj < bj)));
            if (take) { bh = oh; bi = oi; bj = oj; }
          }
          if (lane == 0) {
            const int off = h ? off1 : off0;
            FillOut o;
            o.score = bh;
            o.end_i = bh > 0 ? bi + 1 : 0;
            o.end_j = bh > 0 ? bj - off + 1 : 0;
            o.flags = (P == 1 ? 1 : 0) | (P == 2 && bh > A.sat_limit ? 2 : 0) | (bh == 0 ? 8 : 0) | flags_or;
            A.out[h ? pid1 : pid0] = o;
          }
        }
      }
      const int sg[1] = {single}, a0[1] = {pid0}, a1[1] = {pid1};
      walk_push<1>(wq, wtail, sg, a0, a1);
    }
  };

  auto fetch = [&]() {
    int t = 0;
    if constexpr (COOP) {
      if (wv == 0) {
        if (cepoch > 0) {
          coop_wait(A, &cq->arrived, WALK_FW * cepoch);
          coop_finish();
        }
        if (lane == 0) t = atomicAdd(ctr, 1);
        t = __shfl(t, 0, 64);
        if (lane < 8) cq->prog[lane] = 0;
        if (lane == 0) cq->task = t;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(&cq->epoch, cepoch + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else {
        coop_wait(A, &cq->epoch, cepoch + 1);
        t = __hip_atomic_load(&cq->task, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      ++cepoch;
    } else {
      if (lane == 0) t = atomicAdd(ctr, 1);
      t = __shfl(t, 0, 64);
    }
    if (t >= ntasks) {
      alive = false;
      return;
    }
    if constexpr (GG != GWAVE) t = (GWAVE / GG) * t + group_of<GG>(lane);
    task = t;
    const Task &T = tasks[t];
    n = T.n; S = T.S; NB = T.NB; single = T.single;
    m0 = T.m[0]; m1 = T.m[1]; off0 = T.off[0]; off1 = T.off[1];
    pid0 = T.pair[0]; pid1 = T.pair[1];
    row_off = T.row_off; col_off = T.col_off; ring_off = T.ring_off; tb_off = T.tb_off;
    ring_len = T.ring_len;
    flags_or = T.flags_or;
    s = COOP ? wv : 0; b = 0;
    bestH0 = 0; besti0 = 0; bestj0 = 0; bestH1 = 0; besti1 = 0; bestj1 = 0;
  };

  // H(i, 0) for 1-based row i.
  auto left = [&](int i) -> int { return (LOCAL || i == 0) ? 0 : A.open + (i - 1) * A.ext; };

  fetch();

  while (__any(alive)) {
    if (!alive) continue;
    if constexpr (COOP) {
      if (s >= S) {  // this wave's strips of the task are done
        coop_arrive();
        fetch();
        continue;
      }
    }

    if (b == 0) {
      // ---- strip start: profile slot, left boundary, first block inputs.
      const uint4 rc = s < SSTEP ? *(const uint4 *)(A.rowcodes + row_off + (uint64_t)s * ST + l * R) : rc_next;
      if (s + SSTEP < S)
        rc_next = *(const uint4 *)(A.rowcodes + row_off + (uint64_t)(s + SSTEP) * ST + l * R);
      if constexpr (NIB) build_slot_nib(tab, (uint32_t *)slot, Kp, rc);
      else build_slot<P>(tab, slot, Kp, rc);
      const int row0 = s * ST + l * R;  // 0-based first row of this lane
#pragma unroll
      for (int k = 0; k < R; ++k) {
        H[k] = L::splat(left(row0 + k + 1));
        E[k] = vgap0;
        HoL[k] = gap(H[k], vgo);
        rmax[k] = 0; rcol[k] = 0;
        if (k < R / 2) { kp0[k] = 0; kp1[k] = 0; }
      }
      hd0 = L::splat(left(row0));
      ccodes_n = *(const uint4 *)(A.colcodes + col_off);
      ccodes_nn = *(const uint4 *)(A.colcodes + col_off + 8);
      // Residue of step 0: lane 0 takes column 0, the others a pad code
      // (they are idle until their first column arrives).
      ccur = shr1<GG>((int)(ccodes_n.x & 0xffff), (Kp - 1) | ((Kp - 1) << 8));
      if constexpr (NIB) {
        qa_n = slot32[ccur & 0xff];
        qb_n = slot32[Kp + ((ccur >> 8) & 0xff)];
      } else {
        pa_n = *(const uint2 *)(slot + (ccur & 0xff) * R);
        if constexpr (P == 2) pb_n = *(const uint2 *)(slot + Kp * R + ((ccur >> 8) & 0xff) * R);
      }
      if (s > 0) {
        // ring block 0 = the writer's block GG / 8
        if constexpr (COOP) coop_wait(A, &cq->prog[(s - 1) & 7], coop_prog(s - 1, min(GG / 8 + 1, NB)));
        const uint32_t *rh = A.ring + ring_off + GG;
        const uint32_t *rf = rh + ring_len;
        nH0 = *(const uint4 *)rh; nH1 = *(const uint4 *)(rh + 4);
        nF0 = *(const uint4 *)rf; nF1 = *(const uint4 *)(rf + 4);
      }
    }

    // ---- one block of 8 steps.
    // Residues are read two blocks ahead (block NB is the column slack), the
    // ring one block ahead.
    ccodes = ccodes_n;
    ccodes_n = ccodes_nn;
    if (b + 2 <= NB) ccodes_nn = *(const uint4 *)(A.colcodes + col_off + 8 * (b + 2));
    rH0 = nH0; rH1 = nH1; rF0 = nF0; rF1 = nF1;
    if (b + 1 < NB) {  // prefetch the next block's lane-0 ring inputs
      if (s > 0) {
        if constexpr (COOP) coop_wait(A, &cq->prog[(s - 1) & 7], coop_prog(s - 1, min(b + 1 + GG / 8 + 1, NB)));
        const uint32_t *rh = A.ring + ring_off + GG + 8 * (b + 1);
        const uint32_t *rf = rh + ring_len;
        nH0 = *(const uint4 *)rh; nH1 = *(const uint4 *)(rh + 4);
        nF0 = *(const uint4 *)rf; nF1 = *(const uint4 *)(rf + 4);
      }
    }
    const uint32_t ccw[4] = {ccodes.x, ccodes.y, ccodes.z, ccodes.w};
    // Lane 0's top row of each step: the previous strip's bottom row from the
    // ring or, in the first strip, the boundary.  One branch per block: s is
    // the same in every lane of a wave (a whole-wave task, or a unit of
    // tasks of one shape).
    uint32_t rhw[8], rfw[8];
    if (__builtin_amdgcn_readfirstlane(s) == 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        rhw[q] = (uint32_t)L::bits(L::splat(LOCAL ? 0 : A.open + (8 * b + q) * A.ext));
        rfw[q] = (uint32_t)L::bits(vgap0);
      }
    } else {
      rhw[0] = rH0.x; rhw[1] = rH0.y; rhw[2] = rH0.z; rhw[3] = rH0.w;
      rhw[4] = rH1.x; rhw[5] = rH1.y; rhw[6] = rH1.z; rhw[7] = rH1.w;
      rfw[0] = rF0.x; rfw[1] = rF0.y; rfw[2] = rF0.z; rfw[3] = rF0.w;
      rfw[4] = rF1.x; rfw[5] = rF1.y; rfw[6] = rF1.z; rfw[7] = rF1.w;
    }
    uint32_t tbw[16];
    int rbH[8], rbF[8];

    // The 8 steps, in two versions: with the per-step column test, and
    // without it for blocks where every lane of the wave stays inside its
    // task's columns (most of a strip), which also spares the merges of the
    // registers a masked step leaves alone.
    auto steps = [&](auto full_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int t = 8 * b + q;
        const int j = t - l;
        // This step's residues and profile rows were fetched last step; start
        // the LDS reads of the next step now (lane l's next residue is lane
        // l-1's current one).
        const uint2 pa = pa_n, pb = pb_n;
        const uint32_t qa = qa_n, qb = qb_n;
        const int cthis = ccur;
        // lane 0's residue at step t + 1; at q = 7 the first of the next block,
        // read here, a block after its load was issued (reading it at the
        // block start would wait for the load, and for the stores before it)
        const int code1 = q < 7 ? (int)((ccw[(q + 1) >> 1] >> (16 * ((q + 1) & 1))) & 0xffff)
                                : (int)(ccodes_n.x & 0xffff);
        ccur = shr1<GG>(code1, cthis);
        if constexpr (NIB) {
          qa_n = slot32[ccur & 0xff];
          qb_n = slot32[Kp + ((ccur >> 8) & 0xff)];
        } else {
          pa_n = *(const uint2 *)(slot + (ccur & 0xff) * R);
          if constexpr (P == 2) pb_n = *(const uint2 *)(slot + Kp * R + ((ccur >> 8) & 0xff) * R);
        }
        const int topH = (int)rhw[q], topF = (int)rfw[q];
        const V Hup = L::unbits(shr1<GG>(topH, L::bits(Hbot)));
        const V Fup = L::unbits(shr1<GG>(topF, L::bits(Fbot)));

        if (FULL || (j >= 0 && j < n)) {
          // 4-bit scores: even rows in the low nibbles, odd rows in the high ones.
          const uint32_t aev = qa & 0x0F0F0F0Fu, aod = (qa >> 4) & 0x0F0F0F0Fu;
          const uint32_t bev = qb & 0x0F0F0F0Fu, bod = (qb >> 4) & 0x0F0F0F0Fu;
          // byte scores: copies moved up one byte for the even rows' sign extension
          const uint32_t pas_x = pa.x << 8, pas_y = pa.y << 8, pbs_x = pb.x << 8, pbs_y = pb.y << 8;
          const uint32_t cje = 0xffffu - (uint32_t)j, cjo = 0x7fffu - (uint32_t)j;
          uint32_t ke0 = 0, ke1 = 0;

          V hd = hd0;
          V HoU = gap(Hup, vgo);
          V F = Fup;
          uint32_t xc = 0, yc = 0, acc = 0;
#pragma unroll
          for (int k = 0; k < R; ++k) {
            V Sk;
            if constexpr (NIB) {
              // byte k/2 of the lo / hi half's nibbles -> (lo16, hi16), unbiased
              const uint32_t q8 = (uint32_t)(k >> 1);
              const uint32_t sb = __builtin_amdgcn_perm((k & 1) ? bod : bev, (k & 1) ? aod : aev,
                                                        q8 | (0x0Cu << 8) | ((4u + q8) << 16) | (0x0Cu << 24));
              Sk = L::sub(s2(sb), vbias);
            } else if constexpr (P == 2) {
              const bool hiw = k >= 4;
              Sk = s2(prof2(hiw ? pa.y : pa.x, hiw ? pb.y : pb.x, hiw ? pas_y : pas_x, hiw ? pbs_y : pbs_x, k));
            } else {
              Sk = L::unbits(prof1(pa, k));
            }
            const V HdS = L::add(hd, Sk);
            const V Eext = gap(E[k], vge);
            const V En = L::vmax(HoL[k], Eext);
            const V Fext = gap(F, vge);
            const V Fn = L::vmax(HoU, Fext);
            const V EF = L::vmax(En, Fn);
            V Hn = L::vmax(HdS, EF);
            if constexpr (LOCAL && !FLOOR) Hn = L::vmax(Hn, vzero);
            const V Hon = gap(Hn, vgo);
            if constexpr (TB) {
              // Signs of: H+open - E (e: E extended), H+open - F (f), the
              // diagonal - H (d: H not from the diagonal), E - F (x: F above
              // E).  Taken against the maxima rather than the candidates so
              // that only values still live are read.  Where local H is 0
              // (the floor won) d can differ from the candidates' sign; a
              // walk stops on entering such a cell and never reads it.
              const uint32_t dE = (uint32_t)L::bits(L::sub(HoL[k], En));
              const uint32_t dF = (uint32_t)L::bits(L::sub(HoU, Fn));
              const uint32_t dD = (uint32_t)L::bits(L::sub(HdS, Hn));
              const uint32_t dX = (uint32_t)L::bits(L::sub(En, Fn));
              if constexpr (P == 2) {
                // sign bytes (0x00 / 0xFF) of [e lo, e hi, f lo, f hi] and
                // [d lo, d hi, x lo, x hi]; row k keeps bit k of each byte
                const uint32_t X = __builtin_amdgcn_perm(dF, dE, 0x0B0A0908u);
                const uint32_t Y = __builtin_amdgcn_perm(dX, dD, 0x0B0A0908u);
                // acc | (sign bytes & bit k of each byte): one v_bitop3 each
                // (truth table 0xEA = (a & b) | c), a chain the compiler would
                // otherwise turn into a wider AND / OR3 tree
                const uint32_t mk = 0x01010101u << k;
                xc = __builtin_amdgcn_bitop3_b32(X, mk, xc, 0xEA);
                yc = __builtin_amdgcn_bitop3_b32(Y, mk, yc, 0xEA);
              } else {
                acc = __builtin_amdgcn_alignbit(acc, dE, 31);
                acc = __builtin_amdgcn_alignbit(acc, dF, 31);
                acc = __builtin_amdgcn_alignbit(acc, dD, 31);
                acc = __builtin_amdgcn_alignbit(acc, dX, 31);
              }
            }
            if constexpr (LOCAL) {
              if constexpr (P == 2) {
                const uint32_t hu = u32(Hn);
                const uint32_t c = (k & 1) ? cjo : cje;
                const uint32_t k0 = (hu << 16) | c, k1 = (hu & 0xffff0000u) | c;
                if (k & 1) {
                  kp0[k >> 1] = max(max(kp0[k >> 1], ke0), k0);
                  kp1[k >> 1] = max(max(kp1[k >> 1], ke1), k1);
                } else {
                  ke0 = k0;
                  ke1 = k1;
                }
              } else {
                const int hv = L::bits(Hn);
                if (hv > rmax[k]) { rmax[k] = hv; rcol[k] = j; }
              }
            }
            hd = H[k];
            H[k] = Hn;
            E[k] = En;
            HoL[k] = Hon;
            F = Fn;
            HoU = Hon;
          }
          Hbot = H[R - 1];
          Fbot = F;
          hd0 = Hup;
          if constexpr (TB) {
            if constexpr (P == 2) { tbw[2 * q] = xc; tbw[2 * q + 1] = yc; }
            else tbw[q] = acc;
          }
        }
        rbH[q] = L::bits(Hbot);
        rbF[q] = L::bits(Fbot);
      }
    };
    const bool inside = 8 * b >= GG - 1 && 8 * b + 8 <= n;
    if (__all(inside)) steps(std::true_type{});
    else steps(std::false_type{});

    // ---- block outputs: direction planes (all lanes), ring (the group's
    // last lane); a filler group writes nothing (it shares the offsets of a
    // task).
    if (TB && single != 2) {
      if constexpr (P == 2) {
        uint4 *dst = (uint4 *)(A.tb + tb_off + ((((uint64_t)s * NB + b) * GG + l) * 64));
        dst[0] = make_uint4(tbw[0], tbw[1], tbw[2], tbw[3]);
        dst[1] = make_uint4(tbw[4], tbw[5], tbw[6], tbw[7]);
        dst[2] = make_uint4(tbw[8], tbw[9], tbw[10], tbw[11]);
        dst[3] = make_uint4(tbw[12], tbw[13], tbw[14], tbw[15]);
      } else {
        uint4 *dst = (uint4 *)(A.tb + tb_off + ((((uint64_t)s * NB + b) * GG + l) * 32));
        dst[0] = make_uint4(tbw[0], tbw[1], tbw[2], tbw[3]);
        dst[1] = make_uint4(tbw[4], tbw[5], tbw[6], tbw[7]);
      }
    }
    if (l == GG - 1 && s + 1 < S && single != 2) {
      // The last lane handles column t - (GG-1) at step t; ring index =
      // column + GG, so block b carries indices 8b .. 8b+7: the previous
      // block's q = 7 then q 0..6.
      uint32_t *rh = A.ring + ring_off + 8 * b;
      uint32_t *rf = rh + ring_len;
      *(uint4 *)rh = make_uint4(carryH, rbH[0], rbH[1], rbH[2]);
      *(uint4 *)(rh + 4) = make_uint4(rbH[3], rbH[4], rbH[5], rbH[6]);
      *(uint4 *)rf = make_uint4(carryF, rbF[0], rbF[1], rbF[2]);
      *(uint4 *)(rf + 4) = make_uint4(rbF[3], rbF[4], rbF[5], rbF[6]);
      if constexpr (COOP)  // the next strip (another wave) may read this block
        __hip_atomic_store(&cq->prog[s & 7], coop_prog(s, b + 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    carryH = rbH[7];
    carryF = rbF[7];

    ++b;
    if (b < NB) continue;

    // ---- strip end.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int row0 = s * ST + l * R;
    if constexpr (LOCAL) {
      if constexpr (P == 2) {
        // A pad row (>= m) never holds more than the running maximum of the
        // real row above it, so a pair's key decodes to a real row.
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
          const uint32_t a = kp0[q], c = kp1[q];
          const int ia = row0 + 2 * q + ((a & 0x8000u) ? 0 : 1), ic = row0 + 2 * q + ((c & 0x8000u) ? 0 : 1);
          const int h0 = (int)(a >> 16), h1 = (int)(c >> 16);
          if (ia < m0 && h0 > bestH0) { bestH0 = h0; besti0 = ia; bestj0 = 0x7fff - (int)(a & 0x7fffu); }
          if (ic < m1 && h1 > bestH1) { bestH1 = h1; besti1 = ic; bestj1 = 0x7fff - (int)(c & 0x7fffu); }
        }
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int i = row0 + k;  // 0-based row
          if (i < m0 && rmax[k] > bestH0) { bestH0 = rmax[k]; besti0 = i; bestj0 = rcol[k]; }
        }
      }
    } else {
      // Global: H(m, n) sits in the lane owning row m at the last column.
#pragma unroll
      for (int h = 0; h < P; ++h) {
        const int mh = h ? m1 : m0;
        const int r = mh - 1 - row0;
        if (mh > 0 && r >= 0 && r < R && single != 2 && !(h == 1 && single)) {
          int v = 0;
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (k == r) v = L::bits(H[k]);
          int score = P == 2 ? (int)(int16_t)(h ? ((uint32_t)v >> 16) : ((uint32_t)v & 0xffff)) : v;
          FillOut o;
          o.score = score; o.end_i = mh; o.end_j = n; o.flags = (P == 1 ? 1 : 0) | flags_or;
          A.out[h ? pid1 : pid0] = o;
        }
      }
    }
    s += SSTEP;
    b = 0;
    carryH = 0;
    carryF = 0;
    if (COOP || s < S) continue;

    // ---- task end (local): reduce the group's best cells per half.
    if constexpr (LOCAL) {
#pragma unroll
      for (int h = 0; h < P; ++h) {
        int bh = h ? bestH1 : bestH0, bi = h ? besti1 : besti0, bj = h ? bestj1 : bestj0;
#pragma unroll
        for (int w = GG / 2; w >= 1; w >>= 1) {
          const int x = group_xor<GG>(w);
          const int oh = __shfl_xor(bh, x, 64), oi = __shfl_xor(bi, x, 64), oj = __shfl_xor(bj, x, 64);
          const bool take = oh > bh || (oh == bh && (oi < bi || (oi == bi && oj < bj)));
          if (take) { bh = oh; bi = oi; bj = oj; }
        }
        if (l == 0 && single != 2 && !(h == 1 && single)) {
          const int off = h ? off1 : off0;
          FillOut o;
          o.score = bh;
          o.end_i = bh > 0 ? bi + 1 : 0;
          o.end_j = bh > 0 ? bj - off + 1 : 0;
          o.flags = (P == 1 ? 1 : 0) | (P == 2 && bh > A.sat_limit ? 2 : 0) | (bh == 0 ? 8 : 0) | flags_or;
          A.out[h ? pid1 : pid0] = o;
        }
      }
    }
    if constexpr (TB) {
      if (A.walk) {  // hand the task's pairs to the walker wave
        constexpr int NG = GWAVE / GG;
        int sg[NG], a0[NG], a1[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          sg[g] = NG == 1 ? single : __shfl(single, group_leader<GG>(g), 64);
          a0[g] = NG == 1 ? pid0 : __shfl(pid0, group_leader<GG>(g), 64);
          a1[g] = NG == 1 ? pid1 : __shfl(pid1, group_leader<GG>(g), 64);
        }
        walk_push<NG>(wq, wtail, sg, a0, a1);
      }
    }
    fetch();
  }
}

// Workgroup: one fill wave or, with A.walk, WALK_FW fill waves (waves
// 0..WALK_FW-1) and WALK_FW walker waves (wave WALK_FW + f serves fill wave
// f with its first WALK_LANES lanes).  Waves of a workgroup are dealt to the SIMDs in
// turn, so the fill waves land on distinct SIMDs, each beside one walker.
template <int P, bool LOCAL, bool TB>
__global__ __launch_bounds__(TB ? 128 * WALK_FW : 64) void fill_kernel(FillArgs A) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Substitution table in LDS with 32-byte rows (row r at tab + 32 r).
  int8_t *tab = (int8_t *)lds;
  for (int x = threadIdx.x; x < 32 * 32; x += blockDim.x) {
    const int r = x >> 5, c = x & 31;
    tab[x] = (r < A.Kp && c < A.Kp) ? A.table[r * A.Kp + c] : (int8_t)0;
  }
  const int stride = fill_slot_stride(P, A.Kp, P == 2 && !TB);
  WalkQ *wq = nullptr;
  uint32_t *wlds = nullptr;
  CoopQ *cq = nullptr;
  if constexpr (TB) {
    if (A.walk) {
      const int f = wave & (WALK_FW - 1);
      const int base = 1024 + WALK_FW * 64 * stride;
      wlds = (uint32_t *)(lds + base) + f * WALK_LANES * WSTRIDE;
      WalkQ *wq0 = (WalkQ *)(lds + base + WALK_FW * WALK_LANES * WSTRIDE * 4);
      wq = wq0 + f;
      cq = (CoopQ *)(wq0 + WALK_FW);
      if (wave < WALK_FW && lane == 0) { wq->tail = 0; wq->head = 0; wq->done = 0; wq->claim = 0; }
      if (threadIdx.x == 0) { cq->task = 0; cq->epoch = 0; cq->arrived = 0; }
    }
  }
  __syncthreads();
  if constexpr (TB) {
    if (wave >= WALK_FW) {
      if (lane < WALK_LANES) walk_loop<P>(A, wq, tab, wlds);
      return;
    }
  }
  int8_t *slot = (int8_t *)lds + 1024 + threadIdx.x * stride;
  int wtail = 0;
  // Cooperative tasks first (a workgroup each), then long tasks (a wave
  // each), then G-lane groups, then 8-lane groups take the rest.
  int ncoop = 0;
  if constexpr (TB) {
    if (A.walk && A.ncoop) {
      ncoop = A.ncoop;
      fill_phase<GWAVE, P, LOCAL, TB, true>(A, A.tasks[0], ncoop, A.coop_ctr, tab, slot, wq, wtail, cq);
    }
  }
  if (A.ntasks[0] > ncoop)
    fill_phase<GWAVE, P, LOCAL, TB>(A, A.tasks[0] + ncoop, A.ntasks[0] - ncoop, A.ctr, tab, slot, wq, wtail);
  if (A.ntasks[1])
    fill_phase<G, P, LOCAL, TB>(A, A.tasks[1], A.ntasks[1], A.ctr + 1, tab, slot, wq, wtail);
  if constexpr (P == 2) {
    if (A.ntasks[2]) fill_phase<8, P, LOCAL, TB>(A, A.tasks[2], A.ntasks[2], A.ctr8, tab, slot, wq, wtail);
  }
  if constexpr (TB) {
    if (A.walk) {
      if (lane == 0) __hip_atomic_store(&wq->done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      // then walk beside the walker wave, in the LDS of this wave's profile slots
      const int nl = min(64, 64 * stride / (WSTRIDE * 4));
      if (lane < nl) walk_loop<P>(A, wq, tab, (uint32_t *)(lds + 1024 + wave * 64 * stride));
    }
  }
}

}  // namespace ta
