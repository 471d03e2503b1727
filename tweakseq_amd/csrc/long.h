// Long pairs (tens of thousands of residues per side), global alignment, int32.
//
// Forward pass (long_fwd_kernel): a strip is 512 query rows (64 lanes x R
// rows, the fill kernel's whole-wave layout) across every column, and the
// strips of all long pairs form one device work queue, pair-major.  A wave
// takes the next strip; strip s + 1 of a pair trails strip s along the
// anti-diagonal wavefront on whichever CU it landed, reading s's bottom row
// from a two-slot ring once s's progress counter covers the columns.  Every
// strip a wave waits on was dequeued before its own, by a wave that is
// running, so the pipeline cannot deadlock.  Ring rows and counters cross
// CUs (and the XCDs' separate L2s) with device-coherent (sc1) accesses and
// an explicit wait for completion before a counter moves; no cache
// write-back or invalidate is involved.
//
// The forward pass is score only.  When a traceback follows it also keeps
// checkpoints: the bottom row (H, F) of every KC-th strip, and the column
// (H, E of every row) at every CKX-th column.
//
// Traceback (long_tb_kernel, one workgroup of LW waves per pair) walks the
// matrix from (m, n) in regions.  A region is the rows of one block of KC
// strips up to the walk's row and a band of columns ending at the walk's
// column and starting at a column checkpoint far enough left for a path
// near the diagonal to leave through the block's top: its top boundary is a
// row checkpoint, its left boundary a column checkpoint.  The region is
// recomputed with direction planes (its strips pipelined over the
// workgroup's waves through a private ring and LDS counters), one thread
// walks it until the path leaves through the top or the left, and the next
// region starts there.  The recomputed cells are the forward pass's cells,
// so the walk is the full-matrix walk of the specification (oracle/gotoh.c),
// bit for bit; a region is at most LREG columns wide, so the traceback costs
// about (KC strips x LREG) cells per block of rows instead of the matrix.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "fill.h"
#include "traceback.h"

namespace ta {

constexpr int LW = 8;           // waves of a traceback workgroup
constexpr int KC = LW;          // strips per checkpoint block (one per wave)
constexpr int LST = GWAVE * R;  // query rows per strip
constexpr int LPUB = 8;         // progress is published every LPUB blocks (64 columns)
constexpr int LMAXS = 2047;     // strips per pair the counter tags can name
constexpr int CKX = 1024;       // column checkpoint spacing
constexpr int LMARGIN = 256;    // region columns beyond the diagonal through the walk's rows
constexpr int LREG = KC * LST + LMARGIN + CKX;  // widest traceback region

struct LongPair {
  int32_t pid, m, n, S;
  uint64_t a_off, b_off;  // residue codes (16-byte aligned, zero padded)
  uint64_t ring_off;      // int32 elements: 2 slots x (H, F) x rl
  uint64_t ck_off;        // int32 elements: (S - 1) / KC rows x (H, F) x rl
  uint64_t colck_off;     // int2 elements: (n - 1) / CKX columns x S * LST rows of (H, E)
  int32_t prog_off;       // first of the pair's S progress counters
  int32_t rl;             // elements per row array (column index, + slack)
};

struct LongItem {
  int32_t lp, s;  // long pair, strip
};

struct LongArgs {
  const LongPair *lp;
  int32_t nlp;
  const LongItem *items;
  int32_t nitems;
  int32_t *ctr;         // [0] forward items taken, [1] traceback pairs taken
  const int8_t *table;  // Kp x Kp
  int32_t Kp;
  const uint8_t *codes;
  int32_t *ring, *ckpt, *prog;
  int2 *colck;
  int32_t *wring;       // traceback workgroups: private ring, wring_stride elements each
  uint64_t wring_stride;
  uint8_t *wtb;         // traceback workgroups: direction planes, wtb_stride bytes each
  uint64_t wtb_stride;
  const PairDev *pd;
  FillOut *fo;
  ResultDev *res;
  char *cig;
  int32_t open, ext;
  int32_t ckpt_on;      // keep checkpoints (a traceback follows)
  int32_t spin_limit;   // polls before a wait gives up
  int32_t *err;         // set when a wait gave up (results are then invalid)
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
template <int AUX>
__device__ __forceinline__ uint4 bload4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX));
}
template <int AUX>
__device__ __forceinline__ void bstore4(__amdgpu_buffer_rsrc_t r, uint32_t off, uint4 v) {
  typedef unsigned u4v __attribute__((ext_vector_type(4)));
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4v, v), r, off, 0, AUX);
}
__device__ __forceinline__ uint2 bload2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
// s_waitcnt vmcnt(0): every vector memory access of the wave has completed.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Progress of a strip: ((strip + 1) << 20) | columns whose bottom-row values
// are stored.  A counter slot reused by a later strip only grows.
__device__ __forceinline__ int prog_value(int s, int cols) { return ((s + 1) << 20) | cols; }

// Strips on any CU: one counter per strip in global memory, ring data
// device-coherent.
struct AgentSync {
  static constexpr int AUX = 0x10;  // sc1
  int *above, *mine;
  __device__ int peek() const { return __hip_atomic_load(above, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
  __device__ void publish(int v) const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    wait_vmem();
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(mine, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};

// Strips of one workgroup: counters in LDS (LW slots, strip s in s % LW).
struct GroupSync {
  static constexpr int AUX = 0;
  int *above, *mine;
  __device__ int peek() const { return __hip_atomic_load(above, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
  __device__ void publish(int v) const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    wait_vmem();
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(mine, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
};

enum { TOP_ROW0 = 0, TOP_RING = 1, TOP_CKPT = 2 };

// What one strip sweep covers and where its boundaries come from.  Columns
// are c_left + 1 .. c_left + ncols (1-based); "local" column = column - c_left.
struct Region {
  int s, c_left, ncols;
  int top_kind;
  const int32_t *top;  // ring (local columns) or row checkpoint (absolute columns)
  const int2 *left;    // H, E of column c_left for every row; null: matrix column 0
  int32_t *out, *ck;   // bottom row -> ring (local), row checkpoint (absolute)
  int2 *colck;         // column checkpoints to store (forward pass, c_left == 0)
  uint8_t *tbs;        // direction planes (local columns)
  int NBc;
  bool score;
};

__device__ __forceinline__ uint32_t round16(int x) { return (uint32_t)((x + 15) & ~15); }

// One strip (rows s*LST .. s*LST+511) over the region's columns, swept by
// the calling wave.  Top input: row 0, the strip above through the ring
// (waits on its counter), or a complete checkpoint row; left input: column 0
// or a column checkpoint.  Bottom row goes to the ring and the row
// checkpoint when given, direction planes to `tbs` when TB.
template <class Sync, bool TB>
__device__ void long_strip(const LongArgs &A, const Sync &y, const LongPair &P, const Region &G,
                           const int8_t *tab, int8_t *slot, bool &gave_up) {
  constexpr int GG = GWAVE;
  const int l = threadIdx.x & 63;
  const int Kp = A.Kp, open = A.open, ext = A.ext;
  const int rl = P.rl;
  const int s = G.s, ncols = G.ncols, c_left = G.c_left, top_kind = G.top_kind;
  int32_t *const out = G.out, *const ck = G.ck;
  const auto ra = brsrc(A.codes + P.a_off, round16(P.m));
  const auto rb = brsrc(A.codes + P.b_off + c_left, round16(P.n) - (uint32_t)c_left);
  // A checkpoint row is read from column c_left on: same offsets as a ring.
  const int32_t *top = top_kind == TOP_CKPT ? G.top + c_left : G.top;
  const auto rt = brsrc(top, top ? 4u * (uint32_t)(2 * rl - (top_kind == TOP_CKPT ? c_left : 0)) : 0u);
  const auto ro = brsrc(out, out ? 8u * (uint32_t)rl : 0u);
  const auto rc = brsrc(ck, ck ? 8u * (uint32_t)rl : 0u);
  const int row0 = s * LST + l * R;
  const uint64_t rows_pad = (uint64_t)P.S * LST;

  {  // profile slot of the lane's rows (codes past m read as 0; those rows feed nothing)
    const uint2 c8 = bload2(ra, (uint32_t)row0);
    uint4 w;
    w.x = (c8.x & 0xffu) | (((c8.x >> 8) & 0xffu) << 16);
    w.y = ((c8.x >> 16) & 0xffu) | ((c8.x >> 24) << 16);
    w.z = (c8.y & 0xffu) | (((c8.y >> 8) & 0xffu) << 16);
    w.w = ((c8.y >> 16) & 0xffu) | ((c8.y >> 24) << 16);
    build_slot<1>(tab, slot, Kp, w);
  }
  auto left = [&](int i) -> int { return i == 0 ? 0 : open + (i - 1) * ext; };
  int H[R], E[R], HoL[R];
  int hd0;
  if (G.left) {  // column checkpoint: rows row0 .. row0 + 7 (0-based) at column c_left
    const int4 *lp = (const int4 *)(G.left + row0);
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      const int4 v = lp[q];
      H[2 * q] = v.x; E[2 * q] = v.y; H[2 * q + 1] = v.z; E[2 * q + 1] = v.w;
    }
    hd0 = row0 == 0 ? open + (c_left - 1) * ext : G.left[row0 - 1].x;
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      H[k] = left(row0 + k + 1);
      E[k] = NEG32;
    }
    hd0 = left(row0);
  }
#pragma unroll
  for (int k = 0; k < R; ++k) HoL[k] = H[k] + open;
  int Hbot = 0, Fbot = 0, carryH = 0, carryF = 0;
  const int NB = (ncols + GG + 7) / 8;

  // Columns of the top row known to be stored.
  int avail = top_kind == TOP_RING ? 0 : ncols;
  const int tag = s;  // counter tag of strip s - 1
  auto ensure = [&](int cols) {
    if (cols > ncols) cols = ncols;
    if (avail >= cols || gave_up) return;
    int spins = 0;
    for (;;) {
      const int v = y.peek();
      const int vt = v >> 20;
      avail = vt > tag ? ncols : (vt == tag ? (v & 0xFFFFF) : 0);
      if (avail >= cols) break;
      if (++spins > A.spin_limit) {
        gave_up = true;
        if (l == 0) atomicOr(A.err, 1);
        avail = ncols;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  };
  auto load_top = [&](int blk, uint4 &h0, uint4 &h1, uint4 &f0, uint4 &f1) {
    if (top_kind == TOP_ROW0) return;
    ensure(8 * (blk + 1));
    const uint32_t o = 32u * (uint32_t)blk, of = 4u * (uint32_t)rl + o;
    if (top_kind == TOP_RING) {
      h0 = bload4<Sync::AUX>(rt, o); h1 = bload4<Sync::AUX>(rt, o + 16);
      f0 = bload4<Sync::AUX>(rt, of); f1 = bload4<Sync::AUX>(rt, of + 16);
    } else {
      h0 = bload4<0>(rt, o); h1 = bload4<0>(rt, o + 16);
      f0 = bload4<0>(rt, of); f1 = bload4<0>(rt, of + 16);
    }
  };

  uint4 tH0 = {0, 0, 0, 0}, tH1 = tH0, tF0 = tH0, tF1 = tH0;
  uint4 nH0 = tH0, nH1 = tH0, nF0 = tH0, nF1 = tH0;
  uint2 cc = bload2(rb, 0), ncc = cc;
  load_top(0, nH0, nH1, nF0, nF1);
  // Residue of step 0: lane 0 takes column 0, the others a pad code.
  int ccur = shr1<GG>((int)(cc.x & 0xff), Kp - 1);
  uint2 pa_n = *(const uint2 *)(slot + ccur * R);

  for (int blk = 0; blk < NB; ++blk) {
    cc = ncc;
    tH0 = nH0; tH1 = nH1; tF0 = nF0; tF1 = nF1;
    if (blk + 1 < NB) {
      ncc = bload2(rb, 8u * (uint32_t)(blk + 1));
      load_top(blk + 1, nH0, nH1, nF0, nF1);
    }
    const uint32_t thw[8] = {tH0.x, tH0.y, tH0.z, tH0.w, tH1.x, tH1.y, tH1.z, tH1.w};
    const uint32_t tfw[8] = {tF0.x, tF0.y, tF0.z, tF0.w, tF1.x, tF1.y, tF1.z, tF1.w};
    uint32_t tbw[8];
    int rbH[8], rbF[8];
    // Does a lane reach a checkpoint column (j + 1 a multiple of CKX) here?
    const int hi = 8 * blk + 8;
    const bool ckb = G.colck && hi >= CKX && (hi / CKX) * CKX >= 8 * blk - 62;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int t = 8 * blk + q;
      const int j = t - l;
      const uint2 pa = pa_n;
      const int code1 = q < 7 ? (int)(((q + 1 < 4 ? cc.x : cc.y) >> (8 * ((q + 1) & 3))) & 0xff)
                              : (int)(ncc.x & 0xff);
      ccur = shr1<GG>(code1, ccur);
      pa_n = *(const uint2 *)(slot + ccur * R);
      int topH, topF;
      if (top_kind == TOP_ROW0) {
        topH = open + (c_left + t) * ext;
        topF = NEG32;
      } else {
        topH = (int)thw[q];
        topF = (int)tfw[q];
      }
      const int Hup = shr1<GG>(topH, Hbot);
      const int Fup = shr1<GG>(topF, Fbot);
      tbw[q] = 0;
      if (j >= 0 && j < ncols) {
        int hd = hd0;
        int HoU = Hup + open;
        int F = Fup;
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int HdS = hd + prof1(pa, k);
          const int Eext = E[k] + ext;
          const int En = max(HoL[k], Eext);
          const int Fext = F + ext;
          const int Fn = max(HoU, Fext);
          const int EF = max(En, Fn);
          const int Hn = max(HdS, EF);
          const int Hon = Hn + open;
          if constexpr (TB) {
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(HoL[k] - Eext), 31);
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(HoU - Fext), 31);
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(HdS - EF), 31);
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(En - Fn), 31);
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
        tbw[q] = acc;
        if (ckb && ((j + 1) & (CKX - 1)) == 0 && j + 1 < P.n) {
          int4 *d = (int4 *)(G.colck + (uint64_t)((j + 1) / CKX - 1) * rows_pad + row0);
          d[0] = make_int4(H[0], E[0], H[1], E[1]);
          d[1] = make_int4(H[2], E[2], H[3], E[3]);
          d[2] = make_int4(H[4], E[4], H[5], E[5]);
          d[3] = make_int4(H[6], E[6], H[7], E[7]);
        }
      }
      rbH[q] = Hbot;
      rbF[q] = Fbot;
    }
    if constexpr (TB) {
      uint4 *dst = (uint4 *)(G.tbs + (((uint64_t)l * G.NBc + blk) * 32));
      dst[0] = make_uint4(tbw[0], tbw[1], tbw[2], tbw[3]);
      dst[1] = make_uint4(tbw[4], tbw[5], tbw[6], tbw[7]);
    }
    // The last lane finishes column t - 63 at step t: this block completes
    // columns c0 .. c0 + 7 (the previous block's q = 7, then q = 0 .. 6).
    const int c0 = 8 * blk - GG;
    if (l == GG - 1 && c0 >= 0 && (out || ck)) {
      const uint4 h0 = make_uint4(carryH, rbH[0], rbH[1], rbH[2]), h1 = make_uint4(rbH[3], rbH[4], rbH[5], rbH[6]);
      const uint4 f0 = make_uint4(carryF, rbF[0], rbF[1], rbF[2]), f1 = make_uint4(rbF[3], rbF[4], rbF[5], rbF[6]);
      const uint32_t o = 4u * (uint32_t)c0, of = 4u * (uint32_t)rl + o;
      if (out) {
        bstore4<Sync::AUX>(ro, o, h0); bstore4<Sync::AUX>(ro, o + 16, h1);
        bstore4<Sync::AUX>(ro, of, f0); bstore4<Sync::AUX>(ro, of + 16, f1);
      }
      if (ck) {
        bstore4<0>(rc, o, h0); bstore4<0>(rc, o + 16, h1);
        bstore4<0>(rc, of, f0); bstore4<0>(rc, of + 16, f1);
      }
    }
    carryH = rbH[7];
    carryF = rbF[7];
    if (out && (blk + 1) % LPUB == 0 && blk + 1 < NB) {
      const int done = 8 * (blk + 1) - GG;
      if (done > 0) y.publish(prog_value(s, done < ncols ? done : ncols));
    }
  }
  if (out) y.publish(prog_value(s, ncols));

  if (G.score) {  // global score H(m, n): the lane holding row m - 1, last column
    const int r = P.m - 1 - row0;
    if (r >= 0 && r < R) {
      int v = 0;
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (k == r) v = H[k];
      FillOut o;
      o.score = v; o.end_i = P.m; o.end_j = P.n; o.flags = 1 | 4;  // int32, long path
      A.fo[P.pid] = o;
    }
  }
}

__device__ __forceinline__ void load_table(const LongArgs &A, int8_t *tab) {
  for (int x = threadIdx.x; x < 32 * 32; x += blockDim.x) {
    const int r = x >> 5, c = x & 31;
    tab[x] = (r < A.Kp && c < A.Kp) ? A.table[r * A.Kp + c] : (int8_t)0;
  }
}

// Forward pass: one wave per workgroup, strips off the device queue.
__global__ __launch_bounds__(64, 4) void long_fwd_kernel(LongArgs A) {
  extern __shared__ __align__(16) uint8_t lds[];
  int8_t *tab = (int8_t *)lds;
  int8_t *slot = (int8_t *)lds + 1024 + threadIdx.x * fill_slot_stride(1, A.Kp);
  load_table(A, tab);
  __syncthreads();
  bool gave_up = false;
  for (;;) {
    int it = 0;
    if (threadIdx.x == 0) it = atomicAdd(&A.ctr[0], 1);
    it = __shfl(it, 0, 64);
    if (it >= A.nitems) break;
    const LongItem w = A.items[it];
    const LongPair P = A.lp[w.lp];
    const int s = w.s, rl = P.rl;
    int32_t *ring = A.ring + P.ring_off;
    int *pc = A.prog + P.prog_off;
    const AgentSync y{s > 0 ? pc + s - 1 : pc, pc + s};
    Region G{};
    G.s = s; G.c_left = 0; G.ncols = P.n;
    G.top_kind = s > 0 ? TOP_RING : TOP_ROW0;
    G.top = s > 0 ? ring + ((s - 1) & 1) * 2 * rl : nullptr;
    G.left = nullptr;
    G.out = s + 1 < P.S ? ring + (s & 1) * 2 * rl : nullptr;
    G.ck = (A.ckpt_on && s % KC == KC - 1 && s + 1 < P.S) ? A.ckpt + P.ck_off + (uint64_t)(s / KC) * 2 * rl : nullptr;
    G.colck = (A.ckpt_on && P.n > CKX) ? A.colck + P.colck_off : nullptr;
    G.tbs = nullptr; G.NBc = 0;
    G.score = s == P.S - 1;
    long_strip<AgentSync, false>(A, y, P, G, tab, slot, gave_up);
  }
}

// 4 direction bits of 1-based cell (i, j) of a recomputed block whose first
// strip is s0 (int32 layout: row k's e, f, d, x at bits 31-4k .. 28-4k).
__device__ __forceinline__ int long_code(const uint8_t *tbb, int s0, int NBc, int i, int j) {
  const int r0 = i - 1;
  const int s = r0 / LST - s0, w = r0 % LST, l = w / R, k = w % R;
  const int t = j - 1 + l;
  const uint32_t acc =
      *(const uint32_t *)(tbb + ((((uint64_t)s * GWAVE + l) * NBc + (t >> 3)) * 32) + (t & 7) * 4);
  return (int)(((acc >> (31 - 4 * k)) & 1) | (((acc >> (30 - 4 * k)) & 1) << 1) |
               (((acc >> (29 - 4 * k)) & 1) << 2) | (((acc >> (28 - 4 * k)) & 1) << 3));
}

// Traceback: one workgroup of LW waves per pair (pairs off a counter).
__global__ __launch_bounds__(64 * LW) void long_tb_kernel(LongArgs A) {
  extern __shared__ __align__(16) uint8_t lds[];
  int8_t *tab = (int8_t *)lds;
  int *slots = (int *)(lds + 1024);  // LW progress slots
  int *sh = slots + LW;              // [0] pair, [1] i, [2] j, [3] state
  int8_t *slot = (int8_t *)lds + 1024 + 128 + threadIdx.x * fill_slot_stride(1, A.Kp);
  int32_t *wring = A.wring + blockIdx.x * A.wring_stride;
  uint8_t *wtb = A.wtb + blockIdx.x * A.wtb_stride;
  const int wv = threadIdx.x >> 6;
  load_table(A, tab);
  bool gave_up = false;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) sh[0] = atomicAdd(&A.ctr[1], 1);
    __syncthreads();
    const int q = sh[0];
    if (q >= A.nlp) break;
    const LongPair P = A.lp[q];
    const PairDev pd = A.pd[P.pid];
    const int rl = P.rl;
    RunWriter rw;
    if (threadIdx.x == 0) {
      rw.lo = (uint32_t *)(A.cig + pd.cig_off);
      rw.wp = (uint32_t *)(A.cig + pd.cig_off + pd.cig_cap);
      sh[1] = P.m; sh[2] = P.n; sh[3] = 0;
    }
    for (;;) {
      __syncthreads();
      const int i = sh[1], j = sh[2];
      if (i == 0 || j == 0) break;
      // Region: rows of the block holding row i, columns c_left + 1 .. j.
      const int beta = (i - 1) / (KC * LST);
      const int s0 = beta * KC, top = s0 * LST, s1 = (i - 1) / LST + 1;
      int c_left = j - (i - top) - LMARGIN;
      c_left = c_left <= 0 ? 0 : (c_left / CKX) * CKX;
      const int ncols = j - c_left;
      const int NBc = (ncols + GWAVE + 7) / 8;
      if (threadIdx.x < LW) slots[threadIdx.x] = 0;
      __syncthreads();
      const int s = s0 + wv;
      if (s < s1) {
        const GroupSync y{slots + (s + LW - 1) % LW, slots + s % LW};
        Region G{};
        G.s = s; G.c_left = c_left; G.ncols = ncols;
        G.top_kind = s > s0 ? TOP_RING : (beta ? TOP_CKPT : TOP_ROW0);
        G.top = s > s0 ? wring + ((s - 1) & 1) * 2 * rl
                       : (beta ? A.ckpt + P.ck_off + (uint64_t)(beta - 1) * 2 * rl : nullptr);
        G.left = c_left ? A.colck + P.colck_off + (uint64_t)(c_left / CKX - 1) * P.S * LST : nullptr;
        G.out = s + 1 < s1 ? wring + (s & 1) * 2 * rl : nullptr;
        G.ck = nullptr; G.colck = nullptr;
        G.tbs = wtb + (uint64_t)(s - s0) * GWAVE * NBc * 32; G.NBc = NBc;
        G.score = false;
        long_strip<GroupSync, true>(A, y, P, G, tab, slot, gave_up);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        int ii = i, jj = j, st = sh[3];
        while (ii > top && jj > c_left) {
          const int code = long_code(wtb, s0, NBc, ii, jj - c_left);
          const int g = st ? st : ((code & 4) ? ((code & 8) ? 2 : 1) : 0);
          const int xb = g == 1 ? (code & 1) : (g == 2 ? ((code >> 1) & 1) : 0);
          rw.emit(g, 1);
          st = xb ? g : 0;
          jj -= g != 2;
          ii -= g != 1;
        }
        sh[1] = ii; sh[2] = jj; sh[3] = st;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int i = sh[1], j = sh[2];
      if (i == 0 && j > 0) rw.emit(1, j);
      if (j == 0 && i > 0) rw.emit(2, i);
      rw.finish();
      const FillOut f = A.fo[P.pid];
      ResultDev r;
      r.score = f.score;
      r.a_begin = 0; r.a_end = P.m - 1; r.b_begin = 0; r.b_end = P.n - 1;
      r.cigar_off = (int64_t)((char *)rw.wp - A.cig);
      r.cigar_len = (int32_t)(A.cig + pd.cig_off + pd.cig_cap - (char *)rw.wp);
      r.flags = f.flags | (rw.overflow ? WALK_FAILED : 0);
      A.res[P.pid] = r;
    }
  }
}

// LDS bytes of the two kernels.
__host__ __device__ constexpr int long_fwd_lds(int Kp) { return 1024 + 64 * fill_slot_stride(1, Kp); }
__host__ __device__ constexpr int long_tb_lds(int Kp) { return 1024 + 128 + 64 * LW * fill_slot_stride(1, Kp); }

}  // namespace ta
