// Traceback: direction planes -> coordinates + CIGAR, one thread per pair.
//
// The walk is the state machine of the specification (oracle/gotoh.c follows
// the same one over its own full-matrix codes): in state H follow the
// H-source (diag 'M', E, F); in state E emit 'D' and stay while the
// extend bit is set; in state F emit 'I' likewise.  Global pairs finish along
// row 0 / column 0; local pairs stop on entering a cell whose H is 0, which
// the walker knows by carrying the score backwards (H(i-1,j-1) = H - S,
// gap open / extend undone likewise), so no stop code is stored.
//
// Memory: the fill stores direction planes block-major ([strip][block of 8
// steps][lane]).  While the walk stays in one lane (8 rows) it needs at most
// two consecutive blocks of that lane, so each lane visit is one two-block
// window (two loads), kept in a padded per-thread LDS slice (dynamic indexing
// stays out of scratch).  Residue codes come through 8-byte register
// windows (the current residue in the top byte, shifted a byte per move).
// One iteration emits one CIGAR op and moves with selects, without
// branching on the state; runs leave as binary dwords.  The
// walk is latency-bound, so a wave carries few walkers (more waves per SIMD
// overlap their loads); walkers run longest-path first.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace ta {

struct TbArgs {
  const PairDev *pairs;
  int32_t npairs;
  const int32_t *order;  // walker -> pair (longest walks first); may be null
  const FillOut *fo;
  const PairTB *ptb;
  const uint8_t *codes;  // sequences start 16-byte aligned
  const int8_t *table;   // Kp x Kp
  int32_t Kp;
  const uint8_t *tb;
  char *cig;
  ResultDev *res;
  int32_t local, open, ext;
  int32_t do_cigar;
  int32_t guard_cap;     // most steps a walk may take before it ends flagged (normally above any walk)
  int32_t sys;           // res / cig are in host memory: store them system-coherent (bypassing the GPU caches)
};

// A dword to memory the host reads while the kernel runs (sys) or as usual.
__device__ __forceinline__ void put_word(uint32_t *p, uint32_t v, bool sys) {
  if (sys) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  else *p = v;
}

// The result of pair p.
__device__ __forceinline__ void put_result(const TbArgs &A, int p, const ResultDev &r) {
  if (A.sys) {
    const uint32_t *w = (const uint32_t *)&r;
    uint32_t *d = (uint32_t *)(A.res + p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(ResultDev) / 4); ++k) put_word(d + k, w[k], true);
  } else {
    A.res[p] = r;
  }
}

// ResultDev::flags of a walk that gave up (TA_F_WALK_FAILED in the C ABI).
constexpr int32_t WALK_FAILED = 0x100;

// Run-length CIGAR written backwards from the end of the pair's region as
// one dword per run, (length << 2) | op with op 0 M, 1 D, 2 I (the host turns
// runs into text when it copies results out).
struct RunWriter {
  uint32_t *lo = nullptr;  // start of the pair's region
  uint32_t *wp = nullptr;  // runs live in [wp, end)
  int op = -1;
  int len = 0;
  bool on = true;
  bool overflow = false;
  bool sys = false;  // the region is in host memory

  __device__ __forceinline__ void store() {
    if (!len || !on) return;
    if (wp <= lo) { overflow = true; return; }
    put_word(--wp, ((uint32_t)len << 2) | (uint32_t)op, sys);
  }
  __device__ __forceinline__ void emit(int o, int k) {
    if (o != op) { store(); op = o; len = 0; }
    len += k;
  }
  __device__ __forceinline__ void finish() { store(); len = 0; }
};

// The 8 residue codes of a sequence at positions i-8 .. i-1, position i-1 in
// the top byte (positions below 0 read as 0), from one dword-aligned 12-byte
// load.  A walk steps up one row by shifting the window left a byte, so the
// current residue is always the top byte.
__device__ __forceinline__ uint64_t tail8(const uint8_t *seq, int i) {
  if (i <= 0) return 0;
  const int lo = i - 8;
  const int base = (lo < 0 ? 0 : lo) & ~3;
  const uint32_t *w = (const uint32_t *)(seq + base);
  const uint32_t d0 = w[0], d1 = w[1], d2 = w[2];
  if (lo >= 0) {
    const uint32_t sh = (uint32_t)(lo - base);  // 0..3
    return ((uint64_t)__builtin_amdgcn_alignbyte(d2, d1, sh) << 32) | __builtin_amdgcn_alignbyte(d1, d0, sh);
  }
  return (((uint64_t)d1 << 32) | d0) << (8 * (8 - i));  // positions 0..7, i in 1..7
}

constexpr int WSTRIDE = 68;     // dwords per walker LDS slice: two 32-dword windows + pad (16-byte aligned)
constexpr int TB_WALKERS = 16;  // walkers (threads) per workgroup = per wave
constexpr int EPOCH = 8;        // steps per epoch (a walk moves at most 8 rows / columns)

// The result fields known before any walk.  False when there is nothing to
// walk: the result is then written (long pairs excepted: their own traceback
// kernel writes it).
__device__ __forceinline__ bool walk_prologue(const TbArgs &A, int p, PairDev &pd, FillOut &f, ResultDev &r) {
  pd = A.pairs[p];
  f = A.fo[p];
  r.score = f.score;
  r.flags = f.flags;
  r.cigar_off = (int64_t)(pd.cig_off + pd.cig_cap);
  r.cigar_len = 0;
  const bool local = A.local != 0;
  const bool empty = local && f.score <= 0;
  if (local) {
    r.a_begin = r.a_end = r.b_begin = r.b_end = -1;
    if (!empty) { r.a_end = f.end_i - 1; r.b_end = f.end_j - 1; }
  } else {
    r.a_begin = 0; r.a_end = pd.la - 1; r.b_begin = 0; r.b_end = pd.lb - 1;
  }
  if (empty || (!local && !A.do_cigar)) {
    if (empty) r.flags |= 8;
    put_result(A, p, r);
    return false;
  }
  return A.ptb[p].P != 3;
}

// One walk, resumable epoch by epoch (a walker lane can start another pair
// as soon as its walk ends).  An epoch loads everything the next EPOCH steps
// can touch -- the current lane's two-block window, the window of the lane
// above, and 32-byte windows of both sequences -- waits once, then steps with
// nothing in flight.
template <int P>
struct Walker {
  static constexpr bool p2 = P == 2;
  static constexpr int esz = p2 ? 64 : 32;

  RunWriter cw;
  ResultDev r;
  const uint8_t *pa, *pb, *tbase;
  char *cig_end;
  int p, GG, hsh, NB, off;
  int i, j, state, h, s, l, k, guard;
  int loc, gop, gex;  // mode and gap scores of this walk (taken from the TbArgs at begin)
  bool done;

  __device__ __forceinline__ void begin(const TbArgs &A, int p_, const PairDev &pd, const FillOut &f,
                                        const ResultDev &r_) {
    p = p_;
    r = r_;
    const PairTB d = A.ptb[p];
    cw = RunWriter();
    cw.on = A.do_cigar != 0;
    cw.sys = A.sys != 0;
    cig_end = A.cig + pd.cig_off + pd.cig_cap;
    if (cw.on) {
      cw.lo = (uint32_t *)(A.cig + pd.cig_off);
      cw.wp = (uint32_t *)cig_end;
    }
    pa = A.codes + pd.a_off;
    pb = A.codes + pd.b_off;
    GG = d.GG > 0 ? d.GG : G;
    hsh = 8 * d.half;
    NB = d.NB;
    off = d.off;
    tbase = A.tb + d.tb_off;
    const bool local = A.local != 0;
    loc = A.local; gop = A.open; gex = A.ext;
    i = local ? f.end_i : pd.la;
    j = local ? f.end_j : pd.lb;
    state = 0;  // 0 H, 1 E, 2 F
    h = f.score;
    s = 0; l = 0; k = 0;  // strip, lane, row-in-lane of row i
    if (i > 0) {
      const int r0 = i - 1;
      s = r0 >> (__builtin_ctz(GG) + 3);  // strips of GG * R rows (GG a power of two)
      l = (r0 >> 3) & (GG - 1);
      k = r0 & 7;
    }
    done = false;
    guard = 2 * (pd.la + pd.lb) + 2;  // a walk takes at most la + lb steps
    guard = guard < A.guard_cap ? guard : A.guard_cap;
  }

  __device__ __forceinline__ int first_block(int t) const {
    int ff = (t >> 3) - 1;
    ff = ff > NB - 2 ? NB - 2 : ff;
    return ff < 0 ? 0 : ff;
  }
  __device__ __forceinline__ const uint4 *block_at(int ss, int ll, int bb) const {
    return (const uint4 *)(tbase + (((uint64_t)ss * NB + bb) * GG + ll) * esz);
  }

  // Two blocks of one lane into a slice window: every load issued before
  // the first LDS store.
  __device__ __forceinline__ static void copy_window(uint32_t *dst, const uint4 *src0, const uint4 *src1) {
    if constexpr (p2) {
      const uint4 x0 = src0[0], x1 = src0[1], x2 = src0[2], x3 = src0[3];
      const uint4 y0 = src1[0], y1 = src1[1], y2 = src1[2], y3 = src1[3];
      uint4 *d = (uint4 *)dst;
      d[0] = x0; d[1] = x1; d[2] = x2; d[3] = x3;
      d[4] = y0; d[5] = y1; d[6] = y2; d[7] = y3;
    } else {
      const uint4 x0 = src0[0], x1 = src0[1];
      const uint4 y0 = src1[0], y1 = src1[1];
      uint4 *d = (uint4 *)dst;
      d[0] = x0; d[1] = x1; d[2] = y0; d[3] = y1;
    }
  }

  // Result of the finished walk.
  __device__ __forceinline__ void finish(const TbArgs &A) {
    cw.finish();
    if (guard <= 0 || cw.overflow) r.flags |= WALK_FAILED;
    if (loc) { r.a_begin = i; r.b_begin = j; }
    if (cw.on) {
      r.cigar_off = (int64_t)((char *)cw.wp - A.cig);
      r.cigar_len = (int32_t)(cig_end - (char *)cw.wp);  // bytes of runs
    }
    put_result(A, p, r);
  }

  // Direction code (e f d x in bits 0..3) of row k in a lane's block entry.
  __device__ __forceinline__ int code_at(const uint32_t *entry) const {
    if constexpr (p2) {
      const uint2 v = *(const uint2 *)entry;
      const int sh = hsh + k;
      return (int)(((v.x >> sh) & 1) | (((v.x >> (16 + sh)) & 1) << 1) | (((v.y >> sh) & 1) << 2) |
                   (((v.y >> (16 + sh)) & 1) << 3));
    } else {
      const uint32_t acc = *entry;
      const int pbit = 31 - 4 * k;  // pushes e, f, d, x of row k land at pbit .. pbit-3
      return (int)(((acc >> pbit) & 1) | (((acc >> (pbit - 1)) & 1) << 1) | (((acc >> (pbit - 2)) & 1) << 2) |
                   (((acc >> (pbit - 3)) & 1) << 3));
    }
  }

  // One step from the current cell with its direction code: the op (M 0,
  // D 1, I 2; from state H a gap source is taken at once with the same
  // cell's extend bit), the carried score (local; `diag` gives the cell's
  // substitution score), the run, the move.  True when a local walk has
  // entered a cell whose H is 0 (or a border): the walk is over.
  template <class Diag>
  __device__ __forceinline__ bool take(int code, bool local, int open, int ext, Diag diag) {
    const int g = state ? state : ((code & 4) ? ((code & 8) ? 2 : 1) : 0);
    const int xb = g == 1 ? (code & 1) : (g == 2 ? ((code >> 1) & 1) : 0);
    if (local) h -= g == 0 ? diag() : (xb ? ext : open);
    cw.emit(g, 1);
    state = xb ? g : 0;
    // the move, in selects: a row move steps k down, borrowing from the lane
    // (l), which borrows from the strip (s); GG is a power of two
    const bool up = g != 1;
    j -= g != 2;
    i -= up;
    const bool bk = up && k == 0, bl = bk && l == 0;
    k = up ? ((k - 1) & 7) : k;
    l = bk ? ((l - 1) & (GG - 1)) : l;
    s -= bl;
    return local && g == 0 && (h == 0 || i == 0 || j == 0);
  }

  // The walk's head: borders that end it.  True when it is over (result
  // written).
  __device__ __forceinline__ bool at_border(const TbArgs &A) {
    if (!loc) {
      if (i == 0) { if (j) cw.emit(1, j); finish(A); return true; }
      if (j == 0) { cw.emit(2, i); finish(A); return true; }
    } else if (i <= 0 || j <= 0) {
      r.flags |= WALK_FAILED;
      finish(A);
      return true;
    }
    return false;
  }

  // One epoch; true when the walk has ended and its result is written.
  __device__ __forceinline__ bool epoch(const TbArgs &A, const int8_t *stab, uint32_t *slice) {
    const bool local = loc != 0;
    const int open = gop, ext = gex;
    if (at_border(A)) return true;
    // ---- epoch loads
    const int t0 = j - 1 + off + l;
    const int fA = first_block(t0);
    int sB = s, lB = l - 1, tB = t0 - 1;
    if (lB < 0) { lB = GG - 1; sB = s - 1; tB += GG; }
    const bool hasB = sB >= 0;
    const int fB = hasB ? first_block(tB) : 0;
    const int sA = s, lA = l;
    copy_window(slice, block_at(s, l, fA), block_at(s, l, fA + 1));
    if (hasB) copy_window(slice + 32, block_at(sB, lB, fB), block_at(sB, lB, fB + 1));
    uint64_t ra = 0, rb = 0;  // local: residues of the rows / columns ahead, current one on top
    if (local) {
      ra = tail8(pa, i);
      rb = tail8(pb, j);
    }
    // ---- up to EPOCH steps inside the two windows.  In EPOCH steps a walk
    // climbs at most one lane boundary, so it is in the first window's lane
    // or in the lane above (the second window).
    int e = 0;
#pragma unroll 1
    for (; e < EPOCH && guard > 0; ++e, --guard) {
      if (!local && (i == 0 || j == 0)) break;  // boundary handled by the epoch head
      const int t = j - 1 + off + l;
      const bool inA = s == sA && l == lA;
      const int wi = t - 8 * (inA ? fA : fB);
      if (!(inA || hasB) || wi < 0 || wi >= 16) break;  // left both windows: next epoch
      const int code = code_at(slice + (inA ? 0 : 32) + (wi >> 3) * (p2 ? 16 : 8) + (wi & 7) * (p2 ? 2 : 1));
      const int i0 = i, j0 = j;
      if (take(code, local, open, ext, [&] { return (int)stab[((int)(ra >> 56) << 5) | (int)(rb >> 56)]; })) {
        done = true;
        break;
      }
      ra = i != i0 ? ra << 8 : ra;
      rb = j != j0 ? rb << 8 : rb;
    }
    // An epoch starts inside its own window, so it always takes a step; one
    // that could not (a start cell outside the planes) ends the walk,
    // flagged, rather than looping.
    if (e == 0 && !done) guard = 0;
    if (done || guard <= 0) {
      finish(A);
      return true;
    }
    return false;
  }
};

// One pair, walked to the end by one thread.
__device__ __forceinline__ void walk_pair(const TbArgs &A, int p, const int8_t *stab, uint32_t *slice) {
  PairDev pd;
  FillOut f;
  ResultDev r;
  if (!walk_prologue(A, p, pd, f, r)) return;
  if (A.ptb[p].P == 2) {
    Walker<2> w;
    w.begin(A, p, pd, f, r);
    while (!w.epoch(A, stab, slice)) {}
  } else {
    Walker<1> w;
    w.begin(A, p, pd, f, r);
    while (!w.epoch(A, stab, slice)) {}
  }
}

// A template so that every source including this header may hold a copy
// (the serving kernel's source includes the fill, which includes this).
template <int = 0>
__global__ __launch_bounds__(TB_WALKERS) void traceback_kernel(TbArgs A) {
  __shared__ uint32_t wlds[TB_WALKERS * WSTRIDE];
  __shared__ int8_t stab[1024];
  for (int x = threadIdx.x; x < 1024; x += TB_WALKERS) {
    const int r = x >> 5, c = x & 31;
    stab[x] = (r < A.Kp && c < A.Kp) ? A.table[r * A.Kp + c] : (int8_t)0;
  }
  __syncthreads();

  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.npairs) return;
  walk_pair(A, A.order ? A.order[w] : w, stab, wlds + threadIdx.x * WSTRIDE);
}

}  // namespace ta
