// Traceback: direction planes -> coordinates + CIGAR, one thread per pair.
//
// The walk is the state machine of the specification (oracle/gotoh.c follows
// the same one over its own full-matrix codes): in state H follow the
// H-source (diag 'M', E, F); in state E emit 'D' and stay while the
// extend bit is set; in state F emit 'I' likewise.  Global pairs finish along
// row 0 / column 0; local pairs stop on entering a cell whose H is 0, which
// the walker knows by carrying the score backwards (H(i-1,j-1) = H - S,
// gap open / extend undone likewise), so no stop code is stored.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace ta {

struct TbArgs {
  const PairDev *pairs;
  int32_t npairs;
  const FillOut *fo;
  const PairTB *ptb;
  const uint8_t *codes;
  const int8_t *table;  // Kp x Kp
  int32_t Kp;
  const uint8_t *tb;
  char *cig;
  ResultDev *res;
  int32_t local, open, ext;
  int32_t do_cigar;
};

// 4 direction bits of cell (i, j) (1-based) of a pair: bit0 E-extend,
// bit1 F-extend, bit2 not-diag, bit3 F-rather-than-E.
struct TbReader {
  const uint8_t *tb;
  PairTB d;
  uint64_t cached = ~0ull;
  uint4 blk[4];

  __device__ __forceinline__ int get(int i, int j) {
    const int r0 = i - 1;
    const int s = r0 / STRIP, w = r0 - s * STRIP, l = w / R, k = w - l * R;
    const int t = j - 1 + d.off + l, bi = t >> 3, q = t & 7;
    const int esz = d.P == 2 ? 64 : 32;
    const uint64_t base = d.tb_off + (((uint64_t)s * d.NB + bi) * G + l) * esz;
    if (base != cached) {
      const uint4 *src = (const uint4 *)(tb + base);
      blk[0] = src[0];
      blk[1] = src[1];
      if (d.P == 2) { blk[2] = src[2]; blk[3] = src[3]; }
      cached = base;
    }
    const uint32_t *w32 = (const uint32_t *)blk;
    if (d.P == 2) {
      const uint32_t xc = w32[2 * q], yc = w32[2 * q + 1];
      const int sh = 8 * d.half + k;
      return (int)(((xc >> sh) & 1) | (((xc >> (16 + sh)) & 1) << 1) | (((yc >> sh) & 1) << 2) |
                   (((yc >> (16 + sh)) & 1) << 3));
    }
    const uint32_t acc = w32[q];
    const int p = 31 - 4 * k;  // pushes e, f, d, x of row k land at p, p-1, p-2, p-3
    return (int)(((acc >> p) & 1) | (((acc >> (p - 1)) & 1) << 1) | (((acc >> (p - 2)) & 1) << 2) |
                 (((acc >> (p - 3)) & 1) << 3));
  }
};

struct CigarWriter {
  char *base;   // start of the arena
  char *lo;     // start of this pair's region (writes never go below it)
  char *wp;     // next byte to write is wp[-1]
  bool overflow = false;
  char op = 0;
  int len = 0;
  bool on = true;  // false: walk for coordinates only

  __device__ __forceinline__ void flush() {
    if (!len || !on) return;
    int digits = 1;
    for (int v = len; v >= 10; v /= 10) ++digits;
    if (wp - lo < digits + 1) {
      overflow = true;
      len = 0;
      return;
    }
    *--wp = op;
    int v = len;
    do { *--wp = (char)('0' + v % 10); v /= 10; } while (v);
    len = 0;
  }
  __device__ __forceinline__ void emit(char o, int k = 1) {
    if (k <= 0) return;
    if (o != op) { flush(); op = o; }
    len += k;
  }
};

__global__ __launch_bounds__(64) void traceback_kernel(TbArgs A) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.npairs) return;
  const PairDev pd = A.pairs[p];
  const FillOut f = A.fo[p];
  ResultDev r;
  r.score = f.score;
  r.flags = f.flags;
  r.cigar_off = (int64_t)(pd.cig_off + pd.cig_cap);
  r.cigar_len = 0;
  const bool empty = A.local && f.score <= 0;
  if (A.local) {
    r.a_begin = r.a_end = r.b_begin = r.b_end = -1;
    if (!empty) { r.a_end = f.end_i - 1; r.b_end = f.end_j - 1; }
  } else {
    r.a_begin = 0; r.a_end = pd.la - 1; r.b_begin = 0; r.b_end = pd.lb - 1;
  }
  if (empty || (!A.local && !A.do_cigar)) {
    if (empty) r.flags |= 8;
    A.res[p] = r;
    return;
  }

  TbReader rd;
  rd.tb = A.tb;
  rd.d = A.ptb[p];
  CigarWriter cw;
  cw.base = A.cig;
  cw.on = A.do_cigar != 0;
  cw.wp = cw.on ? A.cig + pd.cig_off + pd.cig_cap : nullptr;
  cw.lo = cw.on ? A.cig + pd.cig_off : nullptr;
  const uint8_t *a = A.codes + pd.a_off;
  const uint8_t *b = A.codes + pd.b_off;

  int i = A.local ? f.end_i : pd.la;
  int j = A.local ? f.end_j : pd.lb;
  int state = 0;  // 0 H, 1 E, 2 F
  int h = f.score;
  // Every iteration moves or changes state, so 2 * (la + lb) + 2 bound it;
  // a walk that leaves the matrix is reported instead of read out of range.
  int guard = 2 * (pd.la + pd.lb) + 2;
  for (; guard > 0; --guard) {
    if (!A.local) {
      if (i == 0) { cw.emit('D', j); break; }
      if (j == 0) { cw.emit('I', i); break; }
    } else if (i <= 0 || j <= 0) {
      r.flags |= 0x100;
      break;
    }
    const int code = rd.get(i, j);
    if (state == 0) {
      if (!(code & 4)) {
        cw.emit('M');
        h -= A.table[a[i - 1] * A.Kp + b[j - 1]];
        --i;
        --j;
        if (A.local && (h == 0 || i == 0 || j == 0)) break;
      } else {
        state = (code & 8) ? 2 : 1;
      }
    } else if (state == 1) {
      cw.emit('D');
      if (code & 1) { h -= A.ext; } else { h -= A.open; state = 0; }
      --j;
    } else {
      cw.emit('I');
      if (code & 2) { h -= A.ext; } else { h -= A.open; state = 0; }
      --i;
    }
  }
  cw.flush();
  if (guard == 0 || cw.overflow) r.flags |= 0x100;
  if (A.local) { r.a_begin = i; r.b_begin = j; }
  if (cw.on) {
    r.cigar_off = (int64_t)(cw.wp - A.cig);
    r.cigar_len = (int32_t)(A.cig + pd.cig_off + pd.cig_cap - cw.wp);
  }
  A.res[p] = r;
}

}  // namespace ta
