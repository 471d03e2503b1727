// One pair at low latency: the editor's "realign selection" path.
//
// In the reference, realigning two selected rows runs the external aligner
// on them (tweakseq/UI/SeqEditMainWin.cpp:1616-1663) and splices the result
// back at the selection (tweakseq/Core/Project.cpp:964-1026).  Here it is
// one launch of one workgroup that takes the letters and returns the CIGAR
// runs, with nothing else on the device:
//  * the letters are read straight from host staging (mapped, pinned
//    memory) and encoded in LDS, the substitution table is built in LDS
//    from a copy compiled into the kernel;
//  * the fill is the striped form of the batch fill in int32 with PR rows
//    per lane: a strip is 64 lanes x PR rows, strip s runs on wave s % PW
//    and trails strip s-1 through a ring of bottom rows in LDS (progress
//    counters, workgroup scope, bounded waits);
//  * direction codes (4 bits per cell) stay in LDS, so the walk never
//    touches HBM: wave 0 loads a 16 x 16-cell window of codes into its
//    lanes (one byte per lane per row pair) and steps through it with
//    readlane on wave-uniform coordinates, reloading when the walk leaves
//    the window;
//  * the result and the runs go straight back to host staging.
// Pairs whose codes do not fit in LDS take the batch path (the caller's
// entry point decides).  Recurrence, boundaries and tie-breaks are those of
// the specification the CPU oracle implements (oracle/gotoh.c); the local
// stop cell (H = 0) is marked in its code as d = 0, x = 1 (a combination
// the H-source never takes), since no walk reads another bit of such a cell.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "../../include/tweakalign.h"
#include "blosum62.h"
#include "common.h"
#include "pair.h"
#include "runtime_internal.h"

namespace ta {

__constant__ int8_t kBlosumDev[] = {TA_BLOSUM62_LOWER};
__constant__ char kProtOrderDev[] = TA_PROT_ORDER;

namespace {

// LDS words of the sync area.
enum { SY_STATUS = 0, SY_SCORE = 1, SY_PROG = 4, SY_BEST = 8, SY_WORDS = 64 };

__device__ __forceinline__ int wave_shr1(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);  // wave_shr:1, lane 0 takes old
}

// (h desc, i asc, j asc): the local end-cell order.
__device__ __forceinline__ bool better(int h, int i, int j, int bh, int bi, int bj) {
  return h > bh || (h == bh && (i < bi || (i == bi && j < bj)));
}

__device__ __forceinline__ int load_code(const int8_t *lmap, char ch, int &bad) {
  const int c = lmap[(uint8_t)ch];
  bad |= c < 0;
  return c < 0 ? PAIR_PAD : c;
}

template <bool LOCAL>
__global__ __launch_bounds__(64 * PAIR_WAVES) void pair_kernel(PairArgs A) {
  extern __shared__ __align__(16) uint8_t lds[];
  const PairLayout &Y = A.lay;
  int8_t *tab = (int8_t *)lds;                   // [a code][b code], 32 x 32
  int8_t *lmap = (int8_t *)(lds + 1024);         // letter -> code, -1 outside the alphabet
  uint8_t *acode = lds + Y.acode;
  uint8_t *bcode = lds + Y.bcode;
  int32_t *sy = (int32_t *)(lds + Y.sync);
  int2 *ring = (int2 *)(lds + Y.ring);
  uint8_t *planes = lds + Y.planes;
  uint32_t *runs = (uint32_t *)(lds + Y.runs);
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m = A.m, n = A.n, T = Y.T, S = Y.S;
  const int open = A.open, ext = A.ext;

  for (int x = tid; x < 1024; x += blockDim.x) {
    const int r = x >> 5, c = x & 31;
    int v = 0;
    if (A.alphabet == TA_PROTEIN) {
      if (r < TA_PROT_K && c < TA_PROT_K) {
        const int hi = r > c ? r : c, lo = r > c ? c : r;
        v = kBlosumDev[hi * (hi + 1) / 2 + lo];
      }
    } else if (r < 4 && c < 4) {
      v = r == c ? A.match : A.mismatch;
    }
    tab[x] = (int8_t)v;
  }
  for (int x = tid; x < 256; x += blockDim.x) {
    const int u = (x >= 'a' && x <= 'z') ? x - 32 : x;
    int code = -1;
    if (A.alphabet == TA_PROTEIN) {
      for (int r = 0; r < TA_PROT_K; ++r)
        if (kProtOrderDev[r] == u) code = r;
      if (u == 'J' || u == 'O' || u == 'U') code = TA_PROT_K - 1;  // scored as X
    } else {
      code = u == 'A' ? 0 : (u == 'C' ? 1 : (u == 'G' ? 2 : (u == 'T' ? 3 : -1)));
    }
    lmap[x] = (int8_t)code;
  }
  if (tid < SY_WORDS) sy[tid] = 0;
  __syncthreads();
  int bad = 0;
  for (int x = tid; x < m; x += blockDim.x) acode[x] = (uint8_t)load_code(lmap, A.a[x], bad);
  for (int x = tid; x < T + 8; x += blockDim.x) bcode[x] = x < n ? (uint8_t)load_code(lmap, A.b[x], bad) : PAIR_PAD;
  if (bad) atomicOr(&sy[SY_STATUS], PAIR_BAD_LETTER);
  __syncthreads();
  if (sy[SY_STATUS] & PAIR_BAD_LETTER) {
    if (tid == 0) A.out->status = sy[SY_STATUS];
    return;
  }

  // ---- fill
  int bh = 0, bi = 0x7fffffff, bj = 0x7fffffff;  // lane's best cell (local), 1-based
  for (int s = wv; s < S; s += PAIR_WAVES) {
    const int r0 = s * PAIR_STRIP + PAIR_R * lane;  // 0-based row of the lane's first row
    const int8_t *t0 = tab + ((r0 < m ? acode[r0] : PAIR_PAD) << 5);
    const int8_t *t1 = tab + ((r0 + 1 < m ? acode[r0 + 1] : PAIR_PAD) << 5);
    auto left = [&](int i) { return (LOCAL || i == 0) ? 0 : open + (i - 1) * ext; };  // H(i, 0), 1-based
    int H0 = left(r0 + 1), H1 = left(r0 + 2), E0 = NEG32, E1 = NEG32;
    int hd = left(r0);  // H(r0, c - 1) of the step to come (1-based row r0 = the row above)
    int Hbot = 0, Fbot = NEG32;
    // Local: one key per row, (H << 16) | (0xFFFF - column): the maximum is
    // the best H at its first column (columns and H stay below 2^16 here).
    uint32_t key0 = 0, key1 = 0;
    const int2 *rin = ring + ((s + PAIR_WAVES - 1) % PAIR_WAVES) * Y.ring_len;
    int2 *rout = ring + (s % PAIR_WAVES) * Y.ring_len;
    int32_t *pin = sy + SY_PROG + (s + PAIR_WAVES - 1) % PAIR_WAVES, *pout = sy + SY_PROG + s % PAIR_WAVES;
    const bool wr = s + 1 < S;
    uint32_t *pl = (uint32_t *)(planes + (size_t)s * T * 64) + lane;  // [t / 4][lane] dwords, byte t % 4
    // the column code of step 0 (lane 0: column 0; the others idle) and its
    // scores; later steps read theirs one step ahead
    int code = wave_shr1(bcode[0], PAIR_PAD);
    int s0n = t0[code], s1n = t1[code];
    uint32_t pk = 0;  // plane bytes of the current 4 steps
    for (int tb = 0; tb < T; tb += 8) {
      int topH[8], topF[8];
      if (s == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          topH[q] = LOCAL ? 0 : open + (tb + q) * ext;  // H(0, tb + q + 1)
          topF[q] = NEG32;
        }
      } else {
        // ring columns [tb, tb + 8) of the strip above: wait for them (bounded)
        const int need = (s << 20) | (tb + 8 < n ? tb + 8 : n);
        int spins = 0;
        while (__hip_atomic_load(pin, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need) {
          if (++spins > A.spin_limit) {
            if (lane == 0) atomicOr(&sy[SY_STATUS], PAIR_GAVE_UP);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int2 v = rin[tb + q < n ? tb + q : n - 1];
          topH[q] = v.x;
          topF[q] = v.y;
        }
      }
      const uint2 bw0 = *(const uint2 *)(bcode + tb), bw1 = *(const uint2 *)(bcode + tb + 8);
      const uint32_t bw[4] = {bw0.x, bw0.y, bw1.x, bw1.y};
      // The 8 steps, in two forms: with the per-lane column test, and without
      // it where every lane is inside the columns (most blocks).
      auto steps = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int t = tb + q;
          const int S0 = s0n, S1 = s1n;
          // the next step's column code (lane l takes lane l-1's) and scores
          const int bnext = (int)((bw[(q + 1) >> 2] >> (8 * ((q + 1) & 3))) & 0xff);
          const int cnext = wave_shr1(bnext, code);
          s0n = t0[cnext];
          s1n = t1[cnext];
          code = cnext;
          const int Hup = wave_shr1(topH[q], Hbot), Fup = wave_shr1(topF[q], Fbot);
          const int c0 = t - lane;  // 0-based column of this lane at this step
          if (FULL || (c0 >= 0 && c0 < n)) {
            // Direction bits are sign bits of differences (|values| < 2^30),
            // pushed into acc: e0 f0 d0 x0 e1 f1 d1 x1 from bit 0 up.  x (F
            // above E) is kept only where d (H not from the diagonal) is set,
            // so that d = 0, x = 1 marks a local stop cell (H = 0), whose
            // other bits no walk reads.
            uint32_t acc = 0;
            // row r0 + 1 (1-based)
            const int HdS0 = hd + S0;
            const int Eo0 = H0 + open, Ee0 = E0 + ext, En0 = max(Eo0, Ee0);
            const int Fo0 = Hup + open, Fe0 = Fup + ext, Fn0 = max(Fo0, Fe0);
            int Hn0 = max(HdS0, max(En0, Fn0));
            if (LOCAL) Hn0 = max(Hn0, 0);
            // row r0 + 2: the row above is the one just computed
            const int HdS1 = H0 + S1;
            const int Eo1 = H1 + open, Ee1 = E1 + ext, En1 = max(Eo1, Ee1);
            const int Fo1 = Hn0 + open, Fe1 = Fn0 + ext, Fn1 = max(Fo1, Fe1);
            int Hn1 = max(HdS1, max(En1, Fn1));
            if (LOCAL) Hn1 = max(Hn1, 0);
            const uint32_t dD0 = (uint32_t)(HdS0 - Hn0), dD1 = (uint32_t)(HdS1 - Hn1);
            uint32_t xd0 = (uint32_t)(En0 - Fn0) & dD0, xd1 = (uint32_t)(En1 - Fn1) & dD1;
            uint32_t dd0 = dD0, dd1 = dD1;
            if (LOCAL) {  // H >= 0: H = 0 sets x and clears d
              xd0 |= (uint32_t)(Hn0 - 1); xd1 |= (uint32_t)(Hn1 - 1);
              dd0 &= (uint32_t)(-Hn0); dd1 &= (uint32_t)(-Hn1);
              key0 = max(key0, ((uint32_t)Hn0 << 16) | (uint32_t)(0xFFFF - c0));
              key1 = max(key1, ((uint32_t)Hn1 << 16) | (uint32_t)(0xFFFF - c0));
            }
            acc = __builtin_amdgcn_alignbit(acc, xd1, 31);
            acc = __builtin_amdgcn_alignbit(acc, dd1, 31);
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(Fo1 - Fe1), 31);
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(Eo1 - Ee1), 31);
            acc = __builtin_amdgcn_alignbit(acc, xd0, 31);
            acc = __builtin_amdgcn_alignbit(acc, dd0, 31);
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(Fo0 - Fe0), 31);
            acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(Eo0 - Ee0), 31);
            pk = (pk & ~(0xffu << (8 * (q & 3)))) | ((acc & 0xffu) << (8 * (q & 3)));
            if (wr && lane == 63) rout[c0] = make_int2(Hn1, Fn1);
            hd = Hup;
            H0 = Hn0; H1 = Hn1; E0 = En0; E1 = En1;
            Hbot = Hn1; Fbot = Fn1;
          }
          if ((q & 3) == 3) pl[(size_t)(t >> 2) * 64] = pk;
        }
      };
      if (tb >= 63 && tb + 8 <= n) steps(std::true_type{});
      else steps(std::false_type{});
      if (wr) {
        int done = tb + 8 - 63;  // lane 63 has stored columns 0 .. tb + 7 - 63
        done = done < 0 ? 0 : (done > n ? n : done);
        if (lane == 63)
          __hip_atomic_store(pout, ((s + 1) << 20) | done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    if (LOCAL) {
      const int rb0 = (int)(key0 >> 16), rc0 = 0x10000 - (int)(key0 & 0xffff);  // 1-based column
      const int rb1 = (int)(key1 >> 16), rc1 = 0x10000 - (int)(key1 & 0xffff);
      if (r0 < m && rb0 > 0 && better(rb0, r0 + 1, rc0, bh, bi, bj)) { bh = rb0; bi = r0 + 1; bj = rc0; }
      if (r0 + 1 < m && rb1 > 0 && better(rb1, r0 + 2, rc1, bh, bi, bj)) { bh = rb1; bi = r0 + 2; bj = rc1; }
    } else {
      if (r0 + 1 == m) sy[SY_SCORE] = H0;
      if (r0 + 2 == m) sy[SY_SCORE] = H1;
    }
  }
  if (LOCAL) {
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
      const int oh = __shfl_xor(bh, w, 64), oi = __shfl_xor(bi, w, 64), oj = __shfl_xor(bj, w, 64);
      if (better(oh, oi, oj, bh, bi, bj)) { bh = oh; bi = oi; bj = oj; }
    }
    if (lane == 0) {
      sy[SY_BEST + 3 * wv] = bh; sy[SY_BEST + 3 * wv + 1] = bi; sy[SY_BEST + 3 * wv + 2] = bj;
    }
  }
  __syncthreads();
  if (wv != 0) return;

  // ---- end cell and walk (wave 0)
  ResultDev r;
  r.flags = TA_F_I32;
  int i = m, j = n;
  if (LOCAL) {
    bh = 0; bi = 0x7fffffff; bj = 0x7fffffff;
    for (int w = 0; w < PAIR_WAVES; ++w) {
      const int oh = sy[SY_BEST + 3 * w], oi = sy[SY_BEST + 3 * w + 1], oj = sy[SY_BEST + 3 * w + 2];
      if (oh > 0 && better(oh, oi, oj, bh, bi, bj)) { bh = oh; bi = oi; bj = oj; }
    }
    r.score = bh;
    if (bh <= 0) {
      r.score = 0;
      r.a_begin = r.a_end = r.b_begin = r.b_end = -1;
      r.flags |= TA_F_EMPTY;
      i = j = 0;
    } else {
      i = bi; j = bj;
      r.a_end = bi - 1; r.b_end = bj - 1;
    }
  } else {
    r.score = sy[SY_SCORE];
    r.a_begin = 0; r.a_end = m - 1; r.b_begin = 0; r.b_end = n - 1;
  }
  const bool walk = LOCAL ? r.score > 0 : A.traceback != 0;
  const int cap = Y.runs_cap;
  int wp = cap, op = -1, len = 0, state = 0, steps = 0;
  bool fail = false, stop = false;
  auto flush = [&]() {
    if (len) {
      --wp;
      runs[wp] = ((uint32_t)len << 2) | (uint32_t)op;  // every lane stores the same word
    }
  };
  auto emit = [&](int o, int k) {
    if (o != op) { flush(); op = o; len = 0; }
    len += k;
  };
  while (walk && !stop && !fail && i > 0 && j > 0) {
    // the 16 x 16 cells above-left of (i, j): lane q holds row pairs rp0 - (q >> 4) (- 4 in v1),
    // column j - 1 - (q & 15) (0-based), one code byte (both rows of the pair)
    const int rp0 = (i - 1) >> 1, cw0 = j - 1;
    int v[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int rp = rp0 - (lane >> 4) - 4 * h, cj = cw0 - (lane & 15);
      int byte = 0;
      if (rp >= 0 && cj >= 0) {
        const int ss = rp >> 6, ll = rp & 63;
        const int t = cj + ll;
        byte = planes[((size_t)ss * T + (t & ~3)) * 64 + 4 * ll + (t & 3)];
      }
      v[h] = byte;
    }
    // Steps inside the window, all on wave-uniform (scalar) values: the
    // code by readlane, the transition by selects.
    for (;;) {
      const int dyp = rp0 - ((i - 1) >> 1), dx = cw0 - (j - 1);
      if (i == 0 || j == 0 || dyp > 7 || dx > 15) break;
      const int lo = __builtin_amdgcn_readlane(v[0], ((dyp & 3) << 4) | dx);
      const int hi = __builtin_amdgcn_readlane(v[1], ((dyp & 3) << 4) | dx);
      const int cd = ((dyp < 4 ? lo : hi) >> (((i - 1) & 1) << 2)) & 15;
      if (LOCAL && state == 0 && (cd & 12) == 8) { stop = true; break; }  // entered a cell whose H is 0
      const int src = ((cd >> 2) & 1) * (1 + ((cd >> 3) & 1));  // H-source: 0 diag, 1 E, 2 F
      const int g = state ? state : src;
      const int xb = g ? (cd >> (g - 1)) & 1 : 0;              // E: bit 0, F: bit 1 (extend)
      if (g != op) { flush(); op = g; len = 0; }
      ++len;
      state = xb ? g : 0;
      j -= g != 2;
      i -= g != 1;
      if (++steps > A.walk_cap) { fail = true; break; }
    }
  }
  if (walk && !fail && !LOCAL) {
    if (i == 0 && j > 0) emit(1, j);
    if (j == 0 && i > 0) emit(2, i);
    i = j = 0;
  }
  if (walk) flush();
  if (LOCAL && r.score > 0) { r.a_begin = i; r.b_begin = j; }
  if (fail) r.flags |= TA_F_WALK_FAILED;
  const int nr = cap - wp;
  uint32_t *dst = (uint32_t *)(A.out + 1);
  if (A.traceback)
    for (int q = lane; q < nr; q += 64) dst[q] = runs[wp + q];
  if (lane == 0) {
    r.cigar_off = 0;
    r.cigar_len = A.traceback ? 4 * nr : 0;
    A.out->r = r;
    A.out->nruns = A.traceback ? nr : 0;
    A.out->status = sy[SY_STATUS];
  }
}

}  // namespace

PairLayout pair_layout(int m, int n) {
  PairLayout L{};
  auto up16 = [](long x) { return (int)((x + 15) & ~15L); };
  L.T = (n + 63 + 7) & ~7;
  L.S = (m + PAIR_STRIP - 1) / PAIR_STRIP;
  L.ring_len = n + 8;
  L.runs_cap = m + n + 2;
  long off = 1280;
  L.acode = (int)off; off += up16(m);
  L.bcode = (int)off; off += up16(L.T + 8);
  L.sync = (int)off; off += 4 * SY_WORDS;
  L.ring = (int)off; off += (long)PAIR_WAVES * L.ring_len * 8;
  L.planes = (int)off; off += (long)L.S * L.T * 64;
  L.runs = (int)off; off += 4L * L.runs_cap;
  L.bytes = off > (1L << 30) ? (1 << 30) : (int)off;
  return L;
}

int pair_launch(const PairArgs &A, hipStream_t st) {
  (void)hipGetLastError();
  if (A.local)
    hipLaunchKernelGGL(pair_kernel<true>, dim3(1), dim3(64 * PAIR_WAVES), A.lay.bytes, st, A);
  else
    hipLaunchKernelGGL(pair_kernel<false>, dim3(1), dim3(64 * PAIR_WAVES), A.lay.bytes, st, A);
  return hipGetLastError() == hipSuccess ? TA_OK : TA_E_DEVICE;
}

}  // namespace ta

// ---------------------------------------------------------------------------
// Host side.
namespace {

using namespace ta;

// Result, CIGAR text and status of a finished pair kernel (record readable
// on the host).
int pair_finish(const PairOut &o, const uint32_t *runs, const ta_params *p, ta_result *out, char *cigar_buf,
                int64_t cap, int64_t *used) {
  if (o.status & PAIR_BAD_LETTER) return TA_E_ALPHABET;
  if (o.status & PAIR_GAVE_UP) return TA_E_DEVICE;
  const ResultDev &r = o.r;
  out->score = r.score;
  out->a_begin = r.a_begin; out->a_end = r.a_end;
  out->b_begin = r.b_begin; out->b_end = r.b_end;
  out->flags = r.flags;
  out->cigar_off = 0;
  out->cigar_len = 0;
  if (p->traceback && o.nruns > 0) {
    const int64_t len = cigar_text(runs, o.nruns, nullptr);
    out->cigar_len = (int32_t)len;
    if (used) *used = len;
    if (!cigar_buf || len > cap) return TA_E_SPACE;
    cigar_text(runs, o.nruns, cigar_buf);
  } else if (used) {
    *used = 0;
  }
  return (r.flags & TA_F_WALK_FAILED) ? TA_E_WALK : TA_OK;
}

// The pair path takes pairs whose codes fit in LDS; others (and empty
// sides, and forced long-pair lanes) go through the batch path.
bool pair_path(const ta_ctx *ctx, const ta_params *p, int32_t la, int32_t lb) {
  if (la <= 0 || lb <= 0 || p->lanes == TA_LANES_LONG) return false;
  if (p->mode == TA_GLOBAL && (int64_t)(la + lb + 2) * -p->gap_extend + 2 * (int64_t)-p->gap_open >= (1ll << 28))
    return false;  // scores near the int32 -infinity: the batch path reports it
  if (p->mode == TA_LOCAL) {  // the kernel's end-cell keys hold H in 16 bits
    const int maxS = p->alphabet == TA_PROTEIN ? 11 : std::max(0, std::max(p->match, p->mismatch));
    if ((int64_t)std::min(la, lb) * maxS >= 65535) return false;
  }
  return pair_layout(la, lb).bytes <= ctx->lds_max;
}

PairArgs pair_args(const ta_params *p, int32_t la, int32_t lb) {
  PairArgs A{};
  A.m = la; A.n = lb;
  A.alphabet = p->alphabet; A.match = p->match; A.mismatch = p->mismatch;
  A.open = p->gap_open; A.ext = p->gap_extend;
  A.local = p->mode == TA_LOCAL; A.traceback = p->traceback != 0;
  A.spin_limit = 1 << 22;
  A.walk_cap = la + lb;  // a walk takes at most la + lb steps
  if (const char *e = getenv("TWEAKALIGN_WALK_GUARD")) A.walk_cap = std::min(A.walk_cap, std::max(0, atoi(e)));
  A.lay = pair_layout(la, lb);
  return A;
}

}  // namespace

int ta_align_pair(ta_ctx *ctx, const ta_params *p, const char *a, int32_t la, const char *b, int32_t lb,
                  ta_result *out, char *cigar_buf, int64_t cap, int64_t *used) {
  if (!ctx || !p || !out || la < 0 || lb < 0 || (la && !a) || (lb && !b)) return TA_E_INVALID;
  if (params_check(p)) return TA_E_INVALID;
  if (!pair_path(ctx, p, la, lb)) {
    const ta_pair q{a, la, b, lb};
    return ta_align_batch(ctx, p, &q, 1, out, cigar_buf, cap, used);
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return TA_E_DEVICE;
  PairArgs A = pair_args(p, la, lb);
  // staging: [PairOut][runs][a][b]
  const size_t o_runs = sizeof(PairOut), o_a = (o_runs + 4 * (size_t)A.lay.runs_cap + 15) & ~(size_t)15;
  const size_t o_b = (o_a + (size_t)la + 15) & ~(size_t)15, need = o_b + (size_t)lb + 16;
  if (ctx->pair_stage_cap < need) {
    if (ctx->pair_stage) (void)hipHostFree(ctx->pair_stage);
    ctx->pair_stage = nullptr;
    ctx->pair_stage_cap = 0;
    const size_t want = std::max(need, (size_t)64 << 10);
    if (hipHostMalloc(&ctx->pair_stage, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      ctx->pair_stage = nullptr;
      return TA_E_NOMEM;
    }
    ctx->pair_stage_cap = want;
  }
  uint8_t *st = (uint8_t *)ctx->pair_stage;
  memcpy(st + o_a, a, (size_t)la);
  memcpy(st + o_b, b, (size_t)lb);
  PairOut *o = (PairOut *)st;
  o->status = 0;
  A.a = (const char *)(st + o_a);
  A.b = (const char *)(st + o_b);
  A.out = o;
  int rc = pair_launch(A, ctx->stream);
  if (rc) return rc;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return TA_E_DEVICE;
  return pair_finish(*o, (const uint32_t *)(st + o_runs), p, out, cigar_buf, cap, used);
}

// ---------------------------------------------------------------------------
// The allocation-free variant.  `scratch` is the only device memory it uses:
// [PairOut][runs][letters of a][letters of b].  In host code its name is a
// host-side stand-in, never a device address: every use goes through the
// address hipGetSymbolAddress gives for the device (pair_scratch_setup).
namespace {

static __device__ __align__(256) uint8_t scratch[TA_PAIR_STATIC_SCRATCH];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

uint8_t *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch[device];
}

// Offsets inside the scratch for an la x lb pair (within its size whenever
// la + lb <= TA_PAIR_STATIC_MAX).
struct StaticLayout {
  size_t runs, a, b, end;
};
StaticLayout static_layout(int32_t la, int32_t lb) {
  StaticLayout L;
  L.runs = sizeof(PairOut);
  L.a = (L.runs + 4 * ((size_t)la + lb + 2) + 15) & ~(size_t)15;
  L.b = (L.a + (size_t)la + 15) & ~(size_t)15;
  L.end = L.b + (size_t)lb;
  return L;
}
static_assert(sizeof(PairOut) + 4 * (TA_PAIR_STATIC_MAX + 2) + 15 + TA_PAIR_STATIC_MAX + 15 <= TA_PAIR_STATIC_SCRATCH,
              "scratch too small for the largest pair it admits");

// Result of a pair with an empty side, as the batch path gives it.
int empty_side(const ta_params *p, int32_t la, int32_t lb, ta_result *out, char *cigar_buf, int64_t cap,
               int64_t *used) {
  *out = ta_result{};
  out->cigar_off = 0;
  if (p->mode == TA_LOCAL) {
    out->a_begin = out->a_end = out->b_begin = out->b_end = -1;
    out->flags = TA_F_EMPTY;
    if (used) *used = 0;
    return TA_OK;
  }
  const int32_t len = la + lb;
  out->score = len ? p->gap_open + (len - 1) * p->gap_extend : 0;
  out->a_begin = 0; out->a_end = la - 1; out->b_begin = 0; out->b_end = lb - 1;
  uint32_t run = len ? ((uint32_t)len << 2) | (la ? 2u : 1u) : 0u;  // la I's or lb D's
  const int64_t need = (p->traceback && len) ? cigar_text(&run, 1, nullptr) : 0;
  out->cigar_len = (int32_t)need;
  if (used) *used = need;
  if (need > 0) {
    if (!cigar_buf || need > cap) return TA_E_SPACE;
    cigar_text(&run, 1, cigar_buf);
  }
  return TA_OK;
}

}  // namespace

void pair_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_done[device]) return;
  g_scratch_done[device] = true;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "tweakalign: no address for the pair scratch on device %d: %s\n", device, hipGetErrorString(e));
    return;
  }
  g_scratch[device] = (uint8_t *)ptr;
}

int ta_align_pair_static(ta_ctx *ctx, const ta_params *p, const char *a, int32_t la, const char *b, int32_t lb,
                         ta_result *out, char *cigar_buf, int64_t cap, int64_t *used) {
  if (!ctx || !p || !out || la < 0 || lb < 0 || (la && !a) || (lb && !b)) return TA_E_INVALID;
  if (params_check(p)) return TA_E_INVALID;
  if ((int64_t)la + lb > TA_PAIR_STATIC_MAX) return TA_E_LIMIT;
  if (la == 0 || lb == 0) return empty_side(p, la, lb, out, cigar_buf, cap, used);
  if (!pair_path(ctx, p, la, lb)) return TA_E_LIMIT;
  uint8_t *scr = scratch_of(ctx->device);
  if (!scr) return TA_E_DEVICE;
  if (hipSetDevice(ctx->device) != hipSuccess) return TA_E_DEVICE;
  const StaticLayout L = static_layout(la, lb);
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(scr + L.a, a, (size_t)la, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(scr + L.b, b, (size_t)lb, hipMemcpyHostToDevice, st) != hipSuccess)
    return TA_E_DEVICE;
  PairArgs A = pair_args(p, la, lb);
  A.a = (const char *)(scr + L.a);
  A.b = (const char *)(scr + L.b);
  A.out = (PairOut *)scr;  // the kernel writes every field the host reads (status included)
  int rc = pair_launch(A, st);
  if (rc) return rc;
  // result and runs to the calling thread's stack
  struct {
    PairOut o;
    uint32_t runs[TA_PAIR_STATIC_MAX + 2];
  } back;
  static_assert(sizeof(PairOut) % 4 == 0, "runs follow the record");
  if (hipMemcpyAsync(&back, scr, sizeof(PairOut) + 4 * ((size_t)la + lb + 2), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return TA_E_DEVICE;
  return pair_finish(back.o, back.runs, p, out, cigar_buf, cap, used);
}
