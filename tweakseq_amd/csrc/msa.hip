// Progressive multiple alignment on the GPU (ta_msa).
//
// tweakseq aligns N > 2 sequences by handing them to an external MSA program
// (tweakseq/UI/SeqEditMainWin.cpp:1647-1659) and reading the aligned FASTA
// back (tweakseq/Core/Project.cpp:908-1032).  This builds the alignment in
// process, to the specification the CPU oracle restates (oracle/msa.c):
//
//  1. global Gotoh scores of all pairs -- one batch of the pairwise engine
//     (n(n-1)/2 pairs, score only: 130,816 pairs for 512 sequences);
//  2. similarities s_ij * 1e6 / max(1, min(self_i, self_j)) (integer), and
//     UPGMA joins on the host with exact integer comparisons (ties to the
//     smallest pair of cluster ids);
//  3. the joins, tree level by tree level: every join of a level is one
//     workgroup of msa_join_kernel, which aligns the columns of the lower
//     cluster's profile (DP rows) with the higher one's (DP columns) by
//     global Gotoh -- match score M(i,j) = cntA[i] . PB[j] with
//     PB[j][x] = sum_y S(x,y) cntB[j][y], gap costs times nA * nB, the
//     pairwise ties (diag > E > F, E / F prefer opening) -- then walks the
//     direction bytes back from (la, lb) to ops; the host applies the ops to
//     the profiles' rows.
//
// Each level first scores every column pair of every join over the whole GPU
// (msa_scores_kernel, 64 x 64-cell tiles; int32: nA * nB is bounded so that
// 24 codes x 11 x nA * nB fits), in the order the sweeps read them.  Then a
// join's workgroup sweeps the DP in the register-strip form of the pairwise
// fill: a strip is
// 64 lanes x JR rows, lane l holds its JR rows' H and E in registers and
// works on column t - l + 1 at step t; the bottom row (H, F) reaches the
// next lane one step later through a wave_shr:1 DPP move, so a wave needs
// no barrier at all.  The workgroup's four waves take strips s, s + 4, ...;
// strip s + 1 reads strip s's last row from an LDS ring once strip s has
// published (an LDS progress counter per strip slot, workgroup scope,
// bounded waits) that it stored the columns it needs.  Values are int32
// where the join's bound allows (most joins), else int64.  Direction bytes
// go out per lane and step, [strip][8-step block][lane][step][row].  The walk
// back runs on thread 0 over 64 x 64 tiles of direction bytes that the
// workgroup stages in LDS, one barrier pair per tile.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <vector>

#include "runtime_internal.h"

namespace {

constexpr int JT = 256;                  // threads per join: 4 waves
constexpr int JW = JT / 64;              // waves per join
constexpr int JRMAX = 8;                 // DP rows per lane at most (a strip: 64 * JR rows)
constexpr int TW = 64;                   // walk tile: TW x TW direction bytes in LDS
constexpr int KS = 24;                   // row stride of counts and PB (codes padded with zeros)
constexpr int64_t WMAX = 8000000;        // nA * nB bound: 24 codes x 11 x nA * nB fits int32

struct JoinJob {
  int32_t la, lb;
  int64_t w;        // nA * nB: gap costs scale
  int64_t cnt_off;  // la x KS: residues of each code in A's columns
  int64_t pb_off;   // lb x KS: PB[j][x] = sum_y S(x, y) cntB[j][y]
  int64_t m_off;    // match scores M(i, j) in the sweep's order: [strip][step][lane][row]
  int64_t tb_off;   // direction bytes src | e << 2 | f << 3, [strip][block][lane][step][row]
  int64_t ops_off;  // la + lb ops, walk order (last column first)
};

struct JoinArgs {
  const JoinJob *jobs;
  const int32_t *cnt, *pb;
  int32_t *mbuf;
  uint8_t *tb;
  uint8_t *ops;
  int32_t *nops;
  int32_t *err;     // | 1: a wait on another wave's progress gave up
  int32_t K, open, ext, spin_limit;
};

// 8-step blocks of a strip of a join with lb columns (lane 63 trails lane 0).
__host__ __device__ inline int join_blocks(int lb) { return (lb + 63 + 7) / 8; }

// Direction byte of cell (i, j), 1-based, in a join's layout.
__device__ __forceinline__ int64_t join_tb_at(int i, int j, int JR, int NB) {
  const int RS = 64 * JR, r = (i - 1) % RS, l = r / JR, t = j - 1 + l;
  return (((((int64_t)((i - 1) / RS) * NB + (t >> 3)) * 64 + l) * 8 + (t & 7)) * JR + (r % JR));
}

// wave_shr:1 (lane l takes lane l - 1's value, lane 0 keeps `old`), for the
// 32- and 64-bit value types.
__device__ __forceinline__ int32_t shr_lane(int32_t old, int32_t v) {
  return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int64_t shr_lane(int64_t old, int64_t v) {
  const int32_t lo = __builtin_amdgcn_update_dpp((int32_t)old, (int32_t)v, 0x138, 0xf, 0xf, false);
  const int32_t hi = __builtin_amdgcn_update_dpp((int32_t)(old >> 32), (int32_t)(v >> 32), 0x138, 0xf, 0xf, false);
  return ((int64_t)hi << 32) | (uint32_t)lo;
}

template <typename T>
struct JoinNeg;
template <>
struct JoinNeg<int32_t> { static constexpr int32_t v = -(1 << 29); };
template <>
struct JoinNeg<int64_t> { static constexpr int64_t v = INT64_MIN / 4; };

// Index of the match score of cell (i + 1, j + 1) (0-based i, j) in the order
// a join's sweep reads them: [strip][step][lane][row], step t = j + lane.
__device__ __forceinline__ int64_t join_m_at(int i, int j, int JR, int NST) {
  const int RS = 64 * JR, l = (i % RS) / JR;
  return (((int64_t)(i / RS) * NST + j + l) * 64 + l) * JR + (i % JR);
}

// PB of B's columns from their residue counts, in place, a thread per
// column: PB[j][x] = sum_y S(x, y) cntB[j][y] (K x K per column, which the
// host used to spend most of a level's preparation on).
__global__ __launch_bounds__(256) void msa_pb_kernel(int32_t *pb, int64_t rows, const int8_t *S, int K) {
  __shared__ int32_t st[KS * KS];  // S(x, y) at [x][y], zero beyond K
  for (int e = threadIdx.x; e < KS * KS; e += blockDim.x) {
    const int x = e / KS, y = e % KS;
    st[e] = (x < K && y < K) ? (int32_t)S[x * K + y] : 0;
  }
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= rows) return;
  int32_t *row = pb + j * KS;
  int32_t c[KS];
#pragma unroll
  for (int y = 0; y < KS; ++y) c[y] = y < K ? row[y] : 0;
  for (int x = 0; x < K; ++x) {
    int32_t v = 0;
#pragma unroll
    for (int y = 0; y < KS; ++y) v += c[y] * st[x * KS + y];
    row[x] = v;
  }
}

// Every join's match scores M(i, j) = cntA[i] . PB[j], tile by tile over
// the whole GPU (a join's own workgroup would compute them on one CU): a
// workgroup takes 64 rows x 64 columns of one join, stages their counts and
// PB rows in LDS, and each thread scores 16 cells.  Tiles: (join, tile index).
__global__ __launch_bounds__(256) void msa_scores_kernel(const JoinJob *jobs, const int2 *tiles, const int32_t *cnt,
                                                         const int32_t *pb, int32_t *mbuf, int JR) {
  __shared__ int32_t ca[64 * KS], pbs[64 * KS];
  const int2 tl = tiles[blockIdx.x];
  const JoinJob J = jobs[tl.x];
  const int ncb = (J.lb + 63) / 64, i0 = (tl.y / ncb) * 64, j0 = (tl.y % ncb) * 64;
  const int ni = J.la - i0 < 64 ? J.la - i0 : 64, nj = J.lb - j0 < 64 ? J.lb - j0 : 64;
  for (int x = threadIdx.x; x < 64 * KS; x += 256) {
    ca[x] = x < ni * KS ? cnt[J.cnt_off + (int64_t)i0 * KS + x] : 0;
    pbs[x] = x < nj * KS ? pb[J.pb_off + (int64_t)j0 * KS + x] : 0;
  }
  __syncthreads();
  const int r = threadIdx.x & 63, NST = 8 * join_blocks(J.lb);
  if (r >= ni) return;
  int32_t a[KS];
#pragma unroll
  for (int x = 0; x < KS; ++x) a[x] = ca[r * KS + x];
  int32_t *mm = mbuf + J.m_off;
  for (int c = threadIdx.x >> 6; c < nj; c += 4) {
    const int32_t *pj = pbs + c * KS;
    int32_t m = 0;
#pragma unroll
    for (int x = 0; x < KS; ++x) m += a[x] * pj[x];
    mm[join_m_at(i0 + r, j0 + c, JR, NST)] = m;
  }
}

// JR: DP rows per lane, the fewest that cover the level's longest row
// profile in one strip per wave where possible; T: int32_t when every join of
// the level keeps its values within +-2^28, else int64_t.
template <int JR, typename T>
__global__ __launch_bounds__(JT) void msa_join_kernel(JoinArgs A) {
  extern __shared__ __align__(16) uint8_t jlds[];  // ring H[lb + 2], F[lb + 2] (T)
  __shared__ uint8_t tile[TW * TW];
  __shared__ int32_t walk[4];                      // i, j, state, finished
  __shared__ int32_t prog[8];                      // strip s publishes ((s + 1) << 20) | columns stored
  constexpr T NEG = JoinNeg<T>::v;
  const JoinJob J = A.jobs[blockIdx.x];
  const int la = J.la, lb = J.lb, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const T go = (T)((int64_t)A.open * J.w), ge = (T)((int64_t)A.ext * J.w);
  constexpr int RS = 64 * JR;  // DP rows per strip
  const int S = (la + RS - 1) / RS, NB = join_blocks(lb);
  T *ringH = (T *)jlds, *ringF = ringH + (lb + 2);
  const int32_t *mm = A.mbuf + J.m_off;
  uint8_t *tb = A.tb + J.tb_off;
  if (tid < 8) prog[tid] = 0;

  // (the match scores were written by msa_scores_kernel, launched before)
  const int NST = 8 * NB;  // steps per strip
  __syncthreads();         // prog

  // ---- the DP, strip by strip on each wave.
  for (int s = wave; s < S; s += JW) {
    const int i0 = s * RS + lane * JR + 1;  // this lane's first row
    T Hl[JR], El[JR];                       // H(i, j - 1), E(i, j - 1)
#pragma unroll
    for (int k = 0; k < JR; ++k) {
      Hl[k] = go + (T)(i0 + k - 1) * ge;  // H(i, 0)
      El[k] = NEG;
    }
    T hd0 = i0 == 1 ? (T)0 : go + (T)(i0 - 2) * ge;  // H(i0 - 1, j - 1) for the next column: H(i0 - 1, 0)
    T Hbot = 0, Fbot = NEG;                           // this lane's last row at its last column
    const bool writer = lane == 63 && s + 1 < S;       // stores the strip's last row for strip s + 1
    // match scores, MD steps ahead (one step ahead left every step waiting
    // on its load); cells outside the pair read slots nobody wrote, and
    // those steps are masked
    constexpr int MD = 4;
    int32_t mring[MD][JR];
    const int32_t *ms = mm + ((int64_t)s * NST * 64 + lane) * JR;
    auto load_m = [&](int t, int32_t (&m)[JR]) {
      const int tt = t < NST ? t : NST - 1;
      const int32_t *src = ms + (int64_t)tt * 64 * JR;
      if constexpr (JR == 2) {
        const int2 v = *(const int2 *)src;
        m[0] = v.x; m[1] = v.y;
      } else {
#pragma unroll
        for (int k = 0; k < JR; k += 4) {
          const int4 v = *(const int4 *)(src + k);
          m[k] = v.x; m[k + 1] = v.y; m[k + 2] = v.z; m[k + 3] = v.w;
        }
      }
    };
#pragma unroll
    for (int d = 0; d < MD; ++d) load_m(d, mring[d]);
    for (int b = 0; b < NB; ++b) {
      // strip s - 1 must have stored the ring up to this block's last column
      if (s > 0) {
        const int need = (s << 20) | min(8 * b + 8, lb);
        int spins = 0;
        while (__hip_atomic_load(&prog[(s - 1) & 7], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need) {
          if (++spins > A.spin_limit) {
            if (lane == 0) atomicOr(A.err, 1);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      // lane 0's top inputs of the block's 8 steps: row 0 or the ring, read
      // once the progress allows (not a dependent LDS read per step)
      T tops[8], topfs[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int c = 8 * b + q + 1;  // lane 0's column at step 8b + q
        if (s == 0) {
          tops[q] = go + (T)(c - 1) * ge;  // H(0, c)
          topfs[q] = NEG;
        } else {
          tops[q] = c <= lb ? ringH[c] : (T)0;
          topfs[q] = c <= lb ? ringF[c] : NEG;
        }
      }
      uint32_t codes[8][JR / 4 > 0 ? JR / 4 : 1];  // direction bytes of the block's steps, packed
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int t = 8 * b + q, j = t - lane + 1;
        int32_t mcur[JR];
#pragma unroll
        for (int k = 0; k < JR; ++k) mcur[k] = mring[q % MD][k];
        load_m(t + MD, mring[q % MD]);
        // lane 0's top: row 0 or the ring; the others: the lane above's bottom row
        const T Hup0 = shr_lane(tops[q], Hbot), Fup0 = shr_lane(topfs[q], Fbot);
        uint32_t bytes = 0;
        if (j >= 1 && j <= lb) {
          T hup = Hup0, fup = Fup0, hd = hd0;
#pragma unroll
          for (int k = 0; k < JR; ++k) {
            const T eo = Hl[k] + go, ee = El[k] + ge;
            const bool eb = ee > eo;
            const T E = eb ? ee : eo;
            const T fo = hup + go, fe = fup + ge;
            const bool fb = fe > fo;
            const T F = fb ? fe : fo;
            T h = hd + (T)mcur[k];
            int src = 0;
            if (E > h) { h = E; src = 1; }
            if (F > h) { h = F; src = 2; }
            hd = Hl[k];
            Hl[k] = h;
            El[k] = E;
            hup = h;
            fup = F;
            bytes |= (uint32_t)(src | ((int)eb << 2) | ((int)fb << 3)) << (8 * (k & 3));
            if ((k & 3) == 3 || k == JR - 1) {
              codes[q][k >> 2] = bytes;
              bytes = 0;
            }
          }
          Hbot = hup;
          Fbot = fup;
          hd0 = Hup0;
          if (writer) {
            ringH[j] = hup;
            ringF[j] = fup;
          }
        } else {
#pragma unroll
          for (int k = 0; k < (JR / 4 > 0 ? JR / 4 : 1); ++k) codes[q][k] = 0;
        }
      }
      // the block's direction bytes: 8 steps x JR rows, contiguous per lane
      uint8_t *dst = tb + ((((int64_t)s * NB + b) * 64 + lane) * 8) * JR;
      if constexpr (JR >= 4) {
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int k = 0; k < JR / 4; ++k) ((uint32_t *)dst)[q * (JR / 4) + k] = codes[q][k];
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) ((uint16_t *)dst)[q] = (uint16_t)codes[q][0];
      }
      // publish the ring columns stored so far (the writer lane's column)
      if (s + 1 < S) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        const int done = min(max(8 * b + 8 - 63, 0), lb);
        if (lane == 63) __hip_atomic_store(&prog[s & 7], ((s + 1) << 20) | done, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
  __syncthreads();

  // ---- walk back from (la, lb): the pairwise state machine (H, E 'D',
  // F 'I').  The workgroup stages the TW x TW tile of direction bytes ending
  // at the walk's cell in LDS; thread 0 walks it until the walk leaves it.
  uint8_t *ops = A.ops + J.ops_off;
  if (tid == 0) { walk[0] = la; walk[1] = lb; walk[2] = 0; walk[3] = 0; }
  int k = 0;  // ops written (thread 0)
  __syncthreads();
  for (;;) {
    const int wi = walk[0], wj = walk[1];
    if (walk[3]) break;
    const int i0 = wi - (TW - 1) > 1 ? wi - (TW - 1) : 1, j0 = wj - (TW - 1) > 1 ? wj - (TW - 1) : 1;
    const int ni = wi - i0 + 1, nj = wj - j0 + 1;
    if (wi >= 1 && wj >= 1)
      for (int c = tid; c < ni * nj; c += JT) {
        const int r = c / nj, q = c - r * nj;
        tile[r * TW + q] = tb[join_tb_at(i0 + r, j0 + q, JR, NB)];
      }
    __syncthreads();
    if (tid == 0) {
      int i = wi, j = wj, state = walk[2];
      bool fin = false;
      for (;;) {
        if (i == 0) { while (j-- > 0) ops[k++] = 'D'; fin = true; break; }
        if (j == 0) { while (i-- > 0) ops[k++] = 'I'; fin = true; break; }
        if (i < i0 || j < j0) break;  // left the tile
        const int code = tile[(i - i0) * TW + (j - j0)];
        if (state == 0) {
          const int s = code & 3;
          if (s == 0) { ops[k++] = 'M'; --i; --j; }
          else state = s;
        } else if (state == 1) {
          ops[k++] = 'D';
          if (!(code & 4)) state = 0;
          --j;
        } else {
          ops[k++] = 'I';
          if (!(code & 8)) state = 0;
          --i;
        }
      }
      walk[0] = i; walk[1] = j; walk[2] = state; walk[3] = fin;
    }
    __syncthreads();
  }
  if (tid == 0) A.nops[blockIdx.x] = k;
}

struct Profile {
  std::vector<int32_t> rows;  // sequence ids
  int32_t len = 0;
  std::vector<int16_t> cell;  // rows x len: code, or -1 for a gap
};

}  // namespace

int ta_msa(ta_ctx *ctx, const ta_params *p, const char *const *seqs, const int32_t *lens, int32_t nseq,
           int32_t *alen, int32_t *cols, int32_t *joins) {
  if (!ctx || !p || !seqs || !lens || nseq <= 0 || !alen || !cols) return TA_E_INVALID;
  if (p->mode != TA_GLOBAL || p->gap_open > 0 || p->gap_extend > 0) return TA_E_INVALID;
  std::vector<int64_t> off(nseq + 1, 0);
  for (int32_t s = 0; s < nseq; ++s) {
    if (lens[s] <= 0 || !seqs[s]) return TA_E_INVALID;
    off[s + 1] = off[s] + lens[s];
  }
  PhaseTrace tr("msa");  // TWEAKALIGN_TRACE=1: host phase times on stderr
  int8_t S[32 * 32];
  int32_t K = 0;
  int rc = ta_matrix(p->alphabet, p->match, p->mismatch, S, &K);
  if (rc) return rc;
  if (K > KS) return TA_E_INVALID;
  int64_t max_abs_s = 1;
  for (int32_t x = 0; x < K * K; ++x) max_abs_s = std::max<int64_t>(max_abs_s, S[x] < 0 ? -S[x] : S[x]);
  std::vector<uint8_t> codes((size_t)off[nseq]);
  for (int32_t s = 0; s < nseq; ++s)
    if ((rc = ta_encode(p->alphabet, seqs[s], lens[s], codes.data() + off[s]))) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return TA_E_DEVICE;
  tr.mark("encode");

  // 1. pair scores on the GPU
  const int32_t nc = 2 * nseq - 1;
  std::vector<int64_t> sum((size_t)nc * nc, 0);
  std::vector<int64_t> self(nseq, 0);
  for (int32_t s = 0; s < nseq; ++s)
    for (int64_t q = off[s]; q < off[s + 1]; ++q) self[s] += S[codes[q] * K + codes[q]];
  if (nseq > 1) {
    std::vector<ta_pair> pairs;
    pairs.reserve((size_t)nseq * (nseq - 1) / 2);
    for (int32_t a = 0; a < nseq; ++a)
      for (int32_t b = a + 1; b < nseq; ++b) pairs.push_back({seqs[a], lens[a], seqs[b], lens[b]});
    ta_params q = *p;
    q.traceback = 0;
    ta_batch *B = nullptr;
    std::vector<ta_result> res(pairs.size());
    rc = batch_create_on(ctx, &q, pairs.data(), (int32_t)pairs.size(), ctx->stream, &B);
    tr.mark("score create");
    if (!rc) rc = ta_batch_run(ctx, B);
    if (!rc) rc = ta_batch_wait(ctx, B);
    tr.mark("score run");
    if (!rc) rc = ta_batch_results(ctx, B, res.data(), nullptr, 0, nullptr);
    if (B) ta_batch_destroy(B);
    if (rc) return rc;
    tr.mark("score results");
    size_t k = 0;
    for (int32_t a = 0; a < nseq; ++a)
      for (int32_t b = a + 1; b < nseq; ++b, ++k) {
        const int64_t d = std::max<int64_t>(1, std::min(self[a], self[b]));
        sum[(size_t)a * nc + b] = sum[(size_t)b * nc + a] = (int64_t)res[k].score * 1000000 / d;
      }
  }

  tr.mark("score sums");
  // 2. UPGMA joins (exact), and each join's tree level.  Row a keeps its
  // best partner b > a (largest mean, ties to the smaller b); the join is the
  // best row's (ties to the smaller a) -- the lexicographically first pair of
  // largest mean.  After a join only rows whose partner died are rescanned;
  // every other row compares its partner with the new cluster.
  std::vector<int32_t> size(nc, 0), level(nc, 0), ja(nseq > 1 ? nseq - 1 : 0), jb(ja.size());
  std::vector<char> live(nc, 0);
  std::vector<int32_t> best(nc, -1);
  for (int32_t s = 0; s < nseq; ++s) { size[s] = 1; live[s] = 1; }
  // mean(a, b) > mean(x, y), exactly
  auto above = [&](int32_t a, int32_t b, int32_t x, int32_t y) {
    return (__int128)sum[(size_t)a * nc + b] * size[x] * size[y] > (__int128)sum[(size_t)x * nc + y] * size[a] * size[b];
  };
  auto rescan = [&](int32_t a, int32_t end) {
    best[a] = -1;
    for (int32_t b = a + 1; b < end; ++b)
      if (live[b] && (best[a] < 0 || above(a, b, a, best[a]))) best[a] = b;
  };
  for (int32_t a = 0; a < nseq; ++a) rescan(a, nseq);
  int32_t top = 0;
  for (int32_t t = 0; t + 1 < nseq; ++t) {
    int32_t ba = -1;
    for (int32_t a = 0; a < nseq + t; ++a)
      if (live[a] && best[a] >= 0 && (ba < 0 || above(a, best[a], ba, best[ba]))) ba = a;
    const int32_t bb = best[ba], c = nseq + t;
    ja[t] = ba; jb[t] = bb;
    if (joins) { joins[2 * t] = ba; joins[2 * t + 1] = bb; }
    size[c] = size[ba] + size[bb];
    level[c] = 1 + std::max(level[ba], level[bb]);
    top = std::max(top, level[c]);
    live[ba] = live[bb] = 0;
    live[c] = 1;
    for (int32_t x = 0; x < c; ++x)
      if (live[x]) sum[(size_t)c * nc + x] = sum[(size_t)x * nc + c] = sum[(size_t)ba * nc + x] + sum[(size_t)bb * nc + x];
    for (int32_t a = 0; a < c; ++a) {
      if (!live[a]) continue;
      if (best[a] == ba || best[a] == bb || best[a] < 0) rescan(a, c + 1);
      else if (above(a, c, a, best[a])) best[a] = c;
    }
  }

  tr.mark("upgma");
  // 3. joins level by level
  std::vector<Profile> prof(nc);
  for (int32_t s = 0; s < nseq; ++s) {
    prof[s].rows = {s};
    prof[s].len = lens[s];
    prof[s].cell.assign(codes.begin() + off[s], codes.begin() + off[s + 1]);
  }
  hipStream_t st = ctx->stream;
  // The substitution table on the device (the PB kernel reads it).
  int8_t *d_S = (int8_t *)device_alloc((size_t)K * K);
  if (!d_S || hipMemcpy(d_S, S, (size_t)K * K, hipMemcpyHostToDevice) != hipSuccess) {
    device_free(d_S);
    return TA_E_DEVICE;
  }
  struct FreeS {
    int8_t *p;
    ~FreeS() { device_free(p); }
  } free_s{d_S};
  // The levels' count tables, kept by the calling thread between levels and
  // calls: fresh multi-MB vectors pay a page fault per 4 KB.
  static thread_local std::vector<int32_t> tl_cnt, tl_pb;
  std::vector<int32_t> &hcnt = tl_cnt, &hpb = tl_pb;
  for (int32_t L = 1; L <= top; ++L) {
    std::vector<int32_t> ts;
    for (int32_t t = 0; t + 1 < nseq; ++t)
      if (level[nseq + t] == L) ts.push_back(t);
    std::vector<JoinJob> jobs(ts.size());
    int64_t tb_total = 0, ops_total = 0, m_total = 0, cnt_total = 0, pb_total = 0;
    int max_la = 0, max_lb = 0;
    bool wide = false;  // some join's DP values need int64
    for (size_t q = 0; q < ts.size(); ++q) {
      max_la = std::max(max_la, prof[ja[ts[q]]].len);
      max_lb = std::max(max_lb, prof[jb[ts[q]]].len);
    }
    // rows per lane: the fewest that give every wave of the join a strip
    const int jr = max_la <= 2 * JT ? 2 : max_la <= 4 * JT ? 4 : JRMAX;
    // sizes and offsets, then each join's counts and PB on the host threads
    for (size_t q = 0; q < ts.size(); ++q) {
      const Profile &PA = prof[ja[ts[q]]], &PB = prof[jb[ts[q]]];
      JoinJob &J = jobs[q];
      J.la = PA.len; J.lb = PB.len;
      J.w = (int64_t)PA.rows.size() * (int64_t)PB.rows.size();
      // joins whose column scores could leave int32
      if (J.w > WMAX || (int64_t)PA.len * PB.len > INT32_MAX) return TA_E_INVALID;
      // int32 DP when every value, and -2^29 plus a gap cost, stays clear of overflow
      const int64_t step = std::max<int64_t>(max_abs_s, (int64_t)-p->gap_open - p->gap_extend) * J.w;
      if ((int64_t)(J.la + J.lb + 2) * step >= (1ll << 28)) wide = true;
      J.cnt_off = cnt_total;
      cnt_total += (int64_t)PA.len * KS;
      J.pb_off = pb_total;
      pb_total += (int64_t)PB.len * KS;
      J.m_off = m_total;  // the sweep's order: strips x steps x 64 lanes x jr rows (16-byte aligned)
      m_total += ((int64_t)(PA.len + 64 * jr - 1) / (64 * jr)) * 8 * join_blocks(PB.len) * 64 * jr;
      J.tb_off = tb_total;
      tb_total += (((int64_t)(PA.len + 64 * jr - 1) / (64 * jr)) * join_blocks(PB.len) * 64 * 8 * jr + 15) & ~15ll;
      J.ops_off = ops_total;
      ops_total += PA.len + PB.len;
    }
    tr.mark("level sizes");
    // each join's tables are cleared by the thread that fills them
    if (hcnt.size() < (size_t)cnt_total) hcnt.resize((size_t)cnt_total);
    if (hpb.size() < (size_t)pb_total) hpb.resize((size_t)pb_total);
    // Column counts of both sides on the host threads, joins dealt in turn;
    // PB (K x K per column of B) is made from B's counts on the GPU.
    const int nthr = (int)std::min<size_t>(16, ts.size());
    parallel_tasks(nthr, [&](int t) {
      for (size_t q = (size_t)t; q < ts.size(); q += (size_t)nthr) {
        const Profile &PA = prof[ja[ts[q]]], &PB = prof[jb[ts[q]]];
        const JoinJob &J = jobs[q];
        int32_t *ca = hcnt.data() + J.cnt_off, *cb = hpb.data() + J.pb_off;
        std::fill(ca, ca + (size_t)PA.len * KS, 0);
        std::fill(cb, cb + (size_t)PB.len * KS, 0);
        for (size_t r = 0; r < PA.rows.size(); ++r)
          for (int32_t i = 0; i < PA.len; ++i) {
            const int c = PA.cell[r * PA.len + i];
            if (c >= 0) ca[(size_t)i * KS + c] += 1;
          }
        for (size_t r = 0; r < PB.rows.size(); ++r)
          for (int32_t j = 0; j < PB.len; ++j) {
            const int c = PB.cell[r * PB.len + j];
            if (c >= 0) cb[(size_t)j * KS + c] += 1;
          }
      }
    });
    tr.mark("level prep");
    const size_t lds = (wide ? sizeof(int64_t) : sizeof(int32_t)) * 2 * (size_t)(max_lb + 2);
    JoinJob *d_jobs = (JoinJob *)device_alloc(sizeof(JoinJob) * jobs.size());
    int32_t *d_cnt = (int32_t *)device_alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, cnt_total));
    int32_t *d_pb = (int32_t *)device_alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, pb_total));
    int32_t *d_m = (int32_t *)device_alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, m_total));
    int32_t *d_err = (int32_t *)device_alloc(sizeof(int32_t));
    uint8_t *d_tb = (uint8_t *)device_alloc((size_t)std::max<int64_t>(1, tb_total));
    uint8_t *d_ops = (uint8_t *)device_alloc((size_t)std::max<int64_t>(1, ops_total));
    int32_t *d_nops = (int32_t *)device_alloc(sizeof(int32_t) * jobs.size());
    std::vector<uint8_t> hops((size_t)std::max<int64_t>(1, ops_total));
    std::vector<int32_t> hnops(jobs.size());
    int32_t herr = 0;
    rc = (d_jobs && d_cnt && d_pb && d_m && d_err && d_tb && d_ops && d_nops) ? TA_OK : TA_E_NOMEM;
    if (!rc && hipMemsetAsync(d_err, 0, sizeof(int32_t), st) != hipSuccess) rc = TA_E_DEVICE;
    if (!rc && (hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JoinJob) * jobs.size(), hipMemcpyHostToDevice, st) ||
                (cnt_total > 0 &&
                 hipMemcpyAsync(d_cnt, hcnt.data(), sizeof(int32_t) * cnt_total, hipMemcpyHostToDevice, st)) ||
                (pb_total > 0 &&
                 hipMemcpyAsync(d_pb, hpb.data(), sizeof(int32_t) * pb_total, hipMemcpyHostToDevice, st))))
      rc = TA_E_DEVICE;
    if (!rc) {
      JoinArgs a;
      a.jobs = d_jobs; a.cnt = d_cnt; a.pb = d_pb; a.mbuf = d_m; a.tb = d_tb; a.ops = d_ops;
      a.nops = d_nops; a.err = d_err;
      a.K = K; a.open = p->gap_open; a.ext = p->gap_extend;  // K <= KS
      a.spin_limit = 1 << 22;
      (void)hipGetLastError();
      // the match scores over the whole GPU, then the joins
      std::vector<int2> tiles;
      for (size_t q = 0; q < jobs.size(); ++q) {
        const int nt = ((jobs[q].la + 63) / 64) * ((jobs[q].lb + 63) / 64);
        for (int t = 0; t < nt; ++t) tiles.push_back(make_int2((int)q, t));
      }
      int2 *d_tiles = (int2 *)device_alloc(sizeof(int2) * tiles.size());
      if (!d_tiles || hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, st))
        rc = TA_E_DEVICE;
      const int64_t pb_rows = pb_total / KS;
      if (!rc && pb_rows > 0)
        hipLaunchKernelGGL(msa_pb_kernel, dim3((unsigned)((pb_rows + 255) / 256)), dim3(256), 0, st, d_pb, pb_rows,
                           d_S, K);
      if (!rc) hipLaunchKernelGGL(msa_scores_kernel, dim3((unsigned)tiles.size()), dim3(256), 0, st, d_jobs, d_tiles, d_cnt,
                                  d_pb, d_m, jr);
      void (*kern)(JoinArgs) =
          wide ? (jr == 2 ? msa_join_kernel<2, int64_t> : jr == 4 ? msa_join_kernel<4, int64_t> : msa_join_kernel<JRMAX, int64_t>)
               : (jr == 2 ? msa_join_kernel<2, int32_t> : jr == 4 ? msa_join_kernel<4, int32_t> : msa_join_kernel<JRMAX, int32_t>);
      if (!rc) hipLaunchKernelGGL(kern, dim3((unsigned)jobs.size()), dim3(JT), lds, st, a);
      if (rc || hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(hops.data(), d_ops, (size_t)ops_total, hipMemcpyDeviceToHost, st) ||
          hipMemcpyAsync(hnops.data(), d_nops, sizeof(int32_t) * jobs.size(), hipMemcpyDeviceToHost, st) ||
          hipMemcpyAsync(&herr, d_err, sizeof(int32_t), hipMemcpyDeviceToHost, st) ||
          hipStreamSynchronize(st))
        rc = TA_E_DEVICE;
      (void)hipStreamSynchronize(st);
      device_free(d_tiles);
      if (!rc && herr) {
        fprintf(stderr, "tweakalign: a wait on another wave's progress gave up (msa join)\n");
        rc = TA_E_DEVICE;
      }
    }
    device_free(d_jobs); device_free(d_cnt); device_free(d_pb); device_free(d_m); device_free(d_err);
    device_free(d_tb); device_free(d_ops); device_free(d_nops);
    if (rc) return rc;
    tr.mark("level gpu");
    // apply each join's ops (walk order: read backwards)
    for (size_t q = 0; q < ts.size(); ++q) {
      const int32_t t = ts[q], c = nseq + t;
      Profile &PA = prof[ja[t]], &PB = prof[jb[t]];
      const int32_t n = hnops[q];
      const uint8_t *o = hops.data() + jobs[q].ops_off;
      Profile C;
      C.rows = PA.rows;
      C.rows.insert(C.rows.end(), PB.rows.begin(), PB.rows.end());
      C.len = n;
      C.cell.assign(C.rows.size() * (size_t)n, -1);
      const size_t na = PA.rows.size(), nb = PB.rows.size();
      int32_t ia = 0, jbb = 0;
      for (int32_t x = 0; x < n; ++x) {
        const uint8_t op = o[n - 1 - x];
        if (op != 'D') {
          for (size_t r = 0; r < na; ++r) C.cell[r * n + x] = PA.cell[r * PA.len + ia];
          ++ia;
        }
        if (op != 'I') {
          for (size_t r = 0; r < nb; ++r) C.cell[(na + r) * n + x] = PB.cell[r * PB.len + jbb];
          ++jbb;
        }
      }
      prof[c] = std::move(C);
      PA = Profile();
      PB = Profile();
    }
    tr.mark("level apply");
  }

  // 4. every residue's column
  const Profile &F = prof[nc - 1];
  *alen = F.len;
  for (size_t r = 0; r < F.rows.size(); ++r) {
    const int32_t s = F.rows[r];
    int64_t k = off[s];
    for (int32_t x = 0; x < F.len; ++x)
      if (F.cell[r * F.len + x] >= 0) cols[k++] = x;
  }
  return TA_OK;
}
