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
// A join first computes every column pair's match score (int32: nA * nB is
// bounded so that 24 codes x 11 x nA * nB fits; a thread's rows' counts sit
// in registers, each column's PB row is a wave-uniform scalar load), then sweeps
// anti-diagonals: thread t owns rows t+1, t+1+256, ... (2, 4 or 8 each;
// longer row profiles go in strips of 2,048 rows, each reading the last row
// of the one before from global memory), keeps H and E of its
// rows' last cells in registers, and reads the row above (H, F at the
// previous diagonal, H at the one before) from a three-slot LDS ring of int64
// values, one barrier per diagonal.  The walk back runs on thread 0 over
// 64 x 64 tiles of direction bytes that the workgroup stages in LDS, one
// barrier pair per tile instead of a global load per step.  Scores and
// direction bytes are diagonal-major, so one diagonal's rows are adjacent.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "runtime_internal.h"

namespace {

constexpr int JT = 256;                  // threads per join
constexpr int JRMAX = 8;                 // DP rows per thread at most (a strip: JT * JR rows)
constexpr int TW = 64;                   // walk tile: TW x TW direction bytes in LDS
constexpr int KS = 24;                   // row stride of counts and PB (codes padded with zeros)
constexpr int64_t JNEG = INT64_MIN / 4;  // -inf
constexpr int64_t WMAX = 8000000;        // nA * nB bound: 24 codes x 11 x nA * nB fits int32

struct JoinJob {
  int32_t la, lb;
  int64_t w;        // nA * nB: gap costs scale
  int64_t cnt_off;  // la x KS: residues of each code in A's columns
  int64_t pb_off;   // lb x KS: PB[j][x] = sum_y S(x, y) cntB[j][y]
  int64_t m_off;    // match scores M(i, j), diagonal-major: [i + j - 2][i - 1], (la + lb - 1) x la
  int64_t tb_off;   // direction bytes src | e << 2 | f << 3, diagonal-major: [i + j][i], (la + lb + 1) x (la + 1)
  int64_t ops_off;  // la + lb ops, walk order (last column first)
  int64_t hb_off;   // 4 x (lb + 1) int64: a strip's last row (H, F), two strips' worth
};

struct JoinArgs {
  const JoinJob *jobs;
  const int32_t *cnt, *pb;
  int32_t *mbuf;
  int64_t *hbot;
  uint8_t *tb;
  uint8_t *ops;
  int32_t *nops;
  int32_t K, open, ext;
};

// JR: DP rows per thread, the fewest that cover the level's longest row
// profile (every unused row slot costs its checks on every diagonal).
template <int JR>
__global__ __launch_bounds__(JT) void msa_join_kernel(JoinArgs A) {
  extern __shared__ int64_t ring[];  // H[3][la + 1], then F[3][la + 1]
  __shared__ uint8_t tile[TW * TW];
  __shared__ int32_t walk[4];        // i, j, state, finished
  const JoinJob J = A.jobs[blockIdx.x];
  const int la = J.la, lb = J.lb, tid = threadIdx.x;
  const int64_t go = (int64_t)A.open * J.w, ge = (int64_t)A.ext * J.w;
  constexpr int RS = JT * JR;  // DP rows per strip
  const int rs = (la < RS ? la : RS) + 1;
  int64_t *Hr = ring, *Fr = ring + 3 * rs;
  const int32_t *cnt = A.cnt + J.cnt_off, *pb = A.pb + J.pb_off;
  int32_t *mm = A.mbuf + J.m_off;
  uint8_t *tb = A.tb + J.tb_off;
  // Both are diagonal-major so that the rows of one diagonal (consecutive
  // threads) read and write consecutive addresses.
  const int64_t ms = la, ts = la + 1;

  // ---- match scores: thread t scores its rows i = t, t + JT, ... (counts
  // in registers) against every column, whose PB row is the same for the
  // whole wave (scalar loads); the stores of a column are adjacent.
  for (int i = tid; i < la; i += JT) {
    int32_t a[KS];
#pragma unroll
    for (int x = 0; x < KS; ++x) a[x] = cnt[(int64_t)i * KS + x];
    for (int j = 0; j < lb; ++j) {
      const int32_t *pj = pb + (int64_t)j * KS;
      int32_t m = 0;
#pragma unroll
      for (int x = 0; x < KS; ++x) m += a[x] * pj[x];
      mm[(int64_t)(i + j) * ms + i] = m;
    }
  }
  __syncthreads();

  // ---- the DP in strips of RS rows (one, unless the rows outnumber the
  // threads' registers), one anti-diagonal d = i + j per barrier.  A strip's
  // first row reads the previous strip's last row from global memory (two
  // buffers by strip parity: a one-row strip reads and writes the same row).
  int64_t *hb = A.hbot + J.hb_off;
  for (int row0 = 0, sp = 0; row0 < la; row0 += RS, sp ^= 1) {
    const int rend = la < row0 + RS ? la : row0 + RS;  // rows row0 + 1 .. rend
    const int64_t *Hin = hb + (int64_t)(sp ^ 1) * 2 * (lb + 1), *Fin = Hin + (lb + 1);
    int64_t *Hout = hb + (int64_t)sp * 2 * (lb + 1), *Fout = Hout + (lb + 1);
    int64_t Hl[JR], El[JR];  // H(i, j-1), E(i, j-1) of each owned row
    int32_t mn[JR];          // match scores of the next diagonal's cells, loaded a diagonal early
#pragma unroll
    for (int r = 0; r < JR; ++r) {
      const int i = row0 + tid + 1 + r * JT;
      Hl[r] = go + (int64_t)(i - 1) * ge;  // H(i, 0)
      El[r] = JNEG;
      mn[r] = (i <= rend && row0 + 2 - i >= 1) ? mm[(int64_t)row0 * ms + (i - 1)] : 0;
    }
    for (int d = row0 + 2; d <= rend + lb; ++d) {
      const int cur = d % 3, prv = (d + 2) % 3, pp = (d + 1) % 3;  // diagonals d, d-1, d-2
      int32_t mc[JR];
#pragma unroll
      for (int r = 0; r < JR; ++r) {
        const int i = row0 + tid + 1 + r * JT, j1 = d + 1 - i;
        mc[r] = mn[r];
        mn[r] = (i <= rend && j1 >= 1 && j1 <= lb) ? mm[(int64_t)(d - 1) * ms + (i - 1)] : 0;
      }
      // Loads of every owned row first (all in flight together), then the
      // cells: a row's ring stores would otherwise order the next row's loads.
      int64_t Hup[JR], Fup[JR], Hdg[JR];
#pragma unroll
      for (int r = 0; r < JR; ++r) {
        const int li = tid + 1 + r * JT, i = row0 + li;  // row in the strip, in the matrix
        const int j = d - i;
        Hup[r] = Fup[r] = Hdg[r] = 0;
        if (i > rend || j < 1 || j > lb) continue;
        if (i == 1) {  // row 0 is the boundary
          Hup[r] = go + (int64_t)(j - 1) * ge;
          Fup[r] = JNEG;
          Hdg[r] = j == 1 ? 0 : go + (int64_t)(j - 2) * ge;
        } else if (li == 1) {  // the previous strip's last row
          Hup[r] = Hin[j];
          Fup[r] = Fin[j];
          Hdg[r] = j == 1 ? go + (int64_t)(i - 2) * ge : Hin[j - 1];
        } else {
          Hup[r] = Hr[prv * rs + li - 1];
          Fup[r] = Fr[prv * rs + li - 1];
          Hdg[r] = j == 1 ? go + (int64_t)(i - 2) * ge : Hr[pp * rs + li - 1];
        }
      }
#pragma unroll
      for (int r = 0; r < JR; ++r) {
        const int li = tid + 1 + r * JT, i = row0 + li;
        const int j = d - i;
        if (i > rend || j < 1 || j > lb) continue;
        const int64_t eo = Hl[r] + go, ee = El[r] + ge;
        const bool eb = ee > eo;
        const int64_t E = eb ? ee : eo;
        const int64_t fo = Hup[r] + go, fe = Fup[r] + ge;
        const bool fb = fe > fo;
        const int64_t F = fb ? fe : fo;
        int64_t h = Hdg[r] + mc[r];
        int src = 0;
        if (E > h) { h = E; src = 1; }
        if (F > h) { h = F; src = 2; }
        Hr[cur * rs + li] = h;
        Fr[cur * rs + li] = F;
        if (i == rend) { Hout[j] = h; Fout[j] = F; }
        Hl[r] = h;
        El[r] = E;
        tb[(int64_t)d * ts + i] = (uint8_t)(src | ((int)eb << 2) | ((int)fb << 3));
      }
      // The ring is the only memory shared within a diagonal: an LDS-only
      // barrier leaves the next diagonal's score loads in flight across it
      // (the direction bytes and a strip's last row are ordered by the full
      // barriers after the strip).
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
    __syncthreads();
  }

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
    const int i0 = wi - (TW - 1) > 0 ? wi - (TW - 1) : 0, j0 = wj - (TW - 1) > 0 ? wj - (TW - 1) : 0;
    const int ni = wi - i0 + 1, nj = wj - j0 + 1;
    for (int c = tid; c < ni * nj; c += JT) {
      const int r = c / nj, q = c - r * nj;
      tile[r * TW + q] = tb[(int64_t)(i0 + r + j0 + q) * ts + (i0 + r)];
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
  int8_t S[32 * 32];
  int32_t K = 0;
  int rc = ta_matrix(p->alphabet, p->match, p->mismatch, S, &K);
  if (rc) return rc;
  if (K > KS) return TA_E_INVALID;
  std::vector<uint8_t> codes((size_t)off[nseq]);
  for (int32_t s = 0; s < nseq; ++s)
    if ((rc = ta_encode(p->alphabet, seqs[s], lens[s], codes.data() + off[s]))) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return TA_E_DEVICE;

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
    if (!rc) rc = ta_batch_run(ctx, B);
    if (!rc) rc = ta_batch_wait(ctx, B);
    if (!rc) rc = ta_batch_results(ctx, B, res.data(), nullptr, 0, nullptr);
    if (B) ta_batch_destroy(B);
    if (rc) return rc;
    size_t k = 0;
    for (int32_t a = 0; a < nseq; ++a)
      for (int32_t b = a + 1; b < nseq; ++b, ++k) {
        const int64_t d = std::max<int64_t>(1, std::min(self[a], self[b]));
        sum[(size_t)a * nc + b] = sum[(size_t)b * nc + a] = (int64_t)res[k].score * 1000000 / d;
      }
  }

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

  // 3. joins level by level
  std::vector<Profile> prof(nc);
  for (int32_t s = 0; s < nseq; ++s) {
    prof[s].rows = {s};
    prof[s].len = lens[s];
    prof[s].cell.assign(codes.begin() + off[s], codes.begin() + off[s + 1]);
  }
  hipStream_t st = ctx->stream;
  for (int32_t L = 1; L <= top; ++L) {
    std::vector<int32_t> ts;
    for (int32_t t = 0; t + 1 < nseq; ++t)
      if (level[nseq + t] == L) ts.push_back(t);
    std::vector<JoinJob> jobs(ts.size());
    std::vector<int32_t> hcnt, hpb;
    int64_t tb_total = 0, ops_total = 0, m_total = 0, cnt_total = 0, pb_total = 0, hb_total = 0;
    int max_la = 0;
    // sizes and offsets, then each join's counts and PB on the host threads
    for (size_t q = 0; q < ts.size(); ++q) {
      const Profile &PA = prof[ja[ts[q]]], &PB = prof[jb[ts[q]]];
      JoinJob &J = jobs[q];
      J.la = PA.len; J.lb = PB.len;
      J.w = (int64_t)PA.rows.size() * (int64_t)PB.rows.size();
      // joins whose column scores could leave int32
      if (J.w > WMAX || (int64_t)PA.len * (PA.len + PB.len + 1) > INT32_MAX) return TA_E_INVALID;
      J.cnt_off = cnt_total;
      cnt_total += (int64_t)PA.len * KS;
      J.pb_off = pb_total;
      pb_total += (int64_t)PB.len * KS;
      J.m_off = m_total;
      m_total += (int64_t)(PA.len + PB.len - 1) * PA.len;
      J.tb_off = tb_total;
      tb_total += (int64_t)(PA.len + PB.len + 1) * (PA.len + 1);
      J.ops_off = ops_total;
      ops_total += PA.len + PB.len;
      J.hb_off = hb_total;
      hb_total += 4 * (int64_t)(PB.len + 1);
      max_la = std::max(max_la, PA.len);
    }
    hcnt.assign((size_t)cnt_total, 0);
    hpb.assign((size_t)pb_total, 0);
    parallel_for(ts.size(), [&](size_t q) {
      const Profile &PA = prof[ja[ts[q]]], &PB = prof[jb[ts[q]]];
      const JoinJob &J = jobs[q];
      int32_t *ca = hcnt.data() + J.cnt_off;
      for (size_t r = 0; r < PA.rows.size(); ++r)
        for (int32_t i = 0; i < PA.len; ++i) {
          const int c = PA.cell[r * PA.len + i];
          if (c >= 0) ca[(size_t)i * KS + c] += 1;
        }
      // PB from B's column counts: K x K per column instead of K per residue
      int32_t *pbj = hpb.data() + J.pb_off;
      std::vector<int32_t> cb((size_t)PB.len * KS, 0);
      for (size_t r = 0; r < PB.rows.size(); ++r)
        for (int32_t j = 0; j < PB.len; ++j) {
          const int c = PB.cell[r * PB.len + j];
          if (c >= 0) cb[(size_t)j * KS + c] += 1;
        }
      for (int32_t j = 0; j < PB.len; ++j)
        for (int32_t y = 0; y < K; ++y) {
          const int32_t n = cb[(size_t)j * KS + y];
          if (!n) continue;
          for (int32_t x = 0; x < K; ++x) pbj[(size_t)j * KS + x] += n * S[x * K + y];
        }
    });
    const int jr = max_la <= 2 * JT ? 2 : max_la <= 4 * JT ? 4 : JRMAX;
    const size_t lds = sizeof(int64_t) * 6 * (size_t)(std::min(max_la, JT * jr) + 1);
    JoinJob *d_jobs = (JoinJob *)device_alloc(sizeof(JoinJob) * jobs.size());
    int32_t *d_cnt = (int32_t *)device_alloc(sizeof(int32_t) * std::max<size_t>(1, hcnt.size()));
    int32_t *d_pb = (int32_t *)device_alloc(sizeof(int32_t) * std::max<size_t>(1, hpb.size()));
    int32_t *d_m = (int32_t *)device_alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, m_total));
    int64_t *d_hb = (int64_t *)device_alloc(sizeof(int64_t) * (size_t)std::max<int64_t>(1, hb_total));
    uint8_t *d_tb = (uint8_t *)device_alloc((size_t)std::max<int64_t>(1, tb_total));
    uint8_t *d_ops = (uint8_t *)device_alloc((size_t)std::max<int64_t>(1, ops_total));
    int32_t *d_nops = (int32_t *)device_alloc(sizeof(int32_t) * jobs.size());
    std::vector<uint8_t> hops((size_t)std::max<int64_t>(1, ops_total));
    std::vector<int32_t> hnops(jobs.size());
    rc = (d_jobs && d_cnt && d_pb && d_m && d_hb && d_tb && d_ops && d_nops) ? TA_OK : TA_E_NOMEM;
    if (!rc && (hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JoinJob) * jobs.size(), hipMemcpyHostToDevice, st) ||
                (!hcnt.empty() &&
                 hipMemcpyAsync(d_cnt, hcnt.data(), sizeof(int32_t) * hcnt.size(), hipMemcpyHostToDevice, st)) ||
                (!hpb.empty() &&
                 hipMemcpyAsync(d_pb, hpb.data(), sizeof(int32_t) * hpb.size(), hipMemcpyHostToDevice, st))))
      rc = TA_E_DEVICE;
    if (!rc) {
      JoinArgs a;
      a.jobs = d_jobs; a.cnt = d_cnt; a.pb = d_pb; a.mbuf = d_m; a.hbot = d_hb; a.tb = d_tb; a.ops = d_ops;
      a.nops = d_nops;
      a.K = K; a.open = p->gap_open; a.ext = p->gap_extend;  // K <= KS
      (void)hipGetLastError();
      void (*kern)(JoinArgs) = jr == 2 ? msa_join_kernel<2> : jr == 4 ? msa_join_kernel<4> : msa_join_kernel<JRMAX>;
      hipLaunchKernelGGL(kern, dim3((unsigned)jobs.size()), dim3(JT), lds, st, a);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(hops.data(), d_ops, (size_t)ops_total, hipMemcpyDeviceToHost, st) ||
          hipMemcpyAsync(hnops.data(), d_nops, sizeof(int32_t) * jobs.size(), hipMemcpyDeviceToHost, st) ||
          hipStreamSynchronize(st))
        rc = TA_E_DEVICE;
    }
    device_free(d_jobs); device_free(d_cnt); device_free(d_pb); device_free(d_m); device_free(d_hb);
    device_free(d_tb); device_free(d_ops); device_free(d_nops);
    if (rc) return rc;
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
