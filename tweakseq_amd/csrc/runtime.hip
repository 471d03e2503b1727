// Host runtime + C ABI of the alignment engine (include/tweakalign.h).
//
// A batch is encoded once on the host, split into DP tasks (packed int16
// couples where the score range allows, int32 singles otherwise), sorted
// longest-first and uploaded; a run is: persistent fill kernels pulling tasks
// off a device counter, an optional int32 rerun of pairs whose int16 pass
// saturated, and the traceback kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <atomic>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/tweakalign.h"
#include "blosum62.h"
#include "common.h"
#include "fill.h"
#include "long.h"
#include "runtime_internal.h"
#include "traceback.h"

using namespace ta;

// ---------------------------------------------------------------------------
// Alphabets.  The protein table is BLOSUM62 in the row order
// A R N D C Q E G H I L K M F P S T W Y V B Z X and the letter map of
// tweakseq/Core/Annotations/Consensus.cpp:34-69 (same numbers, by necessity).
namespace {

const char kProtOrder[] = TA_PROT_ORDER;
const int8_t kBlosumLower[] = {TA_BLOSUM62_LOWER};
constexpr int kProtK = TA_PROT_K;

struct Alphabet {
  int K = 0;
  int8_t S[32 * 32];  // K x K
  int8_t map[256];    // byte -> code, -1 outside the alphabet
  int maxS = 0, minS = 0;
};

Alphabet make_alphabet(int alphabet, int match, int mismatch) {
  Alphabet A;
  memset(A.map, -1, sizeof(A.map));
  if (alphabet == TA_PROTEIN) {
    A.K = kProtK;
    int idx = 0;
    for (int r = 0; r < kProtK; ++r)
      for (int c = 0; c <= r; ++c, ++idx) {
        A.S[r * kProtK + c] = kBlosumLower[idx];
        A.S[c * kProtK + r] = kBlosumLower[idx];
      }
    for (int r = 0; r < kProtK; ++r) {
      A.map[(unsigned char)kProtOrder[r]] = (int8_t)r;
      A.map[(unsigned char)(kProtOrder[r] + 32)] = (int8_t)r;
    }
    for (char c : {'J', 'O', 'U'}) {  // no row of their own: scored as X
      A.map[(unsigned char)c] = 22;
      A.map[(unsigned char)(c + 32)] = 22;
    }
  } else {
    A.K = 4;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) A.S[r * 4 + c] = (int8_t)(r == c ? match : mismatch);
    const char *acgt = "ACGT";
    for (int r = 0; r < 4; ++r) {
      A.map[(unsigned char)acgt[r]] = (int8_t)r;
      A.map[(unsigned char)(acgt[r] + 32)] = (int8_t)r;
    }
  }
  A.maxS = -128;
  A.minS = 127;
  for (int x = 0; x < A.K * A.K; ++x) {
    A.maxS = std::max<int>(A.maxS, A.S[x]);
    A.minS = std::min<int>(A.minS, A.S[x]);
  }
  return A;
}

// Score-only packed couples keep 4-bit scores: the table must span at most
// 16 values, and for local couples the pad code (scored as the minimum) must
// be negative so that leading pad columns hold H at 0.  Other score-only
// tables run in int32.
bool nibble_ok(const Alphabet &A, bool local) { return A.maxS - A.minS <= 15 && (!local || A.minS < 0); }

// (K+1) x (K+1) table of score - minS (0..15), pad row / column at 0.
std::vector<int8_t> nibble_table(const Alphabet &A) {
  const int Kp = A.K + 1;
  std::vector<int8_t> t((size_t)Kp * Kp, 0);
  for (int r = 0; r < A.K; ++r)
    for (int c = 0; c < A.K; ++c) t[(size_t)r * Kp + c] = (int8_t)(A.S[r * A.K + c] - A.minS);
  return t;
}

// (K+1) x (K+1) table with a pad row / column (index K) of PAD_SCORE.
std::vector<int8_t> padded_table(const Alphabet &A) {
  const int Kp = A.K + 1;
  std::vector<int8_t> t((size_t)Kp * Kp, PAD_SCORE);
  for (int r = 0; r < A.K; ++r)
    for (int c = 0; c < A.K; ++c) t[(size_t)r * Kp + c] = A.S[r * A.K + c];
  return t;
}

#define HIPCHK(x)                                                   \
  do {                                                              \
    hipError_t e_ = (x);                                            \
    if (e_ != hipSuccess) {                                         \
      fprintf(stderr, "tweakalign: %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return TA_E_DEVICE;                                           \
    }                                                               \
  } while (0)

// Device memory cache.  Blocks of every size are kept per size class and
// reused: hipFree waits for the whole device, so freeing a batch's planes
// while other batches run would stall them all (the serving loop keeps
// several batches in flight), and hipMalloc per buffer would dominate small
// batches.  Classes are powers of two up to 1 MB, then four per doubling
// (at most 19% rounding).  A request takes a free block of its class or of
// the next classes up to twice its size.  Free blocks stay cached up to half
// the device's memory; past that, and when hipMalloc fails, cached blocks
// are released.
struct MemCache {
  std::mutex mu;
  std::map<std::pair<int, size_t>, std::vector<void *>> free;  // (device, class) -> blocks
  std::unordered_map<void *, std::pair<int, size_t>> live;
  std::map<int, size_t> cached, cap;                              // per device: free bytes, their limit
};
MemCache &memcache() {
  static MemCache *c = new MemCache();  // lives until the process ends
  return *c;
}

size_t size_class(size_t n) {
  size_t c = 256;
  while (c < n && c < (1u << 20)) c <<= 1;
  if (c >= n) return c;
  while (c * 2 < n) c <<= 1;  // c < n <= 2c: classes c * (5, 6, 7, 8) / 4
  for (int q = 5; q <= 8; ++q)
    if (c / 4 * q >= n) return c / 4 * q;
  return 2 * c;
}

// Releases every cached free block of `dev` (lock held).
void cache_release(MemCache &mc, int dev) {
  for (auto it = mc.free.begin(); it != mc.free.end(); ++it) {
    if (it->first.first != dev) continue;
    for (void *p : it->second) (void)hipFree(p);
    it->second.clear();
  }
  mc.cached[dev] = 0;
}

void *cache_alloc(size_t bytes) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  MemCache &mc = memcache();
  const size_t cls = size_class(bytes);
  std::lock_guard<std::mutex> g(mc.mu);
  if (!mc.cap.count(dev)) {
    size_t fr = 0, tot = 0;
    mc.cap[dev] = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 2 : (size_t)64 << 30;
  }
  for (auto it = mc.free.lower_bound({dev, cls}); it != mc.free.end() && it->first.first == dev &&
                                                  it->first.second <= 2 * cls;
       ++it) {
    if (it->second.empty()) continue;
    void *p = it->second.back();
    it->second.pop_back();
    mc.cached[dev] -= it->first.second;
    mc.live[p] = it->first;
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, cls) != hipSuccess) {
    (void)hipGetLastError();
    cache_release(mc, dev);
    if (hipMalloc(&p, cls) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  mc.live[p] = {dev, cls};
  return p;
}

void cache_free(void *p) {
  if (!p) return;
  MemCache &mc = memcache();
  std::lock_guard<std::mutex> g(mc.mu);
  auto it = mc.live.find(p);
  if (it == mc.live.end()) return;
  const auto key = it->second;
  mc.live.erase(it);
  mc.free[key].push_back(p);
  mc.cached[key.first] += key.second;
  if (mc.cached[key.first] > mc.cap[key.first]) cache_release(mc, key.first);
}

template <typename T>
int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return TA_OK;
  *p = (T *)cache_alloc(n * sizeof(T));
  return *p ? TA_OK : TA_E_NOMEM;
}

template <typename T>
int upload(T **dst, const std::vector<T> &src, hipStream_t st) {
  int rc = dalloc(dst, src.size());
  if (rc) return rc;
  if (!src.empty()) HIPCHK(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return TA_OK;
}

template <typename T>
void dfree(T *&p) {
  cache_free((void *)p);
  p = nullptr;
}

// Host staging for the residue codes, the results and the CIGAR runs.  A
// copy to or from pageable memory runs at a fraction of the link's rate (the
// runtime copies it through bounce buffers); with pinned memory the DMA
// engine reads or writes the block itself.  Pinning a block takes
// milliseconds, so blocks are kept and reused: batch creation takes one and
// encodes into it, result fetching takes two and reads from them, and each
// goes back once its copy has finished.  At most 16 free blocks and 1 GB stay
// pinned; blocks above 256 MB come from the heap (pinning them would cost
// about what it saves).
struct Staging {
  uint8_t *p = nullptr;
  size_t cap = 0;
  bool pinned = false;
};
struct StagePool {
  std::mutex mu;
  std::vector<Staging> free;
};
StagePool &stage_pool() {
  static StagePool *s = new StagePool();  // lives until the process ends
  return *s;
}

Staging stage_take(size_t bytes) {
  StagePool &sp = stage_pool();
  {
    std::lock_guard<std::mutex> g(sp.mu);
    int best = -1;  // the smallest free block that fits
    for (int i = 0; i < (int)sp.free.size(); ++i)
      if (sp.free[i].cap >= bytes && (best < 0 || sp.free[i].cap < sp.free[best].cap)) best = i;
    if (best >= 0) {
      const Staging s = sp.free[best];
      sp.free.erase(sp.free.begin() + best);
      return s;
    }
  }
  Staging s;
  s.cap = size_class(std::max(bytes, (size_t)1 << 20));
  if (s.cap <= ((size_t)256 << 20)) {
    if (hipHostMalloc((void **)&s.p, s.cap, hipHostMallocPortable) == hipSuccess) {
      s.pinned = true;
      return s;
    }
    (void)hipGetLastError();
    s.p = nullptr;
  }
  s.p = (uint8_t *)malloc(s.cap);
  if (!s.p) s.cap = 0;
  return s;
}

void stage_give(Staging &s) {
  if (!s.p) return;
  if (!s.pinned) {
    free(s.p);
  } else {
    StagePool &sp = stage_pool();
    std::lock_guard<std::mutex> g(sp.mu);
    sp.free.push_back(s);
    auto by_cap = [](const Staging &x, const Staging &y) { return x.cap < y.cap; };
    size_t bytes = 0;
    for (const Staging &f : sp.free) bytes += f.cap;
    while (sp.free.size() > 16 || bytes > ((size_t)1 << 30)) {  // unpin the smallest, or the largest past 1 GB
      auto it = sp.free.size() > 16 ? std::min_element(sp.free.begin(), sp.free.end(), by_cap)
                                    : std::max_element(sp.free.begin(), sp.free.end(), by_cap);
      bytes -= it->cap;
      (void)hipHostFree(it->p);
      sp.free.erase(it);
    }
  }
  s = Staging{};
}

// A staging block held by one batch creation.  Once a copy from it is
// queued on `stream`, the block goes back only after the stream has drained
// (every return path, device errors included).
struct StageHold {
  Staging s;
  hipStream_t stream = nullptr;
  bool queued = false;
  ~StageHold() {
    if (queued) (void)hipStreamSynchronize(stream);
    stage_give(s);
  }
};

// Host vectors of destroyed batches, kept for the next batches so that
// their pages stay mapped (a fresh multi-MB vector pays a page fault per
// 4 KB, milliseconds per batch): up to 4 per element type, none above 64 MB.
template <typename T>
struct VecBin {
  std::mutex mu;
  std::vector<std::vector<T>> v;
  static VecBin &get() {
    static VecBin *b = new VecBin();
    return *b;
  }
};
template <typename T>
void vec_take(std::vector<T> &x) {
  VecBin<T> &b = VecBin<T>::get();
  std::lock_guard<std::mutex> g(b.mu);
  if (!b.v.empty()) {
    x.swap(b.v.back());
    b.v.pop_back();
  }
  x.clear();
}
template <typename T>
void vec_give(std::vector<T> &x) {
  const size_t bytes = x.capacity() * sizeof(T);
  if (bytes < ((size_t)64 << 10) || bytes > ((size_t)64 << 20)) return;
  VecBin<T> &b = VecBin<T>::get();
  std::lock_guard<std::mutex> g(b.mu);
  if (b.v.size() < 4) {
    b.v.emplace_back();
    b.v.back().swap(x);
  }
}

// Host vectors bound for their own device buffers, gathered into one pinned
// staging block and sent with one asynchronous copy each (a copy from
// pageable memory is staged by the runtime and waits for the link).  The
// vectors must live until commit(); the block until the stream drains.
struct BulkUpload {
  struct Item { void **dst; const void *src; size_t bytes, at; };
  std::vector<Item> items;
  StageHold hold;
  template <typename T>
  void add(T **dst, const std::vector<T> &v) {
    *dst = nullptr;
    items.push_back(Item{(void **)dst, v.data(), v.size() * sizeof(T), 0});
  }
  int commit(hipStream_t st) {
    size_t total = 0;
    for (Item &x : items) {
      x.at = total;
      total += (x.bytes + 255) & ~(size_t)255;
    }
    if (!total) return TA_OK;
    hold.s = stage_take(total);
    if (!hold.s.p) return TA_E_NOMEM;
    for (Item &x : items)
      if (x.bytes) memcpy(hold.s.p + x.at, x.src, x.bytes);
    hold.stream = st;
    for (Item &x : items) {
      if (!x.bytes) continue;
      if (!(*x.dst = cache_alloc(x.bytes))) return TA_E_NOMEM;
      hold.queued = true;
      if (hipMemcpyAsync(*x.dst, hold.s.p + x.at, x.bytes, hipMemcpyHostToDevice, st) != hipSuccess)
        return TA_E_DEVICE;
    }
    return TA_OK;
  }
};


}  // namespace

// The device-memory cache for the library's other sources (runtime_internal.h).
void *device_alloc(size_t bytes) { return cache_alloc(bytes); }
void device_free(void *p) { cache_free(p); }

// ---------------------------------------------------------------------------
const char *ta_strerror(int code) {
  switch (code) {
    case TA_OK: return "ok";
    case TA_E_INVALID: return "invalid argument";
    case TA_E_ALPHABET: return "residue outside the alphabet";
    case TA_E_SPACE: return "output buffer too small";
    case TA_E_DEVICE: return "HIP runtime error";
    case TA_E_NOMEM: return "out of memory";
    case TA_E_BUDGET: return "request exceeds the traceback-memory budget";
    case TA_E_CANCELLED: return "request cancelled";
    case TA_E_TICKET: return "unknown ticket";
    case TA_E_NODEVICE: return "no usable GPU";
    case TA_E_WALK: return "a traceback walk did not complete";
    case TA_E_LIMIT: return "input exceeds the fixed scratch of the allocation-free entry point";
    case TA_E_PENDING: return "pending";
    default: return "unknown error";
  }
}

int ta_matrix(int32_t alphabet, int32_t match, int32_t mismatch, int8_t *out, int32_t *K) {
  if (alphabet != TA_PROTEIN && alphabet != TA_DNA) return TA_E_INVALID;
  Alphabet A = make_alphabet(alphabet, match, mismatch);
  if (K) *K = A.K;
  if (out) memcpy(out, A.S, (size_t)A.K * A.K);
  return TA_OK;
}

int ta_encode(int32_t alphabet, const char *s, int32_t len, uint8_t *out) {
  if (alphabet != TA_PROTEIN && alphabet != TA_DNA) return TA_E_INVALID;
  if (len < 0 || (len > 0 && (!s || !out))) return TA_E_INVALID;
  Alphabet A = make_alphabet(alphabet, 1, -1);
  for (int32_t i = 0; i < len; ++i) {
    int c = A.map[(unsigned char)s[i]];
    if (c < 0) return TA_E_ALPHABET;
    out[i] = (uint8_t)c;
  }
  return TA_OK;
}

int ta_create(ta_ctx **out, int device) {
  if (!out) return TA_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TA_E_NODEVICE;
  if (device < 0 || device >= ndev) return TA_E_INVALID;
  HIPCHK(hipSetDevice(device));
  ta_ctx *c = new ta_ctx();
  c->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->cus = cus;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->lds_max = lds;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return TA_E_DEVICE;
  }
  pair_scratch_setup(device);  // ta_align_pair_static's array on this device (reports its own failure)
  *out = c;
  return TA_OK;
}

void ta_destroy(ta_ctx *c) {
  if (!c) return;
  serve_shutdown(c);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->pair_stage) (void)hipHostFree(c->pair_stage);
  c->seq_arena.reset();  // freed here unless a batch still holds it (none after serve_shutdown)
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int params_check(const ta_params *p) {
  if (!p) return TA_E_INVALID;
  if (p->mode != TA_GLOBAL && p->mode != TA_LOCAL) return TA_E_INVALID;
  if (p->alphabet != TA_PROTEIN && p->alphabet != TA_DNA) return TA_E_INVALID;
  if (p->gap_open > 0 || p->gap_extend > 0 || p->gap_open < -1000 || p->gap_extend < -1000) return TA_E_INVALID;
  if (p->alphabet == TA_DNA && (p->match < -100 || p->match > 100 || p->mismatch < -100 || p->mismatch > 100))
    return TA_E_INVALID;
  if (p->lanes != TA_LANES_AUTO && p->lanes != TA_LANES_I16 && p->lanes != TA_LANES_I32 &&
      p->lanes != TA_LANES_LONG)
    return TA_E_INVALID;
  if (p->lanes == TA_LANES_LONG && p->mode != TA_GLOBAL) return TA_E_INVALID;
  return TA_OK;
}

// ---------------------------------------------------------------------------
// Batch construction.
namespace {

struct HostPair {
  uint64_t a_off, b_off;
  int32_t la, lb;
  int path;  // 3 long-pair pipeline, 2 packed int16, 1 int32, 0 no DP (empty side)
  bool maysat = false;  // path 2, local: the int16 fill may saturate (int32 rerun then)
};

// Global pairs of at least this many cells go through the long-pair
// pipeline: one task would otherwise hold a single wave for the whole sweep
// (and 4 bits per cell of direction planes).
constexpr int64_t kLongCells = 1ll << 25;

bool long_pair(const ta_params &p, int la, int lb) {
  if (p.mode != TA_GLOBAL || la == 0 || lb == 0) return false;
  if (p.lanes == TA_LANES_LONG) return true;
  return p.lanes == TA_LANES_AUTO && (int64_t)la * lb >= kLongCells;
}

// Residue codes one task's strips and columns read (expand_codes_kernel).
struct CodeJob {
  uint64_t row_off, col_off;
  int32_t nrows, ncol, lo, hi, off0, off1;
};

struct Plan {
  // [0] whole-wave groups, [1] G-lane groups (NGW per wave unit), [2] 8-lane
  // groups (8 per wave unit; packed couples only, tasks1[2] stays empty)
  std::vector<Task> tasks2[3], tasks1[3];
  std::vector<CodeJob> cjobs;
  uint64_t rows_n = 0, cols_n = 0;        // rowcodes / colcodes entries planned
  std::vector<PairTB> ptb;
  uint64_t ring_elems = 0, tb_bytes = 0;
};

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Resident one-wave fill workgroups per CU: one per SIMD.  A fill wave keeps
// its SIMD's VALU busy by itself (the recurrence's dependent chain per row is
// a fraction of a step's instructions), and a second wave on the SIMD only
// halves the speed of the long tasks that end the batch (measured on the
// 65,536-pair local batch: 9.1 ms at 4 waves per CU, 9.9 at 8, 10.5 at 12).
// Fill waves per CU: as wanted, within LDS; walking fills have WALK_FW (one
// workgroup per CU).
int fill_waves_per_cu(int P, int Kp, bool nib, int want, bool walk = false) {
  if (walk) return WALK_FW;
  return std::max(1, std::min(want, (160 * 1024) / fill_lds_bytes(P, Kp, nib)));
}

// Append one task (pair lo / hi; hi == lo for a single) to the plan, swept
// by groups of GG lanes.
// n_force / S_force (wave units): columns and strips of the task's unit; extra
// columns lead (local couples only), extra strips are pad rows.

// Raw letters to residue codes in place, 16 bytes per thread step, through a
// 256-entry map (0xff: outside the alphabet, which sets *bad).
__global__ void encode_kernel(uint4 *codes, size_t words, const uint8_t *map, int32_t *bad) {
  __shared__ uint8_t m[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) m[i] = map[i];
  __syncthreads();
  uint32_t any = 0;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (size_t)gridDim.x * blockDim.x) {
    uint4 v = codes[w];
    uint32_t x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t b0 = m[x[q] & 0xff], b1 = m[(x[q] >> 8) & 0xff], b2 = m[(x[q] >> 16) & 0xff], b3 = m[x[q] >> 24];
      any |= b0 | b1 | b2 | b3;
      x[q] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
    }
    codes[w] = make_uint4(x[0], x[1], x[2], x[3]);
  }
  if (__any(any & 0x80) && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}

// The row and column codes of every planned task, expanded on the GPU from
// the uploaded residues: row r of a half is its pair's query residue r (pad
// past m), column j its reference residue j - off (pad before the leading
// columns and past n).  One workgroup per task.
__global__ void expand_codes_kernel(const CodeJob *jobs, int njobs, const PairDev *pairs, const uint8_t *codes,
                                    int pad, uint16_t *rowcodes, uint16_t *colcodes) {
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const CodeJob c = jobs[q];
    const PairDev A = pairs[c.lo], B = pairs[c.hi];
    for (int r = threadIdx.x; r < c.nrows; r += blockDim.x) {
      const int ca = r < A.la ? codes[A.a_off + r] : pad;
      const int cb = r < B.la ? codes[B.a_off + r] : pad;
      rowcodes[c.row_off + r] = (uint16_t)(ca | (cb << 8));
    }
    for (int j = threadIdx.x; j < c.ncol; j += blockDim.x) {
      const int ja = j - c.off0, jb = j - c.off1;
      const int ca = (ja >= 0 && ja < A.lb) ? codes[A.b_off + ja] : pad;
      const int cb = (jb >= 0 && jb < B.lb) ? codes[B.b_off + jb] : pad;
      colcodes[c.col_off + j] = (uint16_t)(ca | (cb << 8));
    }
  }
}

// Allocates a plan's row and column codes on the device and expands them
// there (the residues and pair descriptors are already uploaded).
// Allocates a plan's row and column codes and expands them from a job list
// already on the device.
int expand_jobs(const CodeJob *d_jobs, size_t njobs, uint64_t rows_n, uint64_t cols_n, const PairDev *d_pairs,
                const uint8_t *d_codes, int Kp, uint16_t **d_rows, uint16_t **d_cols, hipStream_t st) {
  int rc = dalloc(d_rows, (size_t)rows_n);
  rc = rc ? rc : dalloc(d_cols, (size_t)cols_n);
  if (rc || !njobs) return rc;
  (void)hipGetLastError();
  hipLaunchKernelGGL(expand_codes_kernel, dim3((unsigned)std::min<size_t>(njobs, 65536)), dim3(256), 0, st,
                     d_jobs, (int)njobs, d_pairs, d_codes, Kp - 1, *d_rows, *d_cols);
  return hipGetLastError() == hipSuccess ? TA_OK : TA_E_DEVICE;
}

int expand_codes(const std::vector<CodeJob> &jobs, uint64_t rows_n, uint64_t cols_n, const PairDev *d_pairs,
                 const uint8_t *d_codes, int Kp, uint16_t **d_rows, uint16_t **d_cols, hipStream_t st) {
  int rc = dalloc(d_rows, (size_t)rows_n);
  rc = rc ? rc : dalloc(d_cols, (size_t)cols_n);
  if (rc || jobs.empty()) return rc;
  CodeJob *d_jobs = nullptr;
  rc = upload(&d_jobs, jobs, st);
  if (rc) return rc;
  (void)hipGetLastError();
  hipLaunchKernelGGL(expand_codes_kernel, dim3((unsigned)std::min<size_t>(jobs.size(), 65536)), dim3(256), 0, st,
                     d_jobs, (int)jobs.size(), d_pairs, d_codes, Kp - 1, *d_rows, *d_cols);
  if (hipGetLastError() != hipSuccess) rc = TA_E_DEVICE;
  // the job list goes back to the cache once the expansion has read it
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = TA_E_DEVICE;
  dfree(d_jobs);
  return rc;
}

// Where a task's row codes, column codes, ring and direction planes start in
// the plan's arrays, or (task_span) how many entries / bytes it takes there.
struct TaskAt {
  uint64_t row, col, ring, tb;
};
// Row codes: S * ST entries; column codes: 8 NB + 8 (a multiple of 8, so
// every task's columns start 16-byte aligned); ring: two arrays of
// 8 NB + 2 GG + 16 when there is more than one strip.
TaskAt task_span(int P, int GG, bool trace, int n, int S) {
  const int NB = ceil_div(n + GG, 8);
  return TaskAt{(uint64_t)S * GG * R, (uint64_t)(8 * NB + 8), S > 1 ? 2ull * (8 * NB + 2 * GG + 16) : 0,
                trace ? (uint64_t)S * tb_strip_bytes(P, NB, GG) : 0};
}

// One task (pair lo / hi; hi == lo for a single) of n columns and S strips
// at the given offsets: the task, its code-expansion job (expanded on the
// GPU) and the pairs' direction-plane records.
Task make_task(const TaskAt &at, CodeJob &job, std::vector<PairTB> &ptb, int P, int GG, bool trace,
               const std::vector<HostPair> &hp, int lo, int hi, bool single, int n, int S) {
  const HostPair &A = hp[lo], &B = hp[hi];
  Task T{};
  T.pair[0] = lo; T.pair[1] = hi;
  T.m[0] = A.la; T.m[1] = B.la;
  T.n = n;
  T.off[0] = T.n - A.lb; T.off[1] = T.n - B.lb;
  T.S = S;
  T.NB = ceil_div(T.n + GG, 8);
  T.single = single ? 1 : 0;
  T.row_off = at.row;
  T.col_off = at.col;
  job = CodeJob{T.row_off, T.col_off, T.S * GG * R, 8 * T.NB + 8, lo, hi, T.off[0], T.off[1]};
  T.ring_len = T.S > 1 ? 8 * T.NB + 2 * GG + 16 : 0;
  T.ring_off = at.ring;
  T.tb_off = trace ? at.tb : 0;
  for (int h = 0; h < (single ? 1 : 2); ++h) {
    PairTB &d = ptb[h ? hi : lo];
    d.tb_off = T.tb_off; d.P = P; d.half = h; d.NB = T.NB; d.off = T.off[h]; d.GG = GG;
  }
  return T;
}

// make_task at the plan's running offsets, appended to dst.
void add_task(Plan &pl, std::vector<Task> &dst, int P, int GG, bool trace, const std::vector<HostPair> &hp,
              int lo, int hi, bool single, int Kp, int n_force = 0, int S_force = 0) {
  const HostPair &A = hp[lo], &B = hp[hi];
  const int n = n_force ? n_force : std::max(A.lb, B.lb);
  const int S = S_force ? S_force : ceil_div(std::max(A.la, B.la), GG * R);
  const TaskAt at{pl.rows_n, pl.cols_n, pl.ring_elems, pl.tb_bytes}, sz = task_span(P, GG, trace, n, S);
  pl.cjobs.emplace_back();
  dst.push_back(make_task(at, pl.cjobs.back(), pl.ptb, P, GG, trace, hp, lo, hi, single, n, S));
  pl.rows_n += sz.row; pl.cols_n += sz.col; pl.ring_elems += sz.ring; pl.tb_bytes += sz.tb;
}

// Positions 0..n-1 in ascending order of key, equal keys in position order:
// an LSD radix sort over 11-bit digits, as many as the largest key needs.
// The planner orders tens of thousands of pairs and units per batch, where
// comparison sorts that read their keys through the pair table cost
// milliseconds.
std::vector<int32_t> order_by_key(std::vector<uint64_t> key) {
  const size_t n = key.size();
  std::vector<int32_t> idx(n), ti(n);
  std::vector<uint64_t> tk(n);
  for (size_t t = 0; t < n; ++t) idx[t] = (int32_t)t;
  uint64_t any = 0;
  for (uint64_t k : key) any |= k;
  for (int sh = 0; sh < 64 && (any >> sh) != 0; sh += 11) {
    std::vector<size_t> at(2049, 0);
    for (uint64_t k : key) ++at[((k >> sh) & 2047) + 1];
    for (int d = 0; d < 2048; ++d) at[d + 1] += at[d];
    for (size_t t = 0; t < n; ++t) {
      const size_t q = at[(key[t] >> sh) & 2047]++;
      ti[q] = idx[t];
      tk[q] = key[t];
    }
    idx.swap(ti);
    key.swap(tk);
  }
  return idx;
}

// `v` reordered by ascending key[v's position].
template <typename T>
void gather_by_key(std::vector<T> &v, std::vector<uint64_t> key) {
  const std::vector<int32_t> o = order_by_key(std::move(key));
  std::vector<T> w;
  w.reserve(v.size());
  for (int32_t t : o) w.push_back(std::move(v[t]));
  v.swap(w);
}

// Lengths are at most 1,000,000 (< 2^20), strips of STRIP rows at most 2^13.
constexpr uint64_t LEN_TOP = (1u << 20) - 1, STRIPS_TOP = (1u << 13) - 1;
static_assert((1000000 + STRIP - 1) / STRIP <= (int)STRIPS_TOP, "strip count exceeds its key field");

// A couple (or single) waiting to be grouped into a wave unit of G-lane tasks.
struct Unit {
  int lo, hi;
  bool single;
  int m, n;  // rows and columns (maxima over the two halves)
};

// Wave units: GWAVE / GG tasks of GG lanes and one shape that one wave runs
// in lockstep.  Local couples are ordered by (strips, columns) and take their
// unit's column count as leading pad columns; global couples need equal
// column counts (H(m, n) is read at the last column), so they are grouped by
// columns.  A unit short of its tasks is filled with copies of its first
// task that write nothing.  Units go out largest first, flattened (GWAVE /
// GG tasks per unit).
void add_units(Plan &pl, std::vector<Task> &dst, int P, int GG, std::vector<Unit> u, bool local, bool trace,
               const std::vector<HostPair> &hp, int Kp) {
  const int NG = GWAVE / GG;
  auto strips = [GG](int m) { return ceil_div(std::max(m, 1), GG * R); };
  {  // local: more strips first, then more columns; global: columns first
    std::vector<uint64_t> key(u.size());
    for (size_t t = 0; t < u.size(); ++t) {
      const uint64_t sk = STRIPS_TOP - (uint64_t)strips(u[t].m), nk = LEN_TOP - (uint64_t)u[t].n;
      key[t] = local ? (sk << 20 | nk) : (nk << 13 | sk);
    }
    gather_by_key(u, std::move(key));
  }
  // The units' extents and shapes first, then their tasks straight into dst
  // in order of cost (more 8-step blocks first, S * NB < 2^40).
  struct Span { uint32_t k, e; int nq, mq; };
  std::vector<Span> spans;
  spans.reserve(u.size() / NG + 1);
  std::vector<uint64_t> key;
  key.reserve(u.size() / NG + 1);
  for (size_t k = 0; k < u.size();) {
    size_t e = k;
    while (e < u.size() && e - k < (size_t)NG && (local || u[e].n == u[k].n)) ++e;
    int nq = 0, mq = 0;
    for (size_t i = k; i < e; ++i) { nq = std::max(nq, u[i].n); mq = std::max(mq, u[i].m); }
    spans.push_back(Span{(uint32_t)k, (uint32_t)e, nq, mq});
    key.push_back((1ull << 40) - (uint64_t)strips(mq) * (uint64_t)ceil_div(nq + GG, 8));
    k = e;
  }
  const std::vector<int32_t> o = order_by_key(std::move(key));
  // Offsets in unit order (every task of a unit has the unit's shape), then
  // the tasks, jobs and plane records on the host threads: each unit's
  // entries, and its pairs' records, are its own.
  struct UnitAt { TaskAt at, sz; size_t dst, job; };
  std::vector<UnitAt> ua(o.size());
  size_t nd = dst.size(), nj = pl.cjobs.size();
  for (size_t x = 0; x < o.size(); ++x) {
    const Span &sp = spans[o[x]];
    const uint64_t cnt = sp.e - sp.k;
    UnitAt &a = ua[x];
    a.at = TaskAt{pl.rows_n, pl.cols_n, pl.ring_elems, pl.tb_bytes};
    a.sz = task_span(P, GG, trace, sp.nq, strips(sp.mq));
    a.dst = nd;
    a.job = nj;
    pl.rows_n += cnt * a.sz.row; pl.cols_n += cnt * a.sz.col;
    pl.ring_elems += cnt * a.sz.ring; pl.tb_bytes += cnt * a.sz.tb;
    nd += NG;
    nj += cnt;
  }
  dst.resize(nd);
  pl.cjobs.resize(nj);
  parallel_for(o.size(), [&](size_t x) {
    const Span &sp = spans[o[x]];
    const UnitAt &a = ua[x];
    for (uint32_t i = sp.k; i < sp.e; ++i) {
      const uint64_t q = i - sp.k;
      const TaskAt at{a.at.row + q * a.sz.row, a.at.col + q * a.sz.col, a.at.ring + q * a.sz.ring,
                      a.at.tb + q * a.sz.tb};
      dst[a.dst + q] = make_task(at, pl.cjobs[a.job + q], pl.ptb, P, GG, trace, hp, u[i].lo, u[i].hi, u[i].single,
                                 sp.nq, strips(sp.mq));
    }
    for (uint32_t q = sp.e - sp.k; q < (uint32_t)NG; ++q) {  // fillers: copies that write nothing
      dst[a.dst + q] = dst[a.dst];
      dst[a.dst + q].single = 2;
    }
  });
}

// Steps a G-lane group needs for a task of these shapes, in 8-step blocks.
double group_cost(int mmax, int n) { return (double)ceil_div(mmax, G * R) * ceil_div(n + G, 8); }

// The same for an 8-lane group (64-row strips, a 7-column skew).
double cost8(int mmax, int n) { return (double)ceil_div(mmax, 8 * R) * ceil_div(n + 8, 8); }

// Packed couples whose 8-lane sweep takes at most this fraction of the mean
// work of a group slot go into 8-lane units (8 tasks per wave; global ones of
// equal column counts): 64-row strips and a 7-column skew compute fewer pad
// cells than 128-row strips and a 15-column skew (65,536-pair log-normal
// local batch: 1.11x the real cells instead of 1.164x; the 130,816-pair
// all-vs-all fill 4.07 -> 3.85 ms), while the 16-lane units keep the tasks
// whose 8-lane units would be long enough to end the run late.  A
// list-scheduling model of the local batch gives the shortest run between
// 0.4x and 0.9x; 8-lane units of every task end it ~50% late.
double eight_lane_limit() {
  const char *e = getenv("TWEAKALIGN_EIGHT_LANE");  // fraction; "0" turns 8-lane units off
  return e ? atof(e) : 0.5;
}


// Whole-wave groups for tasks long enough to set the makespan of a group
// sweep: more than 0.75x the mean work of a group slot (a list-scheduling
// simulation of the 65,536-pair log-normal batch gives 99% slot efficiency
// there, 68% at 1.5x), and tall enough (>= 384 rows) not to waste most of a
// 512-row strip.  With fewer tasks than resident waves (`few`) the run takes
// as long as one task's sweep, so every task taller than one group strip
// gets a whole wave (a lone 300 x 320 pair: 384 steps instead of 1,008 in
// 16-lane strips).
int pick_gg(int mmax, int n, double mean_slot, bool few) {
  if (mmax <= G * R) return G;
  if (few) return GWAVE;
  return (mmax >= 3 * GWAVE * R / 4 && group_cost(mmax, n) > 0.75 * mean_slot) ? GWAVE : G;
}

// Start of a run: work-queue counters to 0, fill outputs to their initial
// values (empty pairs keep theirs), long-pair progress counters to 0.  A
// kernel rather than memset / memcpy nodes, so a captured run is all kernels.
__global__ void init_kernel(int32_t *ctr, int nctr, FillOut *fo, const FillOut *fo_init, int n, int32_t *prog,
                            int nprog) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (int k = tid; k < nctr; k += stride) ctr[k] = 0;
  for (int k = tid; k < n; k += stride) fo[k] = fo_init[k];
  for (int k = tid; k < nprog; k += stride) prog[k] = 0;
}

}  // namespace

struct ta_batch {
  hipStream_t stream = nullptr;  // the context stream, or a serving stream
  ta_params p{};
  int32_t n = 0;
  Alphabet alph;
  int Kp = 0;
  bool trace = false;       // direction planes are stored (local, or CIGAR asked)
  bool may_saturate = false;
  int64_t cells = 0;
  std::vector<HostPair> hp;
  std::vector<FillOut> fo_init;
  FillOut *d_fo_init = nullptr;
  int32_t *d_order = nullptr;  // traceback walker -> pair, longest first
  // Walks inside the fill kernel (direction planes, no int32 rerun possible):
  // the traceback kernel then takes only the pairs no fill task holds.
  bool walk = false;
  // Walking batches of few tasks: the first ncoop[P-1]... whole-wave tasks of
  // each lane width are swept by a workgroup's fill waves together.
  int32_t ncoop2 = 0, ncoop1 = 0;
  int32_t n_rest = 0;
  int32_t *d_order_rest = nullptr;
  std::vector<PairDev> pairs_h;
  uint64_t cig_bytes = 0;
  Plan plan;
  // device
  uint8_t *d_codes = nullptr;
  std::shared_ptr<uint8_t> codes_owner;  // set when d_codes is a resident-sequence arena (held, not freed)
  int8_t *d_table = nullptr;
  int8_t *d_table_nib = nullptr;  // packed couples: 4-bit biased scores
  bool nib = false;
  int fill_waves = 0;             // resident fill waves per CU (0: default_waves)
  bool wave_tasks = false;        // whole-wave groups for every multi-strip task
  Task *d_tasks2[3] = {nullptr, nullptr, nullptr}, *d_tasks1[3] = {nullptr, nullptr, nullptr};
  uint16_t *d_rowcodes = nullptr, *d_colcodes = nullptr;
  uint32_t *d_ring = nullptr;
  uint8_t *d_tb = nullptr;
  PairTB *d_ptb = nullptr;
  PairDev *d_pairs = nullptr;
  FillOut *d_fo = nullptr;
  ResultDev *d_res = nullptr;
  char *d_cig = nullptr;
  int32_t *d_ctr = nullptr;
  // int32 rerun of saturated pairs (allocated on demand)
  Task *d_tasks_r = nullptr;
  uint16_t *d_rowcodes_r = nullptr, *d_colcodes_r = nullptr;
  uint32_t *d_ring_r = nullptr;
  uint8_t *d_tb_r = nullptr;
  int32_t n_rerun = 0;
  // long-pair pipeline
  std::vector<LongPair> longs;
  std::vector<LongItem> litems;
  std::vector<int32_t> litem_first;  // per long pair: its first queue item
  LongPair *d_lp = nullptr;
  LongItem *d_litems = nullptr;
  int32_t *d_lring = nullptr, *d_lck = nullptr, *d_lprog = nullptr, *d_wring = nullptr;
  int2 *d_lcol = nullptr;
  uint8_t *d_wtb = nullptr;
  uint64_t lring_elems = 0, lck_elems = 0, lcol_elems = 0, wring_stride = 0, wtb_stride = 0;
  int32_t lprog_n = 0, long_wgs = 0, spin_limit = 1 << 20;  // about a second of polling
  // Most steps a traceback walk may take (TWEAKALIGN_WALK_GUARD lowers it so
  // that walks end flagged; otherwise above every walk's length).
  int32_t guard_cap = INT32_MAX;
  // captured run
  bool graph_on = false, phases_timed = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // timing
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double ms[4] = {0, 0, 0, 0};
  // plain (one launch per pair) resources, built on first use
  bool plain_ready = false;
  std::vector<Task> plain_tasks;
  Task *d_plain_tasks = nullptr;
  // int32 reruns of the pairs whose int16 fill may saturate: NGW entries
  // each (index plain_rerun[k]), their planes' PairTB in d_pptb_r
  std::vector<int32_t> plain_rerun;
  Task *d_plain_rtasks = nullptr;
  PairTB *d_pptb_r = nullptr;
  uint16_t *d_prow = nullptr, *d_pcol = nullptr;
  uint32_t *d_pring = nullptr;
  uint8_t *d_ptbuf = nullptr;
  PairTB *d_pptb = nullptr;
};

namespace {

bool global_i16_safe(const ta_params &p, int maxS, int la, int lb) {
  const long lo = 3L * -p.gap_open + (long)(la + lb + 2) * -p.gap_extend + 128;
  const long hi = (long)std::min(la, lb) * maxS + maxS;
  return lo < 16000 && hi < 32000;
}
}  // namespace

bool global_int16_safe(const ta_params *p, int maxS, int la, int lb) { return global_i16_safe(*p, maxS, la, lb); }

namespace {

// Resident fill waves per CU a batch asks for: its tuning, else one per SIMD
// when direction planes are written (those kernels keep a SIMD busy alone and
// more waves slowed the 65,536-pair batch: 8.4 ms at 4 per CU, 9.1 at 8),
// else three per SIMD (the score-only kernels are leaner and latency-exposed:
// packed int16 ops issue at 77% of peak with one wave, 90% with two; the
// all-vs-all batch runs 5.6 ms at 4 per CU, 4.3 at 8, 4.0 at 12).
int waves_wanted(const ta_batch *B) { return B->fill_waves ? B->fill_waves : (B->trace ? 4 : 12); }

void free_rerun(ta_batch *B) {
  dfree(B->d_tasks_r); dfree(B->d_rowcodes_r); dfree(B->d_colcodes_r);
  dfree(B->d_ring_r); dfree(B->d_tb_r);
  B->n_rerun = 0;
}

}  // namespace

void ta_batch_destroy(ta_batch *B) {
  if (!B) return;
  // The blocks go back to the cache, where a batch on another stream may take
  // them at once: the batch's last run (every run form ends by recording
  // ev[3]) must have finished first.  The event is the batch's own, so this
  // holds even where the run's stream is gone.
  if (B->ev[3]) (void)hipEventSynchronize(B->ev[3]);
  if (B->codes_owner) B->d_codes = nullptr;
  dfree(B->d_codes); dfree(B->d_table); dfree(B->d_table_nib);
  for (int g = 0; g < 3; ++g) { dfree(B->d_tasks2[g]); dfree(B->d_tasks1[g]); }
  dfree(B->d_rowcodes); dfree(B->d_colcodes); dfree(B->d_ring); dfree(B->d_tb);
  dfree(B->d_ptb); dfree(B->d_pairs); dfree(B->d_fo); dfree(B->d_res); dfree(B->d_cig);
  dfree(B->d_ctr); dfree(B->d_fo_init); dfree(B->d_order); dfree(B->d_order_rest);
  free_rerun(B);
  dfree(B->d_lp); dfree(B->d_litems); dfree(B->d_lring); dfree(B->d_lck); dfree(B->d_lcol); dfree(B->d_lprog);
  dfree(B->d_wring); dfree(B->d_wtb);
  dfree(B->d_plain_tasks); dfree(B->d_prow); dfree(B->d_pcol); dfree(B->d_pring);
  dfree(B->d_ptbuf); dfree(B->d_pptb); dfree(B->d_plain_rtasks); dfree(B->d_pptb_r);
  for (auto &e : B->ev)
    if (e) (void)hipEventDestroy(e);
  if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  vec_give(B->hp);
  vec_give(B->fo_init);
  vec_give(B->pairs_h);
  vec_give(B->plan.ptb);
  vec_give(B->plan.cjobs);
  for (int g = 0; g < 3; ++g) {
    vec_give(B->plan.tasks2[g]);
    vec_give(B->plan.tasks1[g]);
  }
  delete B;
}

// A pair the batch runtime takes: lengths in range, letters present, and
// (global) scores clear of the int32 kernels' -infinity (-2^29).
static bool pair_ok(const ta_params &p, const ta_pair &q) {
  if (q.la < 0 || q.lb < 0 || q.la > 1000000 || q.lb > 1000000 || (q.la && !q.a) || (q.lb && !q.b)) return false;
  return p.mode == TA_LOCAL ||
         (int64_t)(q.la + q.lb + 2) * -p.gap_extend + 2 * (int64_t)-p.gap_open < (1ll << 28);
}

int ta_batch_create(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n, ta_batch **out) {
  return batch_create_on(ctx, p, pairs, n, ctx ? ctx->stream : nullptr, out);
}

int batch_create_on(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n, hipStream_t stream,
                    ta_batch **out, const ta_tuning *force, const ResidentBatch *res) {
  if (!ctx || !p || !out || n < 0 || (n > 0 && !pairs && !res)) return TA_E_INVALID;
  *out = nullptr;
  if (params_check(p)) return TA_E_INVALID;
  if (res && ((int64_t)res->la.size() != n || (n > 0 && !res->arena))) return TA_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  PhaseTrace tr("create");

  ta_batch *B = new ta_batch();
  B->stream = stream;
  B->p = *p;
  B->n = n;
  B->alph = make_alphabet(p->alphabet, p->match, p->mismatch);
  B->Kp = B->alph.K + 1;
  const bool local = p->mode == TA_LOCAL;
  B->trace = local || p->traceback;
  B->nib = nibble_ok(B->alph, local);
  // Launch shape: the autotuner's candidate, else the context's tuning for
  // this kind of batch, else the defaults.
  ta_tuning tu{0, TA_LANES_AUTO, 0};
  if (force) {
    tu = *force;
  } else {
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    int bucket = 0;
    if (res) {  // the bucket from the resident sequences' lengths
      std::vector<ta_pair> lens((size_t)n);
      for (int32_t k = 0; k < n; ++k) lens[k] = ta_pair{nullptr, res->la[k], nullptr, res->lb[k]};
      bucket = ta_tune_bucket(lens.data(), n);
    } else {
      bucket = ta_tune_bucket(pairs, n);
    }
    auto it = ctx->tune.find(tune_key(p, bucket));
    if (it != ctx->tune.end()) tu = it->second;
  }
  B->fill_waves = tu.waves_per_cu;
  B->wave_tasks = tu.wave_tasks != 0;
  const int lanes = p->lanes != TA_LANES_AUTO ? p->lanes : tu.lanes;

  // Encode (sequences shared between pairs are encoded once).  Each
  // sequence also gets the sum over its residues of the best score the
  // residue can reach against any letter (0 if none is positive): no local
  // alignment of it scores more, which decides whether int16 can saturate.
  int best[32] = {0};
  for (int r = 0; r < B->alph.K; ++r)
    for (int c = 0; c < B->alph.K; ++c) best[r] = std::max<int>(best[r], B->alph.S[r * B->alph.K + c]);
  // Unique sequences (by pointer and length) and their offsets in the code
  // array.  Pointers are split by hash over the host threads; each thread
  // scans every pair and keeps a table for its share only (a few hundred KB,
  // cache-resident), mapping a pointer to its latest unique entry, so the
  // same pointer always meets the same thread, in pair order.
  struct Uniq { const char *s; int32_t len; uint64_t off; int64_t bound; };
  // A pair as the checks and the planner see it (resident sequences: the
  // lengths; their letters are on the device).
  auto pair_at = [&](int32_t k) -> ta_pair {
    if (!res) return pairs[k];
    const int32_t la = res->la[k], lb = res->lb[k];
    return ta_pair{la ? "" : nullptr, la, lb ? "" : nullptr, lb};
  };
  for (int32_t k = 0; k < n; ++k)
    if (!pair_ok(*p, pair_at(k))) {
      delete B;
      return TA_E_INVALID;
    }
  tr.mark("validate");
  auto mix = [](const char *s) -> uint64_t {
    uint64_t x = (uint64_t)(uintptr_t)s;
    x = (x ^ (x >> 29)) * 0xbf58476d1ce4e5b9ull;
    return x ^ (x >> 32);
  };
  const int T = n >= 8192 ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
  // a pointer's thread: the hash's top 32 bits scaled to [0, T) (no division)
  auto part_of = [T](uint64_t x) { return (int)(((x >> 32) * (uint64_t)T) >> 32); };
  // The calling thread's buffers, kept between calls: their pages stay
  // mapped (first-touch page faults cost more than the work here).
  // (Named through references: the worker threads below must reach the
  // calling thread's copies, not their own.)
  static thread_local std::vector<int32_t> tl_ua, tl_ub;
  static thread_local std::vector<std::vector<Uniq>> tl_part;
  static thread_local std::vector<Uniq> tl_uq;
  std::vector<int32_t> &ua = tl_ua, &ub = tl_ub;
  std::vector<std::vector<Uniq>> &part = tl_part;
  std::vector<Uniq> &uq = tl_uq;
  size_t ctotal = 0;  // bytes of the unique sequences' codes
  if (!res) {
    ua.resize((size_t)n);
    ub.resize((size_t)n);
    if ((int)part.size() < T) part.resize(T);
    for (int t = 0; t < T; ++t) part[t].clear();
    parallel_tasks(T, [&](int t) {
      struct Slot { const char *s; int32_t idx; };
      // sized for this thread's expected share at most a quarter full; the
      // thread keeps it between calls (the host pool's threads live on)
      size_t cap = 64;
      while (cap < 8 * (size_t)n / (size_t)T) cap <<= 1;
      static thread_local std::vector<Slot> tl_tab;
      std::vector<Slot> &tab = tl_tab;
      if (tab.size() != cap) tab.assign(cap, Slot{nullptr, -1});
      else std::fill(tab.begin(), tab.end(), Slot{nullptr, -1});
      std::vector<Uniq> &mine = part[t];
      auto grow = [&]() {  // keep the table at most half full
        std::vector<Slot> old(2 * cap, Slot{nullptr, -1});
        old.swap(tab);
        cap *= 2;
        for (const Slot &e : old) {
          if (e.idx < 0) continue;
          size_t h = (size_t)mix(e.s) & (cap - 1);
          while (tab[h].idx >= 0) h = (h + 1) & (cap - 1);
          tab[h] = e;
        }
      };
      auto uniq = [&](const char *s, int32_t len, uint64_t x) -> int32_t {
        size_t h = (size_t)x & (cap - 1);
        while (tab[h].idx >= 0 && tab[h].s != s) h = (h + 1) & (cap - 1);
        if (tab[h].idx >= 0 && mine[tab[h].idx].len == len) return tab[h].idx;
        const bool fresh = tab[h].idx < 0;
        mine.push_back(Uniq{s, len, 0, -1});
        tab[h] = Slot{s, (int32_t)mine.size() - 1};
        if (fresh && 2 * mine.size() > cap) grow();
        return (int32_t)mine.size() - 1;
      };
      for (int32_t k = 0; k < n; ++k) {
        const ta_pair &q = pairs[k];
        const uint64_t xa = mix(q.a), xb = mix(q.b);
        if (part_of(xa) == t) ua[k] = uniq(q.a, q.la, xa);
        if (part_of(xb) == t) ub[k] = uniq(q.b, q.lb, xb);
      }
    });
    tr.mark("dedupe hash");
    // Global order: partition by partition; each sequence starts 16-byte
    // aligned (walker windows).  Offsets and the flat array, per partition on
    // the threads, from the partitions' counts and byte totals.
    std::vector<int32_t> base(T + 1, 0);
    std::vector<size_t> bbase(T + 1, 0);
    for (int t = 0; t < T; ++t) {
      base[t + 1] = base[t] + (int32_t)part[t].size();
      size_t bytes = 0;
      for (const Uniq &e : part[t]) bytes += ((size_t)e.len + 15) & ~(size_t)15;
      bbase[t + 1] = bbase[t] + bytes;
    }
    uq.resize((size_t)base[T]);
    parallel_tasks(T, [&](int t) {
      size_t at = bbase[t];
      Uniq *dst = uq.data() + base[t];
      for (const Uniq &e : part[t]) {
        *dst = e;
        dst->off = at;
        at += ((size_t)e.len + 15) & ~(size_t)15;
        ++dst;
      }
    });
    ctotal = bbase[T];
    tr.mark("dedupe offsets");
    if (T > 1)
      parallel_for((size_t)n, [&](size_t k) {
        ua[k] += base[part_of(mix(pairs[k].a))];
        ub[k] += base[part_of(mix(pairs[k].b))];
      });
    tr.mark("dedupe");
  }
  // The walker reads 32-byte windows: 32 bytes of slack after the last
  // sequence.  The raw letters go up as they are, each sequence followed by
  // a letter that encodes to 0 up to the next 16-byte start (each host thread
  // copies its sequences, so the pages are first touched in parallel), and
  // the GPU encodes them in place, flagging letters outside the alphabet.
  const size_t cend = ((ctotal + 15) & ~(size_t)15) + 32;
  StageHold codes;  // raw letters of the unique sequences, 16-byte aligned starts
  if (!res) {
    codes.s = stage_take(cend);
    if (!codes.s.p) {
      delete B;
      return TA_E_NOMEM;
    }
  }
  uint8_t map[256], pad_letter = 0;
  for (int b = 255; b >= 0; --b) {
    const int c = B->alph.map[b];
    map[b] = c < 0 ? 0xff : (uint8_t)c;
    if (c == 0) pad_letter = (uint8_t)b;
  }
  if (!res) {
    parallel_for(uq.size(), [&](size_t u) {
      const Uniq &e = uq[u];
      uint8_t *dst = codes.s.p + e.off;
      if (e.len) memcpy(dst, e.s, (size_t)e.len);
      const size_t next = u + 1 < uq.size() ? uq[u + 1].off : cend;
      memset(dst + e.len, pad_letter, next - (e.off + (size_t)e.len));
    });
    if (uq.empty()) memset(codes.s.p, pad_letter, cend);
  }
  tr.mark("letters");
  // The sum over a sequence's residues of the best score each can reach
  // (0 if none is positive) bounds any local alignment of it; it decides
  // whether int16 can saturate, which only a sequence longer than
  // limit / (best of all letters) can reach, so only such sequences are summed.
  int maxbest = 0;
  for (int r = 0; r < B->alph.K; ++r) maxbest = std::max(maxbest, best[r]);
  const int64_t sat_limit = 32767 - 2L * B->alph.maxS;
  auto bound = [&](int32_t u) -> int64_t {
    Uniq &e = uq[u];
    if ((int64_t)e.len * maxbest <= sat_limit) return (int64_t)e.len * maxbest;
    if (e.bound < 0) {
      int64_t sum = 0;
      const unsigned char *src = (const unsigned char *)e.s;
      for (int32_t i = 0; i < e.len; ++i) sum += map[src[i]] < 0x80 ? best[map[src[i]]] : 0;
      e.bound = sum;
    }
    return e.bound;
  };
  // The letters go up while the host plans: from the pinned block the copy
  // runs on the DMA engine alone; the encoding kernel follows on the stream.
  hipStream_t st = B->stream;
  int rc = TA_OK;
  uint8_t *d_map = nullptr;
  int32_t *d_bad = nullptr;
  int32_t bad_letter = 0;
  if (res) {  // resident sequences: the codes are on the device already
    B->codes_owner = res->arena;
    B->d_codes = res->arena.get();
  } else {
    rc = dalloc(&B->d_codes, cend);
  }
  if (!rc && !res) {
    codes.stream = st;
    codes.queued = true;
    if (hipMemcpyAsync(B->d_codes, codes.s.p, cend, hipMemcpyHostToDevice, st) != hipSuccess) rc = TA_E_DEVICE;
  }
  rc = rc || res ? rc : upload(&d_map, std::vector<uint8_t>(map, map + 256), st);
  rc = rc || res ? rc : dalloc(&d_bad, 1);
  if (!rc && !res && hipMemsetAsync(d_bad, 0, sizeof(int32_t), st) != hipSuccess) rc = TA_E_DEVICE;
  if (!rc && !res) {
    (void)hipGetLastError();
    const size_t words = cend / 16;
    hipLaunchKernelGGL(encode_kernel, dim3((unsigned)std::min<size_t>((words + 255) / 256, 4096)), dim3(256), 0, st,
                       (uint4 *)B->d_codes, words, d_map, d_bad);
    if (hipGetLastError() != hipSuccess) rc = TA_E_DEVICE;
  }
  if (rc) {
    (void)hipStreamSynchronize(st);
    dfree(d_map);
    dfree(d_bad);
    ta_batch_destroy(B);
    return rc;
  }
  tr.mark("upload-encode");
  vec_take(B->hp);
  vec_take(B->fo_init);
  vec_take(B->pairs_h);
  B->hp.resize(n);
  B->fo_init.assign(n, FillOut{0, 0, 0, 0});
  B->pairs_h.resize(n);
  for (int32_t k = 0; k < n; ++k) {
    const ta_pair q = pair_at(k);
    HostPair &h = B->hp[k];
    h.a_off = res ? res->a_off[k] : uq[ua[k]].off;
    h.b_off = res ? res->b_off[k] : uq[ub[k]].off;
    h.la = q.la; h.lb = q.lb;
    B->cells += (int64_t)q.la * q.lb;
    if (q.la == 0 || q.lb == 0) {
      h.path = 0;
      FillOut &f = B->fo_init[k];
      if (local) { f.flags = 8; }
      else {
        const int len = q.la + q.lb;
        f.score = len ? p->gap_open + (len - 1) * p->gap_extend : 0;
        f.end_i = q.la; f.end_j = q.lb;
      }
    } else if (long_pair(*p, q.la, q.lb)) {
      h.path = 3;
    } else if (lanes == TA_LANES_I32 || (!B->nib && !B->trace)) {
      h.path = 1;
    } else if (local && q.lb > 32767) {
      h.path = 1;  // packed end-cell keys hold the column in 15 bits
    } else if (local) {
      h.path = 2;
      h.maysat = (res ? std::min(res->a_bound[k], res->b_bound[k]) : std::min(bound(ua[k]), bound(ub[k]))) > sat_limit;
      if (h.maysat) B->may_saturate = true;
    } else {
      h.path = global_i16_safe(*p, B->alph.maxS, q.la, q.lb) ? 2 : 1;
    }
    PairDev &d = B->pairs_h[k];
    d.a_off = h.a_off; d.b_off = h.b_off; d.la = q.la; d.lb = q.lb;
    d.cig_cap = p->traceback ? 4 * (q.la + q.lb + 1) : 0;  // one dword per run at most
    d.cig_off = B->cig_bytes;
    B->cig_bytes += (uint64_t)d.cig_cap;
  }

  tr.mark("pairs");
  // Couples: local pairs pad to the longer column count; global couples
  // need equal column counts (so H(m, n) of both sits in the last column).
  Plan &pl = B->plan;
  vec_take(pl.ptb);
  vec_take(pl.cjobs);
  for (int g = 0; g < 3; ++g) {
    vec_take(pl.tasks2[g]);
    vec_take(pl.tasks1[g]);
  }
  pl.ptb.assign(n, PairTB{0, 0, 0, 0, 0, 0, 0});
  pl.cjobs.reserve((size_t)n / 2 + NGW);
  std::vector<int> i16, i32;
  for (int32_t k = 0; k < n; ++k) {
    if (B->hp[k].path == 2) i16.push_back(k);
    else if (B->hp[k].path == 1) i32.push_back(k);
  }
  // More strips first, then more columns, then lower pair index (the lists
  // hold ascending pair indices).
  // Local pairs are ranked by 64-row strips (those of 8-lane units), which
  // keeps them in order of 128-row strips too.
  const int rank_rows = local ? 8 * R : STRIP;
  static_assert((1000000 + 8 * R - 1) / (8 * R) < (1 << 15), "strip count exceeds its key field");
  auto sort_bigger = [&](std::vector<int> &v) {
    std::vector<uint64_t> key(v.size());
    for (size_t t = 0; t < v.size(); ++t) {
      const HostPair &h = B->hp[v[t]];
      key[t] = ((1ull << 15) - (uint64_t)ceil_div(h.la, rank_rows)) << 20 | (LEN_TOP - (uint64_t)h.lb);
    }
    gather_by_key(v, std::move(key));
  };
  std::vector<std::pair<int, int>> couples;
  if (local) {
    sort_bigger(i16);
    for (size_t t = 0; t < i16.size(); t += 2)
      couples.push_back({i16[t], t + 1 < i16.size() ? i16[t + 1] : -1});
  } else {
    // Groups of equal columns in ascending order, each ranked as above; the
    // members of a group pair off in turn, the odd one out alone.  One sort
    // over all pairs (the key leads with the column count).
    std::vector<uint64_t> key(i16.size());
    for (size_t t = 0; t < i16.size(); ++t) {
      const HostPair &h = B->hp[i16[t]];
      key[t] = (uint64_t)h.lb << 35 | ((1ull << 15) - (uint64_t)ceil_div(h.la, rank_rows)) << 20;
    }
    gather_by_key(i16, std::move(key));
    for (size_t t = 0; t < i16.size();) {
      const bool two = t + 1 < i16.size() && B->hp[i16[t + 1]].lb == B->hp[i16[t]].lb;
      couples.push_back({i16[t], two ? i16[t + 1] : -1});
      t += two ? 2 : 1;
    }
  }
  tr.mark("couples");
  // Mean group-slot work decides which tasks get whole-wave groups.
  double work = 0;
  auto shape = [&](int lo, int hi) {
    return std::make_pair(std::max(B->hp[lo].la, B->hp[hi].la), std::max(B->hp[lo].lb, B->hp[hi].lb));
  };
  for (auto &c : couples) {
    auto sh = shape(c.first, c.second < 0 ? c.first : c.second);
    work += group_cost(sh.first, sh.second);
  }
  for (int k : i32) work += group_cost(B->hp[k].la, B->hp[k].lb);
  // Walks in the fill kernel: batches with direction planes whose int16 fill
  // cannot saturate (a rerun would replace some pairs' planes after the fill),
  // and more than one fill task: a lone task's walk can only start when its
  // fill ends, and the traceback kernel walks it sooner (a 300 x 320 pair:
  // 0.385 ms from run to results instead of 0.402).
  const char *wenv = getenv("TWEAKALIGN_WALK_IN_FILL");  // "0": walks only after the fill
  B->walk = B->trace && !B->may_saturate && !(wenv && wenv[0] == '0') && couples.size() + i32.size() > 1 &&
            walk_lds_bytes(2, B->Kp) <= ctx->lds_max;
  const int wpc = fill_waves_per_cu(2, B->Kp, !B->trace, waves_wanted(B), B->walk);
  const double slots = (double)ctx->cus * wpc * NGW;
  const double mean_slot = work / slots;
  const bool few = B->wave_tasks || couples.size() + i32.size() <= (size_t)ctx->cus * wpc;
  std::vector<Unit> u2, u1, u8;
  pl.tasks2[0].reserve(couples.size());
  pl.tasks1[0].reserve(i32.size());
  const double lim8 = !few ? eight_lane_limit() * mean_slot : -1.0;
  for (auto &c : couples) {
    const bool single = c.second < 0;
    const int hi = single ? c.first : c.second;
    auto sh = shape(c.first, hi);
    if (pick_gg(sh.first, sh.second, mean_slot, few) == GWAVE)
      add_task(pl, pl.tasks2[0], 2, GWAVE, B->trace, B->hp, c.first, hi, single, B->Kp);
    else if (cost8(sh.first, sh.second) <= lim8 && ceil_div(sh.first, 8 * R) <= (int)STRIPS_TOP)
      u8.push_back(Unit{c.first, hi, single, sh.first, sh.second});  // (strip counts fit the unit sort key)
    else
      u2.push_back(Unit{c.first, hi, single, sh.first, sh.second});
  }
  for (int k : i32) {
    if (pick_gg(B->hp[k].la, B->hp[k].lb, mean_slot, few) == GWAVE)
      add_task(pl, pl.tasks1[0], 1, GWAVE, B->trace, B->hp, k, k, true, B->Kp);
    else
      u1.push_back(Unit{k, k, true, B->hp[k].la, B->hp[k].lb});
  }
  tr.mark("task kinds");
  add_units(pl, pl.tasks2[1], 2, G, u2, local, B->trace, B->hp, B->Kp);
  add_units(pl, pl.tasks2[2], 2, 8, u8, local, B->trace, B->hp, B->Kp);
  add_units(pl, pl.tasks1[1], 1, G, u1, local, B->trace, B->hp, B->Kp);
  tr.mark("wave units");
  // costlier (more 8-step blocks, S * NB < 2^40) first, stable
  auto by_cost = [](std::vector<Task> &v) {
    std::vector<uint64_t> key(v.size());
    for (size_t t = 0; t < v.size(); ++t) key[t] = (1ull << 40) - (uint64_t)v[t].S * v[t].NB;
    gather_by_key(v, std::move(key));
  };
  by_cost(pl.tasks2[0]);
  by_cost(pl.tasks1[0]);
  // With few tasks the run lasts as long as its longest whole-wave task, so
  // tasks of two or more strips get a workgroup's fill waves each (strips
  // pipelined across them), as long as there are at most two per CU.
  const char *cenv = getenv("TWEAKALIGN_COOP");  // "0": never
  if (B->walk && few && !(cenv && cenv[0] == '0')) {
    auto multi = [](const Task &t) { return t.S >= 2; };
    const size_t c2 = std::count_if(pl.tasks2[0].begin(), pl.tasks2[0].end(), multi);
    const size_t c1 = std::count_if(pl.tasks1[0].begin(), pl.tasks1[0].end(), multi);
    if (c2 + c1 > 0 && c2 + c1 <= (size_t)(2 * ctx->cus)) {
      std::stable_partition(pl.tasks2[0].begin(), pl.tasks2[0].end(), multi);
      std::stable_partition(pl.tasks1[0].begin(), pl.tasks1[0].end(), multi);
      B->ncoop2 = (int32_t)c2;
      B->ncoop1 = (int32_t)c1;
    }
  }

  tr.mark("tasks");
  // Long pairs: strips of all pairs form one queue, largest pairs first.
  {
    std::vector<int> lk;
    for (int32_t k = 0; k < n; ++k)
      if (B->hp[k].path == 3) lk.push_back(k);
    std::stable_sort(lk.begin(), lk.end(), [&](int x, int y) {
      return (int64_t)B->hp[x].la * B->hp[x].lb > (int64_t)B->hp[y].la * B->hp[y].lb;
    });
    int max_rl = 0, max_n = 0;
    for (int k : lk) {
      const HostPair &h = B->hp[k];
      LongPair L{};
      L.pid = k; L.m = h.la; L.n = h.lb; L.S = ceil_div(h.la, LST);
      L.a_off = h.a_off; L.b_off = h.b_off;
      L.rl = ((h.lb + 7) & ~7) + 80;  // block loads / stores run up to 71 columns past n
      L.ring_off = B->lring_elems;
      if (L.S > 1) B->lring_elems += 4ull * L.rl;
      L.ck_off = B->lck_elems;
      if (p->traceback) B->lck_elems += (uint64_t)((L.S - 1) / KC) * 2 * L.rl;
      L.colck_off = B->lcol_elems;
      if (p->traceback) B->lcol_elems += (uint64_t)((L.n - 1) / CKX) * L.S * LST;
      L.prog_off = B->lprog_n;
      B->lprog_n += L.S;
      B->litem_first.push_back((int32_t)B->litems.size());
      for (int s = 0; s < L.S; ++s) B->litems.push_back(LongItem{(int32_t)B->longs.size(), s});
      B->longs.push_back(L);
      pl.ptb[k].P = 3;
      max_rl = std::max(max_rl, L.rl);
      max_n = std::max(max_n, L.n);
    }
    if (!lk.empty() && p->traceback) {
      B->wring_stride = 4ull * max_rl;
      B->wtb_stride = (uint64_t)KC * GWAVE * ((std::min(max_n, LREG) + GWAVE + 7) / 8) * 32;
      B->long_wgs = std::max(1, std::min((int)lk.size(), 2 * ctx->cus));
    }
    if (const char *e = getenv("TWEAKALIGN_SPIN_LIMIT")) B->spin_limit = std::max(1, atoi(e));
    if (const char *e = getenv("TWEAKALIGN_WALK_GUARD")) B->guard_cap = std::max(0, atoi(e));
  }

  tr.mark("long pairs");
  // Upload: every host-built table in one pinned block.
  std::vector<int8_t> table = padded_table(B->alph), table_nib;
  BulkUpload up;
  up.add(&B->d_table, table);
  if (B->nib && !B->trace) {
    table_nib = nibble_table(B->alph);
    up.add(&B->d_table_nib, table_nib);
  }
  for (int g = 0; g < 3; ++g) {
    up.add(&B->d_tasks2[g], pl.tasks2[g]);
    up.add(&B->d_tasks1[g], pl.tasks1[g]);
  }
  up.add(&B->d_ptb, pl.ptb);
  up.add(&B->d_pairs, B->pairs_h);
  CodeJob *d_jobs = nullptr;  // freed after the final drain
  up.add(&d_jobs, pl.cjobs);
  up.add(&B->d_fo_init, B->fo_init);
  std::vector<int32_t> order, rest;
  {
    std::vector<uint64_t> key(n);  // longer pairs (la + lb <= 2,000,000) first
    for (int32_t k = 0; k < n; ++k) key[k] = 2000000 - (uint64_t)(B->hp[k].la + B->hp[k].lb);
    order = order_by_key(std::move(key));
    up.add(&B->d_order, order);
    if (B->walk) {  // pairs without a fill task: empty and long pairs
      for (int32_t k : order)
        if (pl.ptb[k].P != 1 && pl.ptb[k].P != 2) rest.push_back(k);
      B->n_rest = (int32_t)rest.size();
      up.add(&B->d_order_rest, rest);
    }
  }
  up.add(&B->d_lp, B->longs);
  up.add(&B->d_litems, B->litems);
  rc = rc ? rc : up.commit(st);
  rc = rc ? rc : expand_jobs(d_jobs, pl.cjobs.size(), pl.rows_n, pl.cols_n, B->d_pairs, B->d_codes, B->Kp,
                             &B->d_rowcodes, &B->d_colcodes, st);
  tr.mark("uploads");
  rc = rc ? rc : dalloc(&B->d_ring, pl.ring_elems);
  rc = rc ? rc : dalloc(&B->d_tb, pl.tb_bytes);
  rc = rc ? rc : dalloc(&B->d_fo, (size_t)std::max(n, 1));
  rc = rc ? rc : dalloc(&B->d_res, (size_t)std::max(n, 1));
  rc = rc ? rc : dalloc(&B->d_cig, B->cig_bytes);
  rc = rc ? rc : dalloc(&B->d_ctr, 16);
  if (!rc && hipMemsetAsync(B->d_ctr, 0, 16 * sizeof(int32_t), st) != hipSuccess) rc = TA_E_DEVICE;
  rc = rc ? rc : dalloc(&B->d_lring, B->lring_elems);
  rc = rc ? rc : dalloc(&B->d_lck, B->lck_elems);
  rc = rc ? rc : dalloc(&B->d_lcol, B->lcol_elems);
  rc = rc ? rc : dalloc(&B->d_lprog, (size_t)B->lprog_n);
  rc = rc ? rc : dalloc(&B->d_wring, B->wring_stride * B->long_wgs);
  rc = rc ? rc : dalloc(&B->d_wtb, B->wtb_stride * B->long_wgs);
  for (auto &e : B->ev)
    if (!rc && hipEventCreate(&e) != hipSuccess) rc = TA_E_DEVICE;
  tr.mark("allocations");
  // Drained on every path: the blocks of a failed creation go back to the
  // cache only after the copies queued into them.
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = TA_E_DEVICE;
  // the letter check, read once the stream has drained (a copy into
  // pageable memory queued earlier would have waited for the upload)
  if (!rc && d_bad && hipMemcpy(&bad_letter, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
    rc = TA_E_DEVICE;
  dfree(d_map);
  dfree(d_bad);
  dfree(d_jobs);
  tr.mark("drain");
  if (!rc && bad_letter) rc = TA_E_ALPHABET;
  if (rc) {
    ta_batch_destroy(B);
    return rc;
  }
  *out = B;
  return TA_OK;
}

int64_t batch_tb_estimate(const ta_params *p, const ta_pair *pairs, int32_t n) {
  if (!p || (p->mode != TA_LOCAL && !p->traceback)) return 0;
  int64_t t = 0;
  for (int32_t k = 0; k < n; ++k) {
    const int la = pairs[k].la, lb = pairs[k].lb;
    if (la <= 0 || lb <= 0) continue;
    if (long_pair(*p, la, lb)) {  // row + column checkpoints, one traceback workgroup's planes
      const int64_t rl = ((lb + 7) & ~7) + 80, S = ceil_div(la, LST);
      t += (S - 1) / KC * 8 * rl + (int64_t)((lb - 1) / CKX) * S * LST * 8 +
           (int64_t)KC * GWAVE * ((std::min(lb, LREG) + GWAVE + 7) / 8) * 32;
    } else {
      t += (int64_t)ceil_div(la, STRIP) * ceil_div(lb + G, 8) * (G * R * 4);
    }
  }
  return t;
}

int ta_batch_info(ta_batch *B, int64_t *cells, int64_t *tb_bytes, int32_t *tasks) {
  if (!B) return TA_E_INVALID;
  if (cells) *cells = B->cells;
  if (tb_bytes)
    *tb_bytes = (int64_t)(B->plan.tb_bytes + B->wtb_stride * B->long_wgs + 4 * B->lck_elems + 8 * B->lcol_elems);
  if (tasks)
    *tasks = (int32_t)(B->plan.tasks2[0].size() + B->plan.tasks2[1].size() + B->plan.tasks2[2].size() +
                       B->plan.tasks1[0].size() + B->plan.tasks1[1].size() + B->litems.size());
  return TA_OK;
}

// ---------------------------------------------------------------------------
// Launch helpers.
namespace {

TbArgs tb_args(ta_batch *B) {
  TbArgs a;
  a.pairs = B->d_pairs; a.npairs = B->n; a.order = nullptr;
  a.fo = B->d_fo; a.ptb = B->d_ptb;
  a.codes = B->d_codes; a.table = B->d_table; a.Kp = B->Kp; a.tb = B->d_tb; a.cig = B->d_cig;
  a.res = B->d_res;
  a.local = B->p.mode == TA_LOCAL; a.open = B->p.gap_open; a.ext = B->p.gap_extend;
  a.do_cigar = B->p.traceback;
  a.guard_cap = B->guard_cap;
  a.sys = 0;
  return a;
}

// `walk`: the batch's own fill (planes in d_tb); each workgroup's walker wave
// walks the pairs its fill wave finishes.
template <int P>
int launch_fill(ta_ctx *ctx, ta_batch *B, Task *const d_tasks[3], const int ntasks[3], uint16_t *rows,
                uint16_t *cols, uint32_t *ring, uint8_t *tb, int32_t *ctr, int max_groups, bool walk = false) {
  if (ntasks[0] + ntasks[1] + ntasks[2] == 0) return TA_OK;
  const bool nib = P == 2 && !B->trace;  // score-only packed kernels keep 4-bit scores
  FillArgs a;
  for (int g = 0; g < 3; ++g) {
    a.tasks[g] = d_tasks[g];
    a.ntasks[g] = ntasks[g];
  }
  a.ctr = ctr;
  a.ctr8 = B->d_ctr + (P == 2 ? 14 : 15);  // 8-lane units come only from the batch plan
  a.table = nib ? B->d_table_nib : B->d_table; a.Kp = B->Kp;
  a.bias = -B->alph.minS;
  a.rowcodes = rows; a.colcodes = cols; a.ring = ring; a.tb = tb; a.out = B->d_fo;
  a.open = B->p.gap_open; a.ext = B->p.gap_extend;
  a.sat_limit = 32767 - 2 * B->alph.maxS;
  walk = walk && B->trace;
  a.walk = walk;
  if (walk) a.wk = tb_args(B);
  a.ncoop = walk ? (P == 2 ? B->ncoop2 : B->ncoop1) : 0;
  a.coop_ctr = B->d_ctr + (P == 2 ? 12 : 13);
  a.err = B->d_ctr + 10;
  a.spin_limit = B->spin_limit * 16;
  const int lds = walk ? walk_lds_bytes(P, B->Kp) : fill_lds_bytes(P, B->Kp, nib);
  const int per_cu = fill_waves_per_cu(P, B->Kp, nib, waves_wanted(B), walk);
  int groups = ntasks[0] + ntasks[1] + ntasks[2];  // one wave per whole-wave task or wave unit
  if (max_groups > 0) groups = std::min(groups, max_groups);
  groups = std::max(1, std::min(groups, ctx->cus * per_cu));
  if (walk) {  // WALK_FW fill waves per workgroup; a cooperative task takes one
    groups = a.ncoop + (groups - std::min(groups, a.ncoop) + WALK_FW - 1) / WALK_FW;
    groups = std::max(1, std::min(groups, ctx->cus));
  }
  const bool local = B->p.mode == TA_LOCAL;
  void (*k)(FillArgs) = nullptr;
  if (local) k = B->trace ? fill_kernel<P, true, true> : fill_kernel<P, true, false>;
  else k = B->trace ? fill_kernel<P, false, true> : fill_kernel<P, false, false>;
  (void)hipGetLastError();  // drop an earlier, already-handled error
  hipLaunchKernelGGL(k, dim3(groups), dim3(walk ? 128 * WALK_FW : 64), lds, B->stream, a);
  HIPCHK(hipGetLastError());
  return TA_OK;
}

int launch_traceback(ta_ctx *ctx, ta_batch *B, const PairDev *pairs, const PairTB *ptb, const uint8_t *tb,
                     int first, int count) {
  if (count <= 0) return TA_OK;
  TbArgs a = tb_args(B);
  const bool whole = first == 0 && count == B->n;
  a.pairs = whole ? pairs : pairs + first; a.npairs = count;
  a.order = whole ? B->d_order : nullptr;
  a.fo = whole ? B->d_fo : B->d_fo + first;
  a.ptb = whole ? ptb : ptb + first;
  a.codes = B->d_codes; a.table = B->d_table; a.Kp = B->Kp; a.tb = tb; a.cig = B->d_cig;
  a.res = whole ? B->d_res : B->d_res + first;
  a.local = B->p.mode == TA_LOCAL; a.open = B->p.gap_open; a.ext = B->p.gap_extend;
  a.do_cigar = B->p.traceback;
  a.guard_cap = B->guard_cap;
  (void)hipGetLastError();
  hipLaunchKernelGGL(traceback_kernel<0>, dim3((count + TB_WALKERS - 1) / TB_WALKERS), dim3(TB_WALKERS), 0,
                     B->stream, a);
  HIPCHK(hipGetLastError());
  return TA_OK;
}

// Long pairs [lp0, lp0 + nlp) of the batch (queue items [it0, it0 + nit)).
LongArgs long_args(ta_batch *B, int it0, int nit, int lp0, int nlp) {
  LongArgs a{};
  a.lp = B->d_lp + lp0; a.nlp = nlp;
  a.items = B->d_litems + it0; a.nitems = nit;
  a.ctr = B->d_ctr + 8;
  a.table = B->d_table; a.Kp = B->Kp; a.codes = B->d_codes;
  a.ring = B->d_lring; a.ckpt = B->d_lck; a.prog = B->d_lprog; a.colck = B->d_lcol;
  a.wring = B->d_wring; a.wring_stride = B->wring_stride;
  a.wtb = B->d_wtb; a.wtb_stride = B->wtb_stride;
  a.pd = B->d_pairs; a.fo = B->d_fo; a.res = B->d_res; a.cig = B->d_cig;
  a.open = B->p.gap_open; a.ext = B->p.gap_extend;
  a.ckpt_on = B->p.traceback;
  a.spin_limit = B->spin_limit;
  a.err = B->d_ctr + 10;
  return a;
}

// Forward pass of long pairs: counters cleared, then one wave per workgroup
// pulling strips.  Items index pairs relative to lp0 = 0, so a subset of
// items must keep the full pair array.
int launch_long_fwd(ta_ctx *ctx, ta_batch *B, int it0, int nit, bool clear) {
  if (nit <= 0) return TA_OK;
  if (clear) {  // the per-pair (plain) form; a batch run starts with init_kernel
    HIPCHK(hipMemsetAsync(B->d_ctr + 8, 0, 3 * sizeof(int32_t), B->stream));
    HIPCHK(hipMemsetAsync(B->d_lprog, 0, sizeof(int32_t) * B->lprog_n, B->stream));
  }
  LongArgs a = long_args(B, it0, nit, 0, (int)B->longs.size());
  const int lds = long_fwd_lds(B->Kp);
  const int per_cu = std::max(1, std::min(16, (160 * 1024) / lds));  // 4 waves per SIMD (117 VGPRs)
  const int grid = std::max(1, std::min(nit, ctx->cus * per_cu));
  (void)hipGetLastError();
  hipLaunchKernelGGL(long_fwd_kernel, dim3(grid), dim3(64), lds, B->stream, a);
  HIPCHK(hipGetLastError());
  return TA_OK;
}

// Traceback of long pairs [lp0, lp0 + nlp): one workgroup of LW waves per pair.
int launch_long_tb(ta_ctx *ctx, ta_batch *B, int lp0, int nlp, bool clear) {
  if (nlp <= 0 || !B->p.traceback) return TA_OK;
  if (clear) HIPCHK(hipMemsetAsync(B->d_ctr + 9, 0, sizeof(int32_t), B->stream));
  LongArgs a = long_args(B, 0, 0, lp0, nlp);
  const int lds = long_tb_lds(B->Kp);
  const int grid = std::max(1, std::min(nlp, B->long_wgs));
  (void)hipGetLastError();
  hipLaunchKernelGGL(long_tb_kernel, dim3(grid), dim3(64 * LW), lds, B->stream, a);
  HIPCHK(hipGetLastError());
  return TA_OK;
}

// Pairs whose int16 pass saturated go through the int32 kernel.
int rerun_saturated(ta_ctx *ctx, ta_batch *B) {
  std::vector<FillOut> fo(B->n);
  HIPCHK(hipMemcpyAsync(fo.data(), B->d_fo, sizeof(FillOut) * B->n, hipMemcpyDeviceToHost, B->stream));
  HIPCHK(hipStreamSynchronize(B->stream));
  std::vector<int> bad;
  for (int k = 0; k < B->n; ++k)
    if (fo[k].flags & TA_F_RERUN) bad.push_back(k);
  free_rerun(B);
  if (bad.empty()) return TA_OK;
  Plan pl;
  pl.ptb = B->plan.ptb;
  std::vector<Unit> u;
  for (int k : bad) {
    if (pick_gg(B->hp[k].la, B->hp[k].lb, 0.0, bad.size() <= (size_t)ctx->cus * 4) == GWAVE) {
      add_task(pl, pl.tasks1[0], 1, GWAVE, B->trace, B->hp, k, k, true, B->Kp);
      pl.tasks1[0].back().flags_or = TA_F_RERUN;
    } else {
      u.push_back(Unit{k, k, true, B->hp[k].la, B->hp[k].lb});
    }
  }
  add_units(pl, pl.tasks1[1], 1, G, u, B->p.mode == TA_LOCAL, B->trace, B->hp, B->Kp);
  for (Task &t : pl.tasks1[1]) t.flags_or = TA_F_RERUN;
  std::vector<Task> all(pl.tasks1[0]);
  all.insert(all.end(), pl.tasks1[1].begin(), pl.tasks1[1].end());
  hipStream_t st = B->stream;
  int rc = TA_OK;
  rc = rc ? rc : upload(&B->d_tasks_r, all, st);
  rc = rc ? rc : expand_codes(pl.cjobs, pl.rows_n, pl.cols_n, B->d_pairs, B->d_codes, B->Kp, &B->d_rowcodes_r,
                              &B->d_colcodes_r, st);
  rc = rc ? rc : dalloc(&B->d_ring_r, pl.ring_elems);
  rc = rc ? rc : dalloc(&B->d_tb_r, pl.tb_bytes);
  if (rc) return rc;
  B->n_rerun = (int)bad.size();
  // Rerun pairs read their planes from the rerun arena: point their PairTB there
  // by offsetting against the main arena base.
  std::vector<PairTB> ptb(B->plan.ptb);
  for (int k : bad) {
    ptb[k] = pl.ptb[k];
    ptb[k].tb_off += (uint64_t)(B->d_tb_r - B->d_tb);
  }
  HIPCHK(hipMemcpyAsync(B->d_ptb, ptb.data(), sizeof(PairTB) * B->n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(B->d_ctr + 4, 0, 2 * sizeof(int32_t), st));
  Task *const tk[3] = {B->d_tasks_r, B->d_tasks_r + pl.tasks1[0].size(), nullptr};
  const int nt[3] = {(int)pl.tasks1[0].size(), (int)pl.tasks1[1].size() / NGW, 0};
  return launch_fill<1>(ctx, B, tk, nt, B->d_rowcodes_r, B->d_colcodes_r, B->d_ring_r, B->d_tb_r, B->d_ctr + 4, 0);
}

}  // namespace

namespace {

int launch_init(ta_batch *B) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(init_kernel, dim3(std::max(1, std::min(1024, (B->n + 255) / 256))), dim3(256), 0, B->stream,
                     B->d_ctr, 16, B->d_fo, B->d_fo_init, B->n, B->d_lprog, B->lprog_n);
  HIPCHK(hipGetLastError());
  return TA_OK;
}

int launch_fills(ta_ctx *ctx, ta_batch *B) {
  const Plan &pl = B->plan;
  const int n2[3] = {(int)pl.tasks2[0].size(), (int)pl.tasks2[1].size() / NGW, (int)pl.tasks2[2].size() / 8};
  const int n1[3] = {(int)pl.tasks1[0].size(), (int)pl.tasks1[1].size() / NGW, 0};
  int rc = launch_fill<2>(ctx, B, B->d_tasks2, n2, B->d_rowcodes, B->d_colcodes, B->d_ring, B->d_tb, B->d_ctr, 0,
                          B->walk);
  if (!rc) rc = launch_fill<1>(ctx, B, B->d_tasks1, n1, B->d_rowcodes, B->d_colcodes, B->d_ring, B->d_tb,
                               B->d_ctr + 2, 0, B->walk);
  if (!rc) rc = launch_long_fwd(ctx, B, 0, (int)B->litems.size(), false);
  return rc;
}

int launch_tracebacks(ta_ctx *ctx, ta_batch *B) {
  int rc = TA_OK;
  if (!B->walk) {
    rc = launch_traceback(ctx, B, B->d_pairs, B->d_ptb, B->d_tb, 0, B->n);
  } else if (B->n_rest) {  // the fill walked every pair it filled
    TbArgs a = tb_args(B);
    a.npairs = B->n_rest;
    a.order = B->d_order_rest;
    (void)hipGetLastError();
    hipLaunchKernelGGL(traceback_kernel<0>, dim3((B->n_rest + TB_WALKERS - 1) / TB_WALKERS), dim3(TB_WALKERS), 0,
                       B->stream, a);
    HIPCHK(hipGetLastError());
  }
  if (!rc) rc = launch_long_tb(ctx, B, 0, (int)B->longs.size(), false);
  return rc;
}

// A captured run: init, fills, tracebacks -- kernels only.  Batches that may
// need the int32 rerun are not captured (the rerun is a host decision).
int run_graph(ta_ctx *ctx, ta_batch *B) {
  hipStream_t st = B->stream;
  if (!B->gexec) {
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = launch_init(B);
    if (!rc) rc = launch_fills(ctx, B);
    if (!rc) rc = launch_tracebacks(ctx, B);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : TA_E_DEVICE;
    }
    B->graph = g;
    HIPCHK(hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0));
  }
  HIPCHK(hipEventRecord(B->ev[0], st));
  HIPCHK(hipGraphLaunch(B->gexec, st));
  HIPCHK(hipEventRecord(B->ev[3], st));
  B->phases_timed = false;
  return TA_OK;
}

}  // namespace

int ta_batch_run(ta_ctx *ctx, ta_batch *B) {
  if (!ctx || !B) return TA_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  if (B->graph_on && !B->may_saturate) return run_graph(ctx, B);
  hipStream_t st = B->stream;
  int rc = launch_init(B);
  if (rc) return rc;
  HIPCHK(hipEventRecord(B->ev[0], st));
  rc = launch_fills(ctx, B);
  if (rc) return rc;
  HIPCHK(hipEventRecord(B->ev[1], st));
  if (B->may_saturate) {
    rc = rerun_saturated(ctx, B);
    if (rc) return rc;
  } else if (B->n_rerun) {
    free_rerun(B);
  }
  HIPCHK(hipEventRecord(B->ev[2], st));
  rc = launch_tracebacks(ctx, B);
  if (rc) return rc;
  HIPCHK(hipEventRecord(B->ev[3], st));
  B->phases_timed = true;
  return TA_OK;
}

int ta_batch_graph(ta_batch *B, int32_t on) {
  if (!B) return TA_E_INVALID;
  B->graph_on = on != 0;
  return TA_OK;
}

int ta_batch_wait(ta_ctx *ctx, ta_batch *B) {
  if (!ctx || !B) return TA_E_INVALID;
  HIPCHK(hipStreamSynchronize(B->stream));
  float t = 0;
  B->ms[0] = B->ms[1] = B->ms[2] = -1.0;  // a captured run is timed as a whole only
  if (B->phases_timed) {
    if (hipEventElapsedTime(&t, B->ev[0], B->ev[1]) == hipSuccess) B->ms[0] = t;
    if (hipEventElapsedTime(&t, B->ev[1], B->ev[2]) == hipSuccess) B->ms[1] = t;
    if (hipEventElapsedTime(&t, B->ev[2], B->ev[3]) == hipSuccess) B->ms[2] = t;
  }
  if (hipEventElapsedTime(&t, B->ev[0], B->ev[3]) == hipSuccess) B->ms[3] = t;
  if (!B->longs.empty() || B->ncoop2 || B->ncoop1) {
    int32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, B->d_ctr + 10, sizeof(err), hipMemcpyDeviceToHost, B->stream));
    HIPCHK(hipStreamSynchronize(B->stream));
    if (err) {
      fprintf(stderr, "tweakalign: a wait on another wave's progress gave up (code %d)\n", err);
      return TA_E_DEVICE;
    }
  }
  return TA_OK;
}

int ta_batch_timings(ta_batch *B, double *ms, int32_t cap) {
  if (!B || !ms) return TA_E_INVALID;
  int k = 0;
  for (; k < cap && k < 4; ++k) ms[k] = B->ms[k];
  return k;
}

// Binary runs ((len << 2) | op, op 0 M / 1 D / 2 I) -> CIGAR text.
int64_t cigar_text(const uint32_t *runs, int n, char *out) {
  static const char ops[4] = {'M', 'D', 'I', '?'};
  int64_t pos = 0;
  char tmp[12];
  for (int k = 0; k < n; ++k) {
    uint32_t len = runs[k] >> 2;
    int d = 0;
    do { tmp[d++] = (char)('0' + len % 10); len /= 10; } while (len);
    if (out)
      while (d) out[pos++] = tmp[--d];
    else
      pos += d;
    if (out) out[pos] = ops[runs[k] & 3];
    ++pos;
  }
  return pos;
}
namespace {
int64_t format_runs(const uint32_t *runs, int n, char *out) { return cigar_text(runs, n, out); }
}  // namespace

namespace {

// Gathers every pair's CIGAR runs (written backwards at the end of its
// region) into one dense array: one wave per pair, dword copies.
__global__ void gather_runs_kernel(const uint32_t *cig, const ResultDev *res, const int64_t *dst_off, uint32_t *dst,
                                   int n) {
  const int p = blockIdx.x;
  if (p >= n) return;
  const int nr = res[p].cigar_len / 4;
  if (nr <= 0) return;
  const uint32_t *src = cig + res[p].cigar_off / 4;
  uint32_t *d = dst + dst_off[p] / 4;
  for (int k = threadIdx.x; k < nr; k += blockDim.x) d[k] = src[k];
}

// Results and CIGAR runs of a finished batch on the host, in pinned staging
// blocks: `res` per pair, `runs` dense, pair k's runs at runs[off[k] / 4 ...]
// (cigar_len bytes).  Only the runs cross PCIe, not the per-pair worst-case
// regions.
struct Fetched {
  StageHold resb, runsb;
  const ResultDev *res = nullptr;
  const uint32_t *runs = nullptr;
  std::vector<int64_t> off;
};

int fetch_results(ta_batch *B, Fetched &F) {
  F.off.assign(B->n, 0);
  std::vector<int64_t> &off = F.off;
  hipStream_t st = B->stream;
  if (B->n) {
    F.resb.s = stage_take(sizeof(ResultDev) * B->n);
    if (!F.resb.s.p) return TA_E_NOMEM;
    F.resb.stream = st;
    F.resb.queued = true;
    HIPCHK(hipMemcpyAsync(F.resb.s.p, B->d_res, sizeof(ResultDev) * B->n, hipMemcpyDeviceToHost, st));
    F.res = (const ResultDev *)F.resb.s.p;
  }
  HIPCHK(hipStreamSynchronize(st));
  if (!B->p.traceback) return TA_OK;
  const ResultDev *res = F.res;
  int64_t total = 0;
  for (int32_t k = 0; k < B->n; ++k) {
    off[k] = total;
    if (res[k].cigar_len > 0) total += res[k].cigar_len;
  }
  if (!total) return TA_OK;
  F.runsb.s = stage_take((size_t)total);
  if (!F.runsb.s.p) return TA_E_NOMEM;
  F.runsb.stream = st;
  F.runsb.queued = true;
  F.runs = (const uint32_t *)F.runsb.s.p;
  int64_t *d_off = nullptr;
  uint32_t *d_runs = nullptr;
  int rc = upload(&d_off, off, st);
  rc = rc ? rc : dalloc(&d_runs, (size_t)(total / 4));
  if (!rc) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(gather_runs_kernel, dim3(B->n), dim3(64), 0, st, (const uint32_t *)B->d_cig, B->d_res,
                       d_off, d_runs, B->n);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(F.runsb.s.p, d_runs, (size_t)total, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = TA_E_DEVICE;
  }
  (void)hipStreamSynchronize(st);  // nothing in flight reads the buffers once they return to the cache
  dfree(d_off);
  dfree(d_runs);
  return rc;
}

void fill_result(ta_result &o, const ResultDev &r) {
  o.score = r.score; o.a_begin = r.a_begin; o.a_end = r.a_end; o.b_begin = r.b_begin; o.b_end = r.b_end;
  o.flags = r.flags;
  o.cigar_off = 0;
  o.cigar_len = 0;
}

}  // namespace

int ta_batch_results(ta_ctx *ctx, ta_batch *B, ta_result *out, char *cigar_buf, int64_t cap, int64_t *used) {
  if (!ctx || !B || (B->n && !out)) return TA_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  PhaseTrace tr("results");
  Fetched F;
  int rc = fetch_results(B, F);
  tr.mark("fetch");
  if (rc) return rc;
  const ResultDev *res = F.res;
  const uint32_t *runs = F.runs;
  const std::vector<int64_t> &off = F.off;
  // CIGAR text lengths, their offsets, then the text, the first and last
  // passes on the host threads.
  const bool tb = B->p.traceback != 0;
  parallel_for((size_t)B->n, [&](size_t k) {
    const ResultDev &r = res[k];
    fill_result(out[k], r);
    if (tb && r.cigar_len > 0) out[k].cigar_len = (int32_t)format_runs(runs + off[k] / 4, r.cigar_len / 4, nullptr);
  });
  int64_t pos = 0;
  for (int32_t k = 0; k < B->n; ++k) {
    out[k].cigar_off = pos;
    pos += out[k].cigar_len;
  }
  if (used) *used = pos;
  if (pos > 0 && (!cigar_buf || pos > cap)) return TA_E_SPACE;
  parallel_for((size_t)B->n, [&](size_t k) {
    if (out[k].cigar_len > 0)
      format_runs(runs + off[k] / 4, res[k].cigar_len / 4, cigar_buf + out[k].cigar_off);
  });
  tr.mark("format");
  for (int32_t k = 0; k < B->n; ++k)
    if (out[k].flags & TA_F_WALK_FAILED) return TA_E_WALK;
  return rc;
}

int batch_results_text(ta_ctx *ctx, ta_batch *B, std::vector<ta_result> &out, std::string &text) {
  if (!ctx || !B) return TA_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  PhaseTrace tr("results");
  Fetched F;
  const int rc = fetch_results(B, F);
  tr.mark("fetch");
  if (rc) return rc;
  const ResultDev *res = F.res;
  const uint32_t *runs = F.runs;
  const std::vector<int64_t> &off = F.off;
  // lengths, offsets, then the text, as ta_batch_results
  out.resize(B->n);
  const bool tb = B->p.traceback != 0;
  parallel_for((size_t)B->n, [&](size_t k) {
    const ResultDev &r = res[k];
    fill_result(out[k], r);
    if (tb && r.cigar_len > 0) out[k].cigar_len = (int32_t)format_runs(runs + off[k] / 4, r.cigar_len / 4, nullptr);
  });
  int64_t pos = 0;
  for (int32_t k = 0; k < B->n; ++k) {
    out[k].cigar_off = pos;
    pos += out[k].cigar_len;
  }
  text.assign((size_t)pos, '\0');
  parallel_for((size_t)B->n, [&](size_t k) {
    if (out[k].cigar_len > 0) format_runs(runs + off[k] / 4, res[k].cigar_len / 4, &text[out[k].cigar_off]);
  });
  tr.mark("format");
  return TA_OK;
}

int ta_align_batch(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n, ta_result *out,
                   char *cigar_buf, int64_t cap, int64_t *used) {
  ta_batch *B = nullptr;
  int rc = ta_batch_create(ctx, p, pairs, n, &B);
  if (rc) return rc;
  rc = ta_batch_run(ctx, B);
  if (!rc) rc = ta_batch_wait(ctx, B);
  if (!rc) rc = ta_batch_results(ctx, B, out, cigar_buf, cap, used);
  ta_batch_destroy(B);
  return rc;
}

// ---------------------------------------------------------------------------
// Resident sequences: codes encoded on the host once, appended to the
// context's device arena (grown by copying into a larger one; batches and
// served requests that still read the old arena hold it until they are
// freed), aligned by handle.  Codes added since the last upload wait in host
// memory and go up in one copy (seq_flush) when a call resolves handles, or
// before an add once 64 MB are pending.

// Uploads the pending codes (the caller holds seq_mu).
static int seq_flush(ta_ctx *ctx) {
  if (ctx->seq_pending.empty()) return TA_OK;
  if (hipMemcpy(ctx->seq_arena.get() + ctx->seq_flushed, ctx->seq_pending.data(), ctx->seq_pending.size(),
                hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    return TA_E_DEVICE;
  }
  ctx->seq_flushed = ctx->seq_used;  // the 32 zero bytes after it are on the device too
  ctx->seq_pending.clear();
  return TA_OK;
}

int ta_seq_add(ta_ctx *ctx, int32_t alphabet, const char *letters, int32_t len, ta_seq *out) {
  if (!ctx || !out || len < 0 || len > 1000000 || (len > 0 && !letters)) return TA_E_INVALID;
  if (alphabet != TA_PROTEIN && alphabet != TA_DNA) return TA_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  const Alphabet A = make_alphabet(alphabet, 1, -1);
  // the codes, code 0 up to the next 16-byte start, then the walkers' slack
  const size_t span = ((size_t)len + 15) & ~(size_t)15;
  int32_t cnt[32] = {0};
  for (int32_t i = 0; i < len; ++i) {
    const int c = A.map[(unsigned char)letters[i]];
    if (c < 0) return TA_E_ALPHABET;
    ++cnt[c];
  }
  std::lock_guard<std::mutex> g(ctx->seq_mu);
  if (ctx->seq_pending.size() >= ((size_t)64 << 20) && seq_flush(ctx)) return TA_E_DEVICE;
  const size_t off = ctx->seq_used, need = off + span + 32;
  if (need > ctx->seq_cap) {
    // a larger arena with the uploaded contents (the pending ones follow at
    // the next flush); the old one is freed once no batch holds it
    const size_t cap = std::max(need, std::max(2 * ctx->seq_cap, (size_t)16 << 20));
    uint8_t *d = nullptr;
    if (hipMalloc((void **)&d, cap) != hipSuccess) {
      (void)hipGetLastError();
      return TA_E_NOMEM;
    }
    std::shared_ptr<uint8_t> arena(d, [](uint8_t *q) { (void)hipFree(q); });
    if (ctx->seq_flushed &&
        hipMemcpy(d, ctx->seq_arena.get(), ctx->seq_flushed, hipMemcpyDeviceToDevice) != hipSuccess) {
      (void)hipGetLastError();
      return TA_E_DEVICE;
    }
    ctx->seq_arena = std::move(arena);
    ctx->seq_cap = cap;
  }
  // pending bytes [seq_flushed, off + span + 32): this sequence's codes
  // replace the previous one's trailing slack
  std::vector<uint8_t> &pend = ctx->seq_pending;
  const size_t at = off - ctx->seq_flushed;
  pend.resize(at + span + 32);
  std::fill(pend.begin() + at, pend.end(), (uint8_t)0);
  for (int32_t i = 0; i < len; ++i) pend[at + i] = (uint8_t)A.map[(unsigned char)letters[i]];
  ctx->seq_used = off + span;
  ctx->seqs.push_back(ta_ctx::SeqRec{off, len, alphabet});
  ctx->seq_counts.insert(ctx->seq_counts.end(), cnt, cnt + 32);
  *out = (ta_seq)(ctx->seqs.size() - 1);
  return TA_OK;
}

int ta_seq_reset(ta_ctx *ctx) {
  if (!ctx) return TA_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->seq_mu);
  ctx->seq_arena.reset();  // freed now, or when the last batch that holds it is
  ctx->seq_cap = ctx->seq_used = ctx->seq_flushed = 0;
  ctx->seq_pending.clear();
  ctx->seq_pending.shrink_to_fit();
  ctx->seqs.clear();
  ctx->seq_counts.clear();
  return TA_OK;
}

int resident_resolve(ta_ctx *ctx, const ta_params *p, const ta_seq_pair *pairs, int32_t n, ResidentBatch &out) {
  if (!ctx || !p || n < 0 || (n > 0 && !pairs) || params_check(p)) return TA_E_INVALID;
  const Alphabet A = make_alphabet(p->alphabet, p->match, p->mismatch);
  int best[32] = {0};  // best score of each residue against any letter (0 if none is positive)
  for (int r = 0; r < A.K; ++r)
    for (int c = 0; c < A.K; ++c) best[r] = std::max<int>(best[r], A.S[r * A.K + c]);
  std::lock_guard<std::mutex> g(ctx->seq_mu);
  const int64_t ns = (int64_t)ctx->seqs.size();
  auto bound = [&](ta_seq h) {
    const int32_t *cnt = ctx->seq_counts.data() + 32 * (size_t)h;
    int64_t sum = 0;
    for (int r = 0; r < A.K; ++r) sum += (int64_t)cnt[r] * best[r];
    return sum;
  };
  out = ResidentBatch();
  out.a_off.resize(n); out.b_off.resize(n);
  out.la.resize(n); out.lb.resize(n);
  out.a_bound.resize(n); out.b_bound.resize(n);
  for (int32_t k = 0; k < n; ++k) {
    const ta_seq a = pairs[k].a, b = pairs[k].b;
    if (a < 0 || b < 0 || a >= ns || b >= ns || ctx->seqs[a].alphabet != p->alphabet ||
        ctx->seqs[b].alphabet != p->alphabet)
      return TA_E_INVALID;
    out.a_off[k] = ctx->seqs[a].off; out.b_off[k] = ctx->seqs[b].off;
    out.la[k] = ctx->seqs[a].len; out.lb[k] = ctx->seqs[b].len;
    out.a_bound[k] = bound(a); out.b_bound[k] = bound(b);
  }
  if (n > 0 && !ctx->seq_arena) return TA_E_INVALID;
  if (n > 0 && seq_flush(ctx)) return TA_E_DEVICE;
  out.arena = ctx->seq_arena;
  return TA_OK;
}

int ta_align_resident(ta_ctx *ctx, const ta_params *p, const ta_seq_pair *pairs, int32_t n, ta_result *out,
                      char *cigar_buf, int64_t cap, int64_t *used) {
  if (!ctx || n < 0 || (n > 0 && (!pairs || !out))) return TA_E_INVALID;
  ResidentBatch rb;
  int rc = resident_resolve(ctx, p, pairs, n, rb);
  if (rc) return rc;
  ta_batch *B = nullptr;
  rc = batch_create_on(ctx, p, nullptr, n, ctx->stream, &B, nullptr, &rb);
  if (rc) return rc;
  rc = ta_batch_run(ctx, B);
  if (!rc) rc = ta_batch_wait(ctx, B);
  if (!rc) rc = ta_batch_results(ctx, B, out, cigar_buf, cap, used);
  ta_batch_destroy(B);
  return rc;
}

// ---------------------------------------------------------------------------
// Plain form: one fill launch (one group, one task) and one traceback launch
// per pair, serialised on the stream -- the baseline batching is measured
// against.
int ta_batch_run_plain(ta_ctx *ctx, ta_batch *B, int32_t first, int32_t count) {
  if (!ctx || !B || first < 0 || count < 0 || first + count > B->n) return TA_E_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = B->stream;
  if (!B->plain_ready) {
    Plan pl;
    pl.ptb.assign(B->n, PairTB{0, 0, 0, 0, 0, 0, 0});
    uint64_t max_tb = 0, max_ring = 0;
    // NGW entries per pair: the task, then fillers for the other groups of
    // its wave (a group task runs in a wave of NGW groups; fillers write
    // nothing).
    std::vector<Task> tasks, rtasks;
    std::vector<PairTB> ptb_r(B->n, PairTB{0, 0, 0, 0, 0, 0, 0});
    std::vector<int32_t> rerun_of(B->n, -1);
    auto fill_unit = [&](std::vector<Task> &v) {
      Task f = v.back();
      f.single = 2;
      for (int q = 1; q < NGW; ++q) v.push_back(f);
    };
    for (int k = 0; k < B->n; ++k) {
      const HostPair &h = B->hp[k];
      const int P = h.path == 1 ? 1 : 2;
      const uint64_t tb0 = pl.tb_bytes, r0 = pl.ring_elems;
      if (h.path == 0 || h.path == 3) {
        if (h.path == 3) pl.ptb[k].P = 3;  // the long-pair kernels write its result
        tasks.push_back(Task{});
        fill_unit(tasks);
        continue;
      }
      const int GG = pick_gg(h.la, h.lb, 0.0, true);
      // Every plain launch reuses the same scratch: rebase to 0.
      auto rebase = [&](std::vector<Task> &v) {
        Task &T = v.back();
        max_tb = std::max(max_tb, pl.tb_bytes - tb0);
        max_ring = std::max(max_ring, pl.ring_elems - r0);
        T.tb_off = 0; T.ring_off = 0;
        pl.ptb[k].tb_off = 0;
        pl.tb_bytes = tb0; pl.ring_elems = r0;
        fill_unit(v);
      };
      add_task(pl, tasks, P, GG, B->trace, B->hp, k, k, true, B->Kp);
      rebase(tasks);
      if (h.maysat) {  // the int32 task a saturated int16 fill reruns as
        const PairTB p16 = pl.ptb[k];
        rerun_of[k] = (int32_t)rtasks.size();
        add_task(pl, rtasks, 1, GG, B->trace, B->hp, k, k, true, B->Kp);
        rtasks.back().flags_or = TA_F_RERUN;
        rebase(rtasks);
        ptb_r[k] = pl.ptb[k];
        pl.ptb[k] = p16;
      }
    }
    int rc = TA_OK;
    rc = rc ? rc : upload(&B->d_plain_tasks, tasks, st);
    rc = rc ? rc : upload(&B->d_plain_rtasks, rtasks, st);
    rc = rc ? rc : upload(&B->d_pptb_r, ptb_r, st);
    rc = rc ? rc : expand_codes(pl.cjobs, pl.rows_n, pl.cols_n, B->d_pairs, B->d_codes, B->Kp, &B->d_prow,
                                &B->d_pcol, st);
    rc = rc ? rc : upload(&B->d_pptb, pl.ptb, st);
    rc = rc ? rc : dalloc(&B->d_pring, max_ring);
    rc = rc ? rc : dalloc(&B->d_ptbuf, max_tb);
    if (rc) return rc;
    B->plain_tasks = tasks;
    B->plain_rerun = rerun_of;
    B->plain_ready = true;
  }
  HIPCHK(hipEventRecord(B->ev[0], st));
  // the counters and the error word (d_ctr[10]) start at 0, as init_kernel
  // leaves them for a batch run
  HIPCHK(hipMemsetAsync(B->d_ctr, 0, 16 * sizeof(int32_t), st));
  if (B->n > 0)
    HIPCHK(hipMemcpyAsync(B->d_fo, B->d_fo_init, sizeof(FillOut) * B->n, hipMemcpyDeviceToDevice, st));
  std::vector<int> lp_of(B->n, -1);
  for (size_t q = 0; q < B->longs.size(); ++q) lp_of[B->longs[q].pid] = (int)q;
  for (int k = first; k < first + count; ++k) {
    const HostPair &h = B->hp[k];
    if (h.path == 3) {
      const int q = lp_of[k];
      int rc = launch_long_fwd(ctx, B, B->litem_first[q], B->longs[q].S, true);
      if (!rc) rc = launch_traceback(ctx, B, B->d_pairs, B->d_pptb, B->d_ptbuf, k, 1);
      if (!rc) rc = launch_long_tb(ctx, B, q, 1, true);
      if (rc) return rc;
      continue;
    }
    const PairTB *ptb = B->d_pptb;
    if (h.path) {
      HIPCHK(hipMemsetAsync(B->d_ctr + 6, 0, 2 * sizeof(int32_t), st));
      const bool wave = B->plain_tasks[NGW * k].S && pick_gg(h.la, h.lb, 0.0, true) == GWAVE;
      Task *const tk[3] = {B->d_plain_tasks + NGW * k, B->d_plain_tasks + NGW * k, nullptr};
      const int nt[3] = {wave ? 1 : 0, wave ? 0 : 1, 0};
      int rc = h.path == 1 ? launch_fill<1>(ctx, B, tk, nt, B->d_prow, B->d_pcol, B->d_pring, B->d_ptbuf, B->d_ctr + 6, 1)
                           : launch_fill<2>(ctx, B, tk, nt, B->d_prow, B->d_pcol, B->d_pring, B->d_ptbuf, B->d_ctr + 6, 1);
      if (rc) return rc;
      if (B->plain_rerun[k] >= 0) {
        // The int16 fill may have saturated: as the batched form, rerun the
        // pair in int32 when it flagged so (a host round trip per such pair).
        FillOut fo{};
        HIPCHK(hipMemcpyAsync(&fo, B->d_fo + k, sizeof(FillOut), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (fo.flags & TA_F_RERUN) {
          HIPCHK(hipMemsetAsync(B->d_ctr + 6, 0, 2 * sizeof(int32_t), st));
          Task *const rt[3] = {B->d_plain_rtasks + B->plain_rerun[k], B->d_plain_rtasks + B->plain_rerun[k], nullptr};
          rc = launch_fill<1>(ctx, B, rt, nt, B->d_prow, B->d_pcol, B->d_pring, B->d_ptbuf, B->d_ctr + 6, 1);
          if (rc) return rc;
          ptb = B->d_pptb_r;
        }
      }
    }
    int rc = launch_traceback(ctx, B, B->d_pairs, ptb, B->d_ptbuf, k, 1);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(B->ev[1], st));
  HIPCHK(hipEventRecord(B->ev[2], st));
  HIPCHK(hipEventRecord(B->ev[3], st));
  return TA_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: candidate launch shapes timed on the caller's pairs, the fastest
// kept per kind of batch on the context.
int ta_tune_bucket(const ta_pair *pairs, int32_t n) {
  if (!pairs || n <= 0) return 0;
  // The median (element n / 2 in sorted order) is below a limit exactly when
  // more than n / 2 lengths are: counts against the three limits, no sort.
  int32_t c[3] = {0, 0, 0};
  for (int32_t k = 0; k < n; ++k) {
    const int32_t v = std::max(pairs[k].la, pairs[k].lb);
    c[0] += v < 256;
    c[1] += v < 1024;
    c[2] += v < 4096;
  }
  const int32_t h = n / 2;
  return c[0] > h ? 0 : (c[1] > h ? 1 : (c[2] > h ? 2 : 3));
}

int ta_tune_set(ta_ctx *ctx, const ta_params *p, int32_t bucket, const ta_tuning *t) {
  if (!ctx || !p || !t || bucket < 0 || bucket > 3 || t->waves_per_cu < 0 || t->waves_per_cu > 32) return TA_E_INVALID;
  if (t->lanes != TA_LANES_AUTO && t->lanes != TA_LANES_I16 && t->lanes != TA_LANES_I32) return TA_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  ctx->tune[tune_key(p, bucket)] = *t;
  return TA_OK;
}

int ta_tune_get(ta_ctx *ctx, const ta_params *p, int32_t bucket, ta_tuning *t) {
  if (!ctx || !p || !t || bucket < 0 || bucket > 3) return TA_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  auto it = ctx->tune.find(tune_key(p, bucket));
  if (it == ctx->tune.end()) return TA_E_INVALID;
  *t = it->second;
  return TA_OK;
}

int ta_autotune(ta_ctx *ctx, const ta_params *p, const ta_pair *pairs, int32_t n, int32_t repeats, ta_tuning *best,
                double *best_ms) {
  if (!ctx || !p || !pairs || n <= 0) return TA_E_INVALID;
  repeats = std::max(1, std::min(repeats, 50));
  std::vector<int32_t> lanes_c;
  if (p->lanes == TA_LANES_AUTO) lanes_c = {TA_LANES_I16, TA_LANES_I32};
  else lanes_c = {p->lanes};
  const int Kp = make_alphabet(p->alphabet, p->match, p->mismatch).K + 1;
  ta_tuning win{0, TA_LANES_AUTO, 0};
  double win_ms = 1e300;
  for (int32_t lanes : lanes_c) {
    const int P = lanes == TA_LANES_I32 ? 1 : 2;
    int last_eff = -1;
    bool walk_done[2] = {false, false};  // walking fills ignore the wave count: measure one
    for (int waves : {4, 8, 12}) {
      const int eff = fill_waves_per_cu(P, Kp, P == 2 && !(p->mode == TA_LOCAL || p->traceback), waves);
      if (eff == last_eff) continue;
      last_eff = eff;
      for (int wt = 0; wt < 2; ++wt) {
        const ta_tuning cand{waves, lanes, wt};
        ta_batch *B = nullptr;
        int rc = batch_create_on(ctx, p, pairs, n, ctx->stream, &B, &cand);
        if (rc) return rc;
        if (B->walk && walk_done[wt]) {
          ta_batch_destroy(B);
          continue;
        }
        if (B->walk) walk_done[wt] = true;
        double ms = 1e300;
        for (int r = 0; r <= repeats && !rc; ++r) {  // first run warms up
          rc = ta_batch_run(ctx, B);
          if (!rc) rc = ta_batch_wait(ctx, B);
          if (!rc && r > 0) ms = std::min(ms, B->ms[3]);
        }
        ta_batch_destroy(B);
        if (rc) return rc;
        if (ms < win_ms) { win_ms = ms; win = cand; }
      }
    }
  }
  {
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    ctx->tune[tune_key(p, ta_tune_bucket(pairs, n))] = win;
  }
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return TA_OK;
}
