// tweakalign-valu-peak: measured VALU throughput of the instruction kinds the
// DP fill is built from, on every CU of the device, to set beside the
// theoretical int32 peak (CUs x 64 lanes x clock) the roofline uses.
//
// Each kernel runs CHAINS independent dependency chains per lane through an
// unrolled loop of one instruction kind, so the SIMD is issue-bound, not
// latency-bound.  One wave per SIMD (4 per CU) and two per SIMD are timed.
// Output: one JSON line, lane-operations per second per kind.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

namespace {

constexpr int CHAINS = 8;
constexpr int ITERS = 4096;
constexpr int UNROLL = 16;

typedef short s2v __attribute__((ext_vector_type(2)));

// v_max_i32 chains: x = max(x, y_k) with y_k changing every step (v_add).
__global__ __launch_bounds__(64) void k_i32(const int *in, int *out) {
  int x[CHAINS], y[CHAINS];
  for (int c = 0; c < CHAINS; ++c) { x[c] = in[threadIdx.x + c]; y[c] = in[threadIdx.x + 64 + c]; }
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) {
        y[c] = y[c] + x[c];        // v_add_u32
        x[c] = max(x[c], y[c]);    // v_max_i32
      }
  }
  int r = 0;
  for (int c = 0; c < CHAINS; ++c) r ^= x[c] ^ y[c];
  out[blockIdx.x * 64 + threadIdx.x] = r;
}

// Packed int16: the same pattern with saturating v_pk_add_i16 / v_pk_max_i16.
__global__ __launch_bounds__(64) void k_pk16(const int *in, int *out) {
  s2v x[CHAINS], y[CHAINS];
  for (int c = 0; c < CHAINS; ++c) {
    x[c] = __builtin_bit_cast(s2v, in[threadIdx.x + c]);
    y[c] = __builtin_bit_cast(s2v, in[threadIdx.x + 64 + c]);
  }
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) {
        y[c] = __builtin_elementwise_add_sat(y[c], x[c]);
        x[c] = __builtin_elementwise_max(x[c], y[c]);
      }
  }
  int r = 0;
  for (int c = 0; c < CHAINS; ++c) r ^= __builtin_bit_cast(int, x[c]) ^ __builtin_bit_cast(int, y[c]);
  out[blockIdx.x * 64 + threadIdx.x] = r;
}

// v_perm_b32 + v_add_u32.
__global__ __launch_bounds__(64) void k_perm(const int *in, int *out) {
  uint32_t x[CHAINS], y[CHAINS];
  for (int c = 0; c < CHAINS; ++c) { x[c] = in[threadIdx.x + c]; y[c] = in[threadIdx.x + 64 + c]; }
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) {
        y[c] = y[c] + x[c];
        x[c] = __builtin_amdgcn_perm(y[c], x[c], 0x0A050801u);
      }
  }
  uint32_t r = 0;
  for (int c = 0; c < CHAINS; ++c) r ^= x[c] ^ y[c];
  out[blockIdx.x * 64 + threadIdx.x] = (int)r;
}

// DPP row_shr:1 moves + v_add_u32.
__global__ __launch_bounds__(64) void k_dpp(const int *in, int *out) {
  int x[CHAINS], y[CHAINS];
  for (int c = 0; c < CHAINS; ++c) { x[c] = in[threadIdx.x + c]; y[c] = in[threadIdx.x + 64 + c]; }
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) {
        y[c] = y[c] + x[c];
        x[c] = __builtin_amdgcn_update_dpp(x[c], y[c], 0x111, 0xf, 0xf, false);
      }
  }
  int r = 0;
  for (int c = 0; c < CHAINS; ++c) r ^= x[c] ^ y[c];
  out[blockIdx.x * 64 + threadIdx.x] = r;
}

double time_kernel(void (*k)(const int *, int *), int blocks, const int *in, int *out) {
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  hipLaunchKernelGGL(k, dim3(blocks), dim3(64), 0, 0, in, out);  // warm-up
  (void)hipEventRecord(a, 0);
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k, dim3(blocks), dim3(64), 0, 0, in, out);
  (void)hipEventRecord(b, 0);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / 3.0;
}

}  // namespace

int main() {
  int cus = 0, clk_khz = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) {
    fprintf(stderr, "tweakalign-valu-peak: no usable GPU\n");
    return 1;
  }
  (void)hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, 0);
  int *in = nullptr, *out = nullptr;
  if (hipMalloc(&in, 4096) != hipSuccess || hipMalloc(&out, (size_t)cus * 8 * 64 * sizeof(int)) != hipSuccess) {
    fprintf(stderr, "tweakalign-valu-peak: allocation failed\n");
    return 1;
  }
  (void)hipMemset(in, 1, 4096);
  // VALU instructions per lane per launch (2 per chain step).
  const double instr = 2.0 * CHAINS * UNROLL * (double)ITERS;
  struct K { const char *name; void (*k)(const int *, int *); };
  const K ks[] = {{"int32_add_max", k_i32}, {"pk_i16_add_max", k_pk16}, {"perm_add", k_perm}, {"dpp_add", k_dpp}};
  std::string js = "{\"cus\": " + std::to_string(cus) + ", \"clock_mhz\": " + std::to_string(clk_khz / 1000) +
                   ", \"peak_int32_gops_theoretical\": " + std::to_string(cus * 64.0 * clk_khz / 1e6);
  for (int waves_per_simd = 1; waves_per_simd <= 2; ++waves_per_simd) {
    const int blocks = cus * 4 * waves_per_simd;
    for (const K &k : ks) {
      const double ms = time_kernel(k.k, blocks, in, out);
      const double gops = instr * 64.0 * blocks / (ms * 1e-3) / 1e9;  // lane-instructions per second
      char buf[160];
      snprintf(buf, sizeof(buf), ", \"%s_w%d_ginstr_lane_s\": %.1f", k.name, waves_per_simd, gops);
      js += buf;
    }
  }
  js += "}";
  printf("%s\n", js.c_str());
  (void)hipFree(in);
  (void)hipFree(out);
  return 0;
}
