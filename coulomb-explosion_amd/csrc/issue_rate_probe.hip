// VALU issue-rate probe for gfx950: SIMD cycles per wave64 instruction for the
// instruction classes of the integrator's step, at 1, 2 and 8 resident waves
// per SIMD.  Each lane runs 8 independent chains of one operation for 4096
// iterations (the chains give a single wave enough independent work); the
// time comes from HIP events over 5 launches of a grid of 1, 2 or 8 blocks of
// 256 threads per CU.  Cycles = time x the device clock x SIMDs / wave
// instructions; the loop's own counter and branch (~1 instruction per 8) are
// included.  Build: make -C coulomb-explosion_amd issue_rate_probe
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kChains = 8;
constexpr int kIters = 4096;

enum Op { kFma, kMul, kSub, kCmpSel, kRsq, kRcp, kExp, kFmaRsq };

template <int OP>
__global__ void __launch_bounds__(256) chains(float* out, float seed) {
  float a[kChains], b[kChains], d[kChains];
  for (int c = 0; c < kChains; ++c) {
    a[c] = seed + 0.001f * static_cast<float>(threadIdx.x + c);
    b[c] = 0.999f + 1e-6f * static_cast<float>(c * threadIdx.x);
    d[c] = 1e-4f * static_cast<float>(c + 1) + 1e-7f * static_cast<float>(threadIdx.x);
  }
  for (int i = 0; i < kIters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      if constexpr (OP == kFma) a[c] = __builtin_fmaf(a[c], b[c], d[c]);
      if constexpr (OP == kMul) a[c] = a[c] * b[c];
      if constexpr (OP == kSub) a[c] = a[c] - d[c];
      if constexpr (OP == kCmpSel) a[c] = (a[c] < b[c]) ? d[c] : a[c];
      if constexpr (OP == kRsq) a[c] = __builtin_amdgcn_rsqf(a[c]);
      if constexpr (OP == kRcp) a[c] = __builtin_amdgcn_rcpf(a[c]);
      if constexpr (OP == kExp) a[c] = __builtin_amdgcn_exp2f(-a[c]);
      if constexpr (OP == kFmaRsq) {
        a[c] = __builtin_fmaf(a[c], b[c], d[c]);
        b[c] = __builtin_amdgcn_rsqf(b[c]);
      }
    }
  }
  float s = 0.0f;
  for (int c = 0; c < kChains; ++c) s += a[c] + b[c] + d[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

template <int OP>
void measure(const char* name, int instr_per_chain_step, float* out, int cus, int simds_per_cu,
             double clock_hz, hipEvent_t e0, hipEvent_t e1) {
  std::printf("%-24s", name);
  for (int waves : {1, 2, 8}) {
    const int blocks = cus * waves;  // 256 threads = one wave per SIMD per block
    hipLaunchKernelGGL(chains<OP>, dim3(blocks), dim3(256), 0, nullptr, out, 0.5f);
    check(hipGetLastError(), "launch");
    check(hipEventRecord(e0, nullptr), "event");
    for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(chains<OP>, dim3(blocks), dim3(256), 0, nullptr, out, 0.5f);
    check(hipEventRecord(e1, nullptr), "event");
    check(hipEventSynchronize(e1), "sync");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    const double wave_instr = 5.0 * blocks * (256 / 64) * double(kChains) * kIters * instr_per_chain_step;
    const double cycles = ms * 1e-3 * clock_hz * double(cus) * simds_per_cu / wave_instr;
    std::printf("  %6.2f", cycles);
  }
  std::printf("\n");
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
  int clock_khz = 0;
  check(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, 0), "clock");
  const int cus = prop.multiProcessorCount;
  float* out = nullptr;
  check(hipMalloc(&out, sizeof(float) * size_t(cus) * 8 * 256), "hipMalloc");
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  std::printf("%s, %d CUs, clock %.0f MHz; SIMD cycles per wave64 instruction\n", prop.gcnArchName,
              cus, clock_khz / 1e3);
  std::printf("%-24s  %6s  %6s  %6s\n", "waves per SIMD:", "1", "2", "8");
  const double hz = clock_khz * 1e3;
  measure<kFma>("v_fma_f32", 1, out, cus, 4, hz, e0, e1);
  measure<kMul>("v_mul_f32", 1, out, cus, 4, hz, e0, e1);
  measure<kSub>("v_sub_f32", 1, out, cus, 4, hz, e0, e1);
  measure<kCmpSel>("v_cmp + v_cndmask", 2, out, cus, 4, hz, e0, e1);
  measure<kRsq>("v_rsq_f32", 1, out, cus, 4, hz, e0, e1);
  measure<kRcp>("v_rcp_f32", 1, out, cus, 4, hz, e0, e1);
  measure<kExp>("v_exp_f32", 1, out, cus, 4, hz, e0, e1);
  measure<kFmaRsq>("v_fma + v_rsq (mixed)", 2, out, cus, 4, hz, e0, e1);
  check(hipFree(out), "hipFree");
  return 0;
}
