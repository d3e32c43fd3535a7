// scmf_calib.hip — kernels with a known VALU instruction stream, for reading
// the SQ issue counters of the engine's kernels against a measured rate.
//
// Each kernel runs `iters` iterations of independent dependency chains (the
// loop counter lives in scalar registers).  VALU instructions per loop
// iteration, as hipcc -O3 emits them for gfx950:
//   kind 0: 8 v_fmamk_f32
//   kind 1: 8 v_fma_f64
//   kind 2: 8 v_mad_u64_u32 + 8 v_xor_b32 (the 32x32->64 products of Philox)
//   kind 3: 4 v_fma_f64 + 4 v_lshrrev_b32 + 4 v_xor_b32 (fp64 and integer
//           streams that may issue together)
// Timed with HIP events at full occupancy they give the cycles a wave64 VALU
// instruction of each kind holds a SIMD; under rocprofv3 they give the ratio
// of SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU and SQ_ACTIVE_INST_VALU2 (dual issue)
// to that rate.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int CAL_BLOCK = 256;

template <int KIND>
__global__ void __launch_bounds__(CAL_BLOCK) k_valu(float* out, int iters, float seed) {
  const int t = blockIdx.x * CAL_BLOCK + threadIdx.x;
  if constexpr (KIND == 0) {
    float a[8];
    for (int k = 0; k < 8; ++k) a[k] = seed + (float)(t + k);
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = __builtin_fmaf(a[k], 0.99999994f, 1.0e-7f);
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) s += a[k];
    out[t] = s;
  } else if constexpr (KIND == 1) {
    double a[8];
    for (int k = 0; k < 8; ++k) a[k] = (double)seed + (double)(t + k);
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = __builtin_fma(a[k], 0.9999999999999999, 1.0e-17);
    double s = 0.0;
    for (int k = 0; k < 8; ++k) s += a[k];
    out[t] = (float)s;
  } else if constexpr (KIND == 2) {
    uint32_t a[8];
    for (int k = 0; k < 8; ++k) a[k] = (uint32_t)t * 2654435761u + (uint32_t)k + (uint32_t)seed;
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint64_t p = (uint64_t)a[k] * 0xD2511F53u + (uint64_t)(k + 1);
        a[k] = (uint32_t)p ^ (uint32_t)(p >> 32);
      }
    uint32_t s = 0;
    for (int k = 0; k < 8; ++k) s ^= a[k];
    out[t] = (float)s;
  } else {
    double a[4];
    uint32_t b[4];
    for (int k = 0; k < 4; ++k) {
      a[k] = (double)seed + (double)(t + k);
      b[k] = (uint32_t)t + (uint32_t)k;
    }
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        a[k] = __builtin_fma(a[k], 0.9999999999999999, 1.0e-17);
        b[k] = b[k] ^ (b[k] >> 3);
      }
    double s = 0.0;
    for (int k = 0; k < 4; ++k) s += a[k] + (double)b[k];
    out[t] = (float)s;
  }
}

const void* kernel_of(int kind) {
  switch (kind) {
    case 0: return (const void*)k_valu<0>;
    case 1: return (const void*)k_valu<1>;
    case 2: return (const void*)k_valu<2>;
    default: return (const void*)k_valu<3>;
  }
}

}  // namespace

extern "C" {

// Runs kernel `kind` on `blocks` workgroups of 256 lanes for `iters`
// iterations, `reps` times; *ms = mean device time per launch (HIP events).
// Returns 0, or a HIP error code.
int scmf_calib_valu(int kind, int blocks, int iters, int reps, float* ms) {
  if (kind < 0 || kind > 3 || blocks <= 0 || iters <= 0 || reps <= 0 || !ms) return -1;
  float* out = nullptr;
  hipError_t e = hipMalloc(&out, sizeof(float) * (size_t)blocks * CAL_BLOCK);
  if (e != hipSuccess) return (int)e;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  float seed = 1.0f;
  void* args[] = {(void*)&out, (void*)&iters, (void*)&seed};
  e = hipLaunchKernel(kernel_of(kind), dim3(blocks), dim3(CAL_BLOCK), args, 0, nullptr);  // warm-up
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipEventRecord(a, nullptr);
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    e = hipLaunchKernel(kernel_of(kind), dim3(blocks), dim3(CAL_BLOCK), args, 0, nullptr);
  if (e == hipSuccess) e = hipEventRecord(b, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  *ms = t / (float)reps;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  return (int)e;
}

}  // extern "C"
