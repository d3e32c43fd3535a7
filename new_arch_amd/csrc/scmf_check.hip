// scmf_check.hip — the device builds of the shared arithmetic primitives
// (scmf_prims.h) over caller-given inputs, so tests can compare them bit for
// bit with the host builds the oracle uses.  The device build of some of them
// differs in code (scalar-register fma addends in exp, a fixed-range
// correctly rounded sqrt in the Box-Muller normals, three-input XORs in the
// Philox rounds) but must not differ in results.  Not part of the engine:
// libscmf_check.so is test support.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scmf_prims.h"

namespace {

constexpr int CHK_BLOCK = 256;

__global__ void __launch_bounds__(CHK_BLOCK) k_normal3f(const scmf::u32x4* __restrict__ w, int64_t n,
                                                       float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * CHK_BLOCK + threadIdx.x;
  if (i < n) scmf::normal3f(w[i], &out[3 * i], &out[3 * i + 1], &out[3 * i + 2]);
}

__global__ void __launch_bounds__(CHK_BLOCK) k_exp(const double* __restrict__ x, int64_t n,
                                                  double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * CHK_BLOCK + threadIdx.x;
  if (i < n) y[i] = scmf::exp_cw(x[i]);
}

// in: counter words c0..c3 and key words k0, k1 of each block
__global__ void __launch_bounds__(CHK_BLOCK) k_philox(const uint32_t* __restrict__ in, int64_t n,
                                                     uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * CHK_BLOCK + threadIdx.x;
  if (i < n) {
    const uint32_t* c = in + 6 * i;
    const scmf::u32x4 r = scmf::philox4x32_10(c[0], c[1], c[2], c[3], c[4], c[5]);
    for (int j = 0; j < 4; ++j) out[4 * i + j] = r.v[j];
  }
}

template <class In, class Out, class K>
int run(K kern, const In* in_host, size_t in_elems, int64_t n, Out* out_host, size_t out_elems) {
  if (n <= 0) return 0;
  In* din = nullptr;
  Out* dout = nullptr;
  hipError_t e = hipMalloc(&din, sizeof(In) * in_elems);
  if (e == hipSuccess) e = hipMalloc(&dout, sizeof(Out) * out_elems);
  if (e == hipSuccess) e = hipMemcpy(din, in_host, sizeof(In) * in_elems, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const unsigned blocks = (unsigned)((n + CHK_BLOCK - 1) / CHK_BLOCK);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(CHK_BLOCK), 0, nullptr, din, n, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out_host, dout, sizeof(Out) * out_elems, hipMemcpyDeviceToHost);
  if (din) (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  return (int)e;
}

}  // namespace

extern "C" {

// out[3 i .. 3 i + 2] = normal3f of the Philox block w[4 i .. 4 i + 3].
// Returns 0 or a HIP error code.
int scmf_check_normal3f(const uint32_t* w, int64_t n, float* out) {
  if (n < 0 || (n > 0 && (!w || !out))) return -1;
  return run(k_normal3f, (const scmf::u32x4*)w, (size_t)n, n, out, 3 * (size_t)n);
}

// out[4 i .. 4 i + 3] = Philox4x32-10 of counter in[6 i .. 6 i + 3] and key
// in[6 i + 4], in[6 i + 5].  Returns 0 or a HIP error code.
int scmf_check_philox(const uint32_t* in, int64_t n, uint32_t* out) {
  if (n < 0 || (n > 0 && (!in || !out))) return -1;
  return run(k_philox, in, 6 * (size_t)n, n, out, 4 * (size_t)n);
}

// y[i] = exp_cw(x[i]).  Returns 0 or a HIP error code.
int scmf_check_exp(const double* x, int64_t n, double* y) {
  if (n < 0 || (n > 0 && (!x || !y))) return -1;
  return run(k_exp, x, (size_t)n, n, y, (size_t)n);
}

}  // extern "C"
