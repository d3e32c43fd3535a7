// scmf_prims.h — arithmetic primitives shared, bit for bit, by the HIP kernels
// and the CPU oracle.
//
// What lives here is the *definition* of the injected random streams and of the
// special functions the Monte Carlo path needs:
//   * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As
//     Easy as 1, 2, 3", SC'11), keyed by the run seed, counter-addressed by
//     (chain, sweep, move, slot, system);
//   * uniform transforms of its 32-bit outputs;
//   * Box–Muller normals (Box & Muller 1958) built from our own log and sincos;
//   * our own exp (Cody & Waite 1980 argument reduction + Taylor polynomial).
// The reference draws from cuRAND (curand_uniform, curand_normal4,
// curand_uniform_double, curand_normal4_double; device_function.cu:163-183,
// 683, 749) and calls CUDA's exp (device_function.cu:447).  Neither can be
// reproduced here, so both sides consume these streams instead.
//
// Every operation below is a single IEEE-754 operation (+ - * / sqrt fma,
// conversions, floor) so g++ (x86-64 SSE2) and hipcc (gfx950) produce the same
// bits.  Both compilers must build this with FP contraction off; fma is only
// ever requested explicitly.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SCMF_HD __host__ __device__ __forceinline__
#else
#define SCMF_HD static inline
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace scmf {

// ---------------------------------------------------------------- Philox4x32-10
struct u32x4 {
  uint32_t v[4];
};

// a ^ b ^ c: on gfx950 one v_bitop3_b32 (truth table 0x96), else two XORs
SCMF_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

SCMF_HD void philox_round(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                          uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = xor3(hi1, c1, k0);
  const uint32_t n2 = xor3(hi0, c3, k1);
  c0 = n0;
  c1 = lo1;
  c2 = n2;
  c3 = lo0;
}

// Ten rounds with the Weyl key schedule (0x9E3779B9, 0xBB67AE85).
SCMF_HD u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                            uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    philox_round(c0, c1, c2, c3, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  u32x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// Stream slots.  Counter layout: {chain, sweep, (index << 2) | slot, system}.
//   SLOT_PICK, index k: the bead picks of moves 2k (words 0, 1) and 2k + 1
//                       (words 2, 3), 53-bit doubles;
//   SLOT_DISP, index i: move i's displacement from the top 24 bits of its words
//                       (three normals for smart MC, three uniforms for trial
//                       moves) and its accept uniform from the low bytes of
//                       words 0, 1, 2 (bits no displacement uses);
//   SLOT_INIT:          the initial configuration (start point, then one block
//                       per placed bead).
// A move thus costs one Philox block plus half of one for its pick.
enum : uint32_t {
  SLOT_PICK = 0,
  SLOT_DISP = 1,
  SLOT_INIT = 2,
  INIT_SWEEP = 0xFFFFFFFFu  // sweep word reserved for the initial configuration
};

SCMF_HD u32x4 stream_block(uint64_t seed, uint32_t system, uint32_t chain, uint32_t sweep,
                           uint32_t index, uint32_t slot) {
  return philox4x32_10(chain, sweep, (index << 2) | slot, system, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}

// ---------------------------------------------------------------- uniforms
// (0, 1] with 53 random bits: ((hi27 << 26) | lo26) + 1, times 2^-53.  Exact.
SCMF_HD double u01_53(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)(m + 1) * 1.1102230246251565404e-16;  // 2^-53
}
// (0, 1] float with 24 random bits.  Exact.
SCMF_HD float u01_24f(uint32_t a) { return (float)((a >> 8) + 1u) * 5.9604644775390625e-8f; }
// [0, 1) float with 24 random bits.  Exact.
SCMF_HD float u01_24f_closed0(uint32_t a) { return (float)(a >> 8) * 5.9604644775390625e-8f; }
// (0, 1) double from 32 bits: (a + 1/2) 2^-32.  Exact.
SCMF_HD double u01_32_open(uint32_t a) {
  return ((double)a + 0.5) * 2.3283064365386962890625e-10;
}

// The pick uniform of move i from its SLOT_PICK block (index i >> 1).
SCMF_HD double pick_u01(const u32x4& w, uint32_t i) {
  return (i & 1u) ? u01_53(w.v[2], w.v[3]) : u01_53(w.v[0], w.v[1]);
}
// The bead a move with pick uniform u picks on a chain of n beads:
// (int)(u n), the last bead for u = 1 (device_function.cu:749-751).
SCMF_HD int pick_bead(double u, int n) {
  const int b = (int)(u * n);
  return b == n ? n - 1 : b;
}
// pick_bead(pick_u01(w, i), n) with n53 = n 2^-53, in fewer operations: u
// is m 2^-53 for the 53-bit integer m = hi 2^26 + lo + 1 (one exact fma of
// the two exactly converted halves), and m n53 is the real number u n, so
// its single rounding is that of u * n.
SCMF_HD int pick_of(const u32x4& w, uint32_t i, int n, double n53) {
  const uint32_t a = (i & 1u) ? w.v[2] : w.v[0];
  const uint32_t b = (i & 1u) ? w.v[3] : w.v[1];
  const double m = __builtin_fma((double)(a >> 5), 67108864.0, (double)((b >> 6) + 1u));
  const int k = (int)(m * n53);
  return k == n ? n - 1 : k;
}
// The accept uniform of a move, (0, 1] with 24 bits, from the low bytes of
// words 0, 1, 2 of its SLOT_DISP block.  Exact.
SCMF_HD float accept_u01(const u32x4& disp) {
  const uint32_t b = (disp.v[0] & 0xFFu) | ((disp.v[1] & 0xFFu) << 8) | ((disp.v[2] & 0xFFu) << 16);
  return (float)(b + 1u) * 5.9604644775390625e-8f;
}

// ---------------------------------------------------------------- log
// Natural log for finite x > 0 (normal range).  x = 2^e m, m in [sqrt(1/2),
// sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1), summed as the odd Taylor
// series of atanh to s^23 (|s| <= 0.1716, truncation < 1e-19).  The split of
// ln 2 into a 29-bit head and a tail keeps e*LN2_HI exact for |e| < 2^24.
SCMF_HD double log_pos(double x) {
  union {
    double d;
    uint64_t u;
  } bits;
  bits.d = x;
  int e = (int)((bits.u >> 52) & 0x7FF) - 1023;
  bits.u = (bits.u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;  // m in [1, 2)
  double m = bits.d;
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double s = f / (m + 1.0);
  const double z = s * s;
  // atanh(s)/s = sum_k z^k / (2k+1), k = 0..11, Horner from the top.
  double p = 1.0 / 23.0;
  p = __builtin_fma(p, z, 1.0 / 21.0);
  p = __builtin_fma(p, z, 1.0 / 19.0);
  p = __builtin_fma(p, z, 1.0 / 17.0);
  p = __builtin_fma(p, z, 1.0 / 15.0);
  p = __builtin_fma(p, z, 1.0 / 13.0);
  p = __builtin_fma(p, z, 1.0 / 11.0);
  p = __builtin_fma(p, z, 1.0 / 9.0);
  p = __builtin_fma(p, z, 1.0 / 7.0);
  p = __builtin_fma(p, z, 1.0 / 5.0);
  p = __builtin_fma(p, z, 1.0 / 3.0);
  // log m = 2 s + 2 s z p'  where p' is the series without its leading 1.
  const double two_s = 2.0 * s;
  const double logm = __builtin_fma(two_s * z, p, two_s);
  const double LN2_HI = 0x1.62e42fep-1;           // ln 2 with its low 24 bits cleared
  const double LN2_LO = 0x1.f473de6af278fp-30;    // ln 2 - LN2_HI
  const double de = (double)e;
  return __builtin_fma(de, LN2_HI, __builtin_fma(de, LN2_LO, logm));
}

// ---------------------------------------------------------------- sincos(2 pi u)
// For u in [0, 1): quarter-turn reduction is exact because u is a multiple of
// 2^-32; the remaining angle t = 2 pi v, |v| <= 1/8, goes through the Taylor
// series of sin (to t^19) and cos (to t^18).
SCMF_HD void sincos_2pi(double u, double* s_out, double* c_out) {
  const double q = __builtin_floor(u * 4.0 + 0.5);  // nearest quarter turn, 0..4
  const double v = u - q * 0.25;                     // exact, |v| <= 1/8
  const double t = v * 6.283185307179586232;         // 2 pi
  const double t2 = t * t;
  double sp = -1.0 / 121645100408832000.0;           // -1/19!
  sp = __builtin_fma(sp, t2, 1.0 / 355687428096000.0);
  sp = __builtin_fma(sp, t2, -1.0 / 1307674368000.0);
  sp = __builtin_fma(sp, t2, 1.0 / 6227020800.0);
  sp = __builtin_fma(sp, t2, -1.0 / 39916800.0);
  sp = __builtin_fma(sp, t2, 1.0 / 362880.0);
  sp = __builtin_fma(sp, t2, -1.0 / 5040.0);
  sp = __builtin_fma(sp, t2, 1.0 / 120.0);
  sp = __builtin_fma(sp, t2, -1.0 / 6.0);
  const double sn = __builtin_fma(t * t2, sp, t);
  double cp = 1.0 / 6402373705728000.0;              // 1/18!
  cp = __builtin_fma(cp, t2, -1.0 / 20922789888000.0);
  cp = __builtin_fma(cp, t2, 1.0 / 87178291200.0);
  cp = __builtin_fma(cp, t2, -1.0 / 479001600.0);
  cp = __builtin_fma(cp, t2, 1.0 / 3628800.0);
  cp = __builtin_fma(cp, t2, -1.0 / 40320.0);
  cp = __builtin_fma(cp, t2, 1.0 / 720.0);
  cp = __builtin_fma(cp, t2, -1.0 / 24.0);
  cp = __builtin_fma(cp, t2, 0.5);
  const double cs = __builtin_fma(-t2, cp, 1.0);
  const int quad = ((int)q) & 3;
  double s, c;
  if (quad == 0) {
    s = sn;
    c = cs;
  } else if (quad == 1) {
    s = cs;
    c = -sn;
  } else if (quad == 2) {
    s = -sn;
    c = -cs;
  } else {
    s = -cs;
    c = sn;
  }
  *s_out = s;
  *c_out = c;
}

// ---------------------------------------------------------------- normals
// Three standard normals from one Philox block: (w0, w1) -> Box–Muller pair,
// (w2, w3) -> the cosine branch only (the fourth normal is not needed).
SCMF_HD void normal3(const u32x4& w, double* n0, double* n1, double* n2) {
  const double r0 = __builtin_sqrt(-2.0 * log_pos(u01_32_open(w.v[0])));
  double s0, c0;
  sincos_2pi((double)w.v[1] * 2.3283064365386962890625e-10, &s0, &c0);
  const double r1 = __builtin_sqrt(-2.0 * log_pos(u01_32_open(w.v[2])));
  double s1, c1;
  sincos_2pi((double)w.v[3] * 2.3283064365386962890625e-10, &s1, &c1);
  *n0 = r0 * c0;
  *n1 = r0 * s0;
  *n2 = r1 * c1;
}

// ---------------------------------------------------------------- fp32 normals
// Natural log in fp32 for x in (0, 1], with no division: x = 2^e m,
// m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), log m = f + f^2 Q(f) with Q
// the degree-8 polynomial fitted to (log1p(f) - f) / f^2 on that range
// (least squares on Chebyshev nodes, weighted by the effect on log m and
// reweighted towards the largest errors, then rounded to fp32), evaluated by
// Horner's rule with FMAs.  LN2 split so e*LN2F_HI is exact.  Every input of
// the displacement normals, 2^-24 k for k = 1..2^24, is within 0.9 ulp of the
// exact logarithm, and log 1 = 0.
SCMF_HD float logf_pos(float x) {
  union {
    float f;
    uint32_t u;
  } bits;
  bits.f = x;
  int e = (int)((bits.u >> 23) & 0xFF) - 127;
  bits.u = (bits.u & 0x007FFFFFu) | 0x3F800000u;
  float m = bits.f;
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  float q = -0x1.107b3p-4f;
  q = __builtin_fmaf(q, f, 0x1.f7638ap-4f);
  q = __builtin_fmaf(q, f, -0x1.10e41p-3f);
  q = __builtin_fmaf(q, f, 0x1.2474e8p-3f);
  q = __builtin_fmaf(q, f, -0x1.5432ccp-3f);
  q = __builtin_fmaf(q, f, 0x1.998ad2p-3f);
  q = __builtin_fmaf(q, f, -0x1.00038ap-2f);
  q = __builtin_fmaf(q, f, 0x1.55557ep-2f);
  q = __builtin_fmaf(q, f, -0x1.fffffap-2f);
  const float logm = __builtin_fmaf(f * f, q, f);
  const float LN2F_HI = 0x1.62e4p-1f;             // ln 2 with its low 8 bits cleared
  const float LN2F_LO = 0x1.7f7d1cp-20f;          // ln 2 - LN2F_HI
  const float fe = (float)e;
  return __builtin_fmaf(fe, LN2F_HI, __builtin_fmaf(fe, LN2F_LO, logm));
}

// sin and cos of 2 pi k 2^-24 for a 24-bit integer k: the nearest quarter
// turn is split off in integer arithmetic (exact), the rest, |t| <= pi/4,
// goes through Taylor series to t^11 (sin) and t^10 (cos).
SCMF_HD void sincosf_turn24(uint32_t k, float* s_out, float* c_out) {
  const uint32_t q = (k + (1u << 21)) >> 22;                   // 0..4
  const int32_t v = (int32_t)k - (int32_t)(q << 22);           // |v| <= 2^21
  const float t = (float)v * (5.9604644775390625e-8f * 6.28318530717958648f);
  const float t2 = t * t;
  float sp = -1.0f / 39916800.0f;
  sp = __builtin_fmaf(sp, t2, 1.0f / 362880.0f);
  sp = __builtin_fmaf(sp, t2, -1.0f / 5040.0f);
  sp = __builtin_fmaf(sp, t2, 1.0f / 120.0f);
  sp = __builtin_fmaf(sp, t2, -1.0f / 6.0f);
  const float sn = __builtin_fmaf(t * t2, sp, t);
  float cp = -1.0f / 3628800.0f;
  cp = __builtin_fmaf(cp, t2, 1.0f / 40320.0f);
  cp = __builtin_fmaf(cp, t2, -1.0f / 720.0f);
  cp = __builtin_fmaf(cp, t2, 1.0f / 24.0f);
  cp = __builtin_fmaf(cp, t2, -0.5f);
  const float cs = __builtin_fmaf(t2, cp, 1.0f);
  // quarter turn q: odd quadrants swap sin and cos; sin is negated in
  // quadrants 2 and 3, cos in 1 and 2 (sign-bit flips, as negation is), so
  // the selection needs no branch: (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn)
  const uint32_t quad = q & 3u;
  const float a = (quad & 1u) ? cs : sn;
  const float b = (quad & 1u) ? sn : cs;
  const uint32_t sneg = (quad & 2u) << 30, cneg = ((quad + 1u) & 2u) << 30;
  *s_out = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, a) ^ sneg);
  *c_out = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, b) ^ cneg);
}

// Correctly rounded sqrt of x = -2 log(u), u a 24-bit uniform in (0, 1]: x
// is +-0 or at least 1.19e-7, never subnormal, infinite or NaN.  On the GPU:
// v_sqrt_f32 (within 1 ulp) and the two fma residual checks of its
// neighbours that the compiler's IEEE expansion applies, without that
// expansion's scaling of tiny arguments and its special-value check, which
// such an x never needs (+-0 comes back as +-0 through the checks).
SCMF_HD float sqrtf_bm(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float y = __builtin_amdgcn_sqrtf(x);
  const float ym = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, y) - 1u);
  const float yp = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, y) + 1u);
  const float rm = __builtin_fmaf(-ym, y, x);
  const float rp = __builtin_fmaf(-yp, y, x);
  const float r = rm <= 0.0f ? ym : y;
  return rp > 0.0f ? yp : r;
#else
  return __builtin_sqrtf(x);
#endif
}

// Three standard normals in fp32 from one Philox block (Box–Muller on
// 24-bit uniforms): (w0, w1) -> a pair, (w2, w3) -> the cosine branch.
// This is the displacement noise of a smart-MC move (the reference draws it
// with curand_normal4 in fp32, device_function.cu:683).
SCMF_HD void normal3f(const u32x4& w, float* n0, float* n1, float* n2) {
  const float u0 = u01_24f(w.v[0]);  // (0, 1]
  const float r0 = sqrtf_bm(-2.0f * logf_pos(u0));
  float s0, c0;
  sincosf_turn24(w.v[1] >> 8, &s0, &c0);
  const float u2 = u01_24f(w.v[2]);
  const float r1 = sqrtf_bm(-2.0f * logf_pos(u2));
  float s1, c1;
  sincosf_turn24(w.v[3] >> 8, &s1, &c1);
  *n0 = r0 * c0;
  *n1 = r0 * s0;
  *n2 = r1 * c1;
}

// ---------------------------------------------------------------- exp
// fma(p, r, c) for a constant c.  On the GPU the addend is read from a scalar
// register pair (one v_fma_f64); left to itself the compiler keeps every
// fp64 coefficient of a polynomial in a vector register pair, hoisted out of
// the enclosing loop, and copies it before a two-operand fma, which costs a
// move per coefficient and spills in a kernel short of registers.  The
// arithmetic is the same single fused multiply-add.
SCMF_HD double fma_c(double p, double r, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double out;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(out) : "v"(p), "v"(r), "s"(c));
  return out;
#else
  return __builtin_fma(p, r, c);
#endif
}

// exp(x) in double.  k = nearest integer to x/ln2; r = x - k ln2 in two steps
// (Cody & Waite 1980), |r| <= 0.347; exp(r) by its Taylor series to r^14
// (truncation < 2e-18 relative); scale by 2^k in two halves so gradual
// underflow is handled.  Overflow returns +inf, deep underflow returns 0.
SCMF_HD double exp_cw(double x) {
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.2) return 0.0;
  if (!(x == x)) return x;
  const double kd = __builtin_floor(x * 1.4426950408889634074 + 0.5);
  const double LN2_HI = 0x1.62e42fep-1;
  const double LN2_LO = 0x1.f473de6af278fp-30;
  const double r = __builtin_fma(-kd, LN2_LO, __builtin_fma(-kd, LN2_HI, x));
  double p = 1.0 / 87178291200.0;  // 1/14!
  p = fma_c(p, r, 1.0 / 6227020800.0);
  p = fma_c(p, r, 1.0 / 479001600.0);
  p = fma_c(p, r, 1.0 / 39916800.0);
  p = fma_c(p, r, 1.0 / 3628800.0);
  p = fma_c(p, r, 1.0 / 362880.0);
  p = fma_c(p, r, 1.0 / 40320.0);
  p = fma_c(p, r, 1.0 / 5040.0);
  p = fma_c(p, r, 1.0 / 720.0);
  p = fma_c(p, r, 1.0 / 120.0);
  p = fma_c(p, r, 1.0 / 24.0);
  p = fma_c(p, r, 1.0 / 6.0);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  const double er = __builtin_fma(p, r, 1.0);
  const int k = (int)kd;
  const int k1 = k / 2;
  const int k2 = k - k1;
  union {
    double d;
    uint64_t u;
  } s1, s2;
  s1.u = (uint64_t)(k1 + 1023) << 52;
  s2.u = (uint64_t)(k2 + 1023) << 52;
  return (er * s1.d) * s2.d;
}

}  // namespace scmf
