// scmf_bin.h — nearest-grid-point binning, exactly as the reference computes
// it, with a cheap path for the common case.
//
// Reference (coord_to_index_unified_high, device_function.cu:313-360), per axis:
//   r >= 0:  k = (int)(r/L);      i = (int)((r - L*k)/d) % n
//   r <  0:  k = (int)(-r/L) + 1; i = (int)((k*L + r)/d) % n
// with d = L/n rounded to double.  That costs two fp64 divisions and an integer
// modulo per axis.
//
// Fast path: with m = floor(r/L) and t = r - m L (exact real arithmetic), both
// reference branches return floor(t/d) mod n whenever r/L and t/d lie farther
// from an integer than the rounding errors of either computation.  The fast
// path evaluates q1 = r * (1/L), t = fma(-m, L, r), q2 = t * (1/d), checks both
// fractional parts against a bound that covers the reference's error and its
// own (below), and otherwise falls back to the verbatim computation.
//
// Error bound.  Let u = 2^-53.  The reference's k is right when
// dist(r/L, Z) > u |r/L|; its t carries |t_ref - t| <= 2u (L|m| + L) and its
// quotient |q_ref - t/d| <= 2u n (|m| + 2) (n ~ L/d).  The fast path's own
// errors are no larger.  Requiring dist(q1, Z) > 2^-49 (|q1| + 1) and
// dist(q2, Z) > 2^-48 n (|m| + 2), with |q1| < 2^30, leaves a margin above both.
//
// In-box path (fold).  For 0 <= r < L (1 - 2^-50) the reference's quotient
// r/L rounds below 1, so k = 0, r - L*k = r exactly, and the cell is
// trunc(fl(r/d)) mod n.  q = r * (1/d) differs from r/d by at most
// 2^-52 q and fl(r/d) by at most 2^-53 q, with q < n + 1; when q lies
// farther than 2^-50 (n + 2) from an integer both truncate to the same
// cell (and q < n, so the mod is the identity).  Everything else takes
// fold_fast.
//
// Power-of-two cells (fold_pow2).  When d = 2^-k exactly, r/d = r 2^k has no
// rounding in fp64 nor, for fp32 r, in fp32; so for 0 <= r < L (1 - 2^-50)
// the reference's cell is trunc(r 2^k) whenever that is below n.  The test
// returns false otherwise and the caller takes fold.
//
// Power-of-two cells in a power-of-two box (L = 2^j, n = 2^(j+k)): r/L, the
// integer k, L k and the division by d are exact, so for r >= 0 the
// reference's cell is floor(r/d) mod n = floor(r 2^k) & (n - 1) (with
// r - L k exact: a multiple of r's last bit below 2^j).  For r < 0 only the
// sum s = k L + r, 0 < s <= L, can round, by at most 2^(j-53), and rounding
// never crosses a multiple of d from above (those are representable); so the
// cell is again floor(r 2^k) & (n - 1) unless r/d lies within n 2^-53 below
// an integer.  fp32 r: s is exact when |r| >= 2^(j-29) (s is a multiple of
// r's last bit, 2^(e-23), and below 2^(j+1)).  fp64 r (an old + d sum): the
// fractional part f of r/d must stay below 1 - n 2^-52.  Both need |r/d| <
// 2^30, inside the reference's own integer range.  Beads outside the box
// (unwrapped coordinates) then take this path instead of fold.
#pragma once
#include <stdint.h>

#include "scmf_prims.h"

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace scmf {

struct Axis {
  double L, d, invL, invd, guard2;  // guard2 = 2^-48 * n
  double lhi, gin, gin1;            // L (1 - 2^-50); 2^-50 (n + 2); 1 - gin
  float invd_f, lhi_f;              // 1/d and lhi rounded to fp32 (used when pow2)
  int n;
  int pow2;                         // d is an exact power of two
  int wrap2;                        // and L = n d with n a power of two
  float rneg_f;                     // -2^(j-29): fp32 r <= this bins exactly (wrap2)
  double gw1;                       // 1 - n 2^-52: bound on frac(r/d) for fp64 r < 0 (wrap2)
};

// The reference's computation, verbatim.  Coordinates it cannot bin
// (non-finite, or beyond 2^31 box lengths, where its int conversions
// overflow) land in cell 0.
SCMF_HD int fold_exact(double r, double L, double d, int n) {
  int i = 0;
  if (r >= 0) {
    const double q = r / L;
    if (!(q < 2147483647.0)) return 0;
    const int k = (int)q;
    i = (int)((r - L * k) / d);
    i = i % n;
  } else if (r < 0) {
    const double q = -r / L;
    if (!(q < 2147483646.0)) return 0;
    const int k = (int)q + 1;
    i = (int)((k * L + r) / d);
    i = i % n;
  }
  if ((unsigned)i >= (unsigned)n) i = 0;
  return i;
}

// Same result as fold_exact for every input.
SCMF_HD int fold_fast(double r, const Axis a) {
  const double q1 = r * a.invL;
  const double m = __builtin_floor(q1);
  const double f1 = q1 - m;  // exact (Sterbenz) for |q1| < 2^52
  const double g1 = 0x1p-49 * (__builtin_fabs(q1) + 1.0);
  if (__builtin_fabs(q1) < 1073741824.0 && f1 > g1 && f1 < 1.0 - g1) {
    const double t = __builtin_fma(-m, a.L, r);
    const double q2 = t * a.invd;
    const double c = __builtin_floor(q2);
    const double f2 = q2 - c;
    const double g2 = a.guard2 * (__builtin_fabs(m) + 2.0);
    if (f2 > g2 && f2 < 1.0 - g2) {
      int i = (int)c;
      if (i >= a.n) i -= a.n;
      if ((unsigned)i < (unsigned)a.n) return i;
    }
  }
  return fold_exact(r, a.L, a.d, a.n);
}

// Same result as fold_exact for every input; the cheap case first.
SCMF_HD int fold(double r, const Axis a) {
  if (r >= 0.0 && r < a.lhi) {
    const double q = r * a.invd;
    const double c = __builtin_floor(q);
    const double f = q - c;
    if (f > a.gin && f < a.gin1) {
      const int i = (int)c;
      if ((unsigned)i < (unsigned)a.n) return i;
    }
  }
  return fold_fast(r, a);
}

// Power-of-two cells, fp32 coordinate: *i = the reference cell when true.
// (Every fp32 r < (float)lhi is below lhi itself.)
SCMF_HD bool fold_pow2(float r, const Axis a, int* i) {
  if (a.wrap2) {
    const float q = __builtin_floorf(r * a.invd_f);
    const bool ok = (r >= 0.0f || r <= a.rneg_f) && __builtin_fabsf(q) < 1073741824.0f;
    *i = (int)(ok ? q : 0.0f) & (a.n - 1);
    return ok;
  }
  const bool in = r >= 0.0f && r < a.lhi_f;
  const int c = (int)((in ? r : 0.0f) * a.invd_f);
  *i = c;
  return in && c < a.n;
}

// Power-of-two cells, fp64 coordinate.
SCMF_HD bool fold_pow2(double r, const Axis a, int* i) {
  if (a.wrap2) {
    const double q = r * a.invd;
    const double c = __builtin_floor(q);
    const bool ok = (r >= 0.0 || q - c < a.gw1) && __builtin_fabs(c) < 1073741824.0;
    *i = (int)(ok ? c : 0.0) & (a.n - 1);
    return ok;
  }
  const bool in = r >= 0.0 && r < a.lhi;
  const int c = (int)((in ? r : 0.0) * a.invd);
  *i = c;
  return in && c < a.n;
}

SCMF_HD Axis make_axis(double L, int n) {
  Axis a;
  a.L = L;
  a.n = n;
  a.d = L / n;
  a.invL = 1.0 / L;
  a.invd = 1.0 / a.d;
  a.guard2 = 0x1p-48 * (double)n;
  a.lhi = L * (1.0 - 0x1p-50);
  a.gin = 0x1p-50 * ((double)n + 2.0);
  a.gin1 = 1.0 - a.gin;
  int e = 0;
  const double m = __builtin_frexp(a.d, &e);
  a.pow2 = (m == 0.5 && e > -120 && e < 120) ? 1 : 0;
  a.invd_f = (float)a.invd;
  a.lhi_f = (float)a.lhi;
  a.wrap2 = a.pow2 && n > 0 && (n & (n - 1)) == 0 && a.d * n == L ? 1 : 0;
  a.rneg_f = a.wrap2 ? -(float)(L * 0x1p-29) : 0.0f;
  a.gw1 = 1.0 - 0x1p-52 * (double)n;
  return a;
}

}  // namespace scmf
