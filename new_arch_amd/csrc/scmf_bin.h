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
#pragma once
#include <stdint.h>

#include "scmf_prims.h"

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace scmf {

struct Axis {
  double L, d, invL, invd, guard2;  // guard2 = 2^-48 * n
  int n;
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

SCMF_HD Axis make_axis(double L, int n) {
  Axis a;
  a.L = L;
  a.n = n;
  a.d = L / n;
  a.invL = 1.0 / L;
  a.invd = 1.0 / a.d;
  a.guard2 = 0x1p-48 * (double)n;
  return a;
}

}  // namespace scmf
