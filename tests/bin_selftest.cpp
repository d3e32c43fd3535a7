// Harness for tests/test_bin_fast.py: compares the engine's fast NGP binning
// (new_arch_amd/csrc/scmf_bin.h, fold and fold_fast) with the reference
// computation it must reproduce (fold_exact) on random and adversarial
// coordinates.
// Usage: bin_selftest L n count seed  -> prints "<checked> <mismatches> <slow>"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>

#include "../new_arch_amd/csrc/scmf_bin.h"

int main(int argc, char** argv) {
  const double L = atof(argv[1]);
  const int n = atoi(argv[2]);
  const long long count = atoll(argv[3]);
  std::mt19937_64 rng(strtoull(argv[4], nullptr, 10));
  const scmf::Axis a = scmf::make_axis(L, n);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long long bad = 0, checked = 0;
  auto check = [&](double r) {
    ++checked;
    const int f = scmf::fold_fast(r, a);
    const int g = scmf::fold(r, a);
    const int e = scmf::fold_exact(r, a.L, a.d, a.n);
    int pd = e, pf = e;
    if (a.pow2) {  // power-of-two path: when it claims a cell, the cell is right
      int c;
      if (scmf::fold_pow2(r, a, &c)) pd = c;
      const float rf = (float)r;
      if (scmf::fold_pow2(rf, a, &c) && c != scmf::fold_exact((double)rf, a.L, a.d, a.n)) pf = -1;
    }
    if (f != e || g != e || pd != e || pf != e) {
      if (bad < 10)
        fprintf(stderr, "mismatch r=%.17g fast=%d fold=%d pow2=%d pow2f=%d exact=%d\n", r, f, g,
                pd, pf, e);
      ++bad;
    }
  };
  for (long long i = 0; i < count; ++i) {
    const int kind = (int)(i % 14);
    double r;
    if (kind == 0) {
      r = (U(rng) * 2 - 1) * 50 * L;                    // unwrapped coordinates
    } else if (kind == 1) {
      r = (float)((U(rng) * 2 - 1) * 8 * L);           // fp32 positions
    } else if (kind == 2) {                             // cell edges +- a few ulp
      const long long k = (long long)((U(rng) * 2 - 1) * 20 * n);
      r = k * a.d;
      const int s = (int)(U(rng) * 9) - 4;
      for (int t = 0; t < abs(s); ++t) r = nextafter(r, s > 0 ? INFINITY : -INFINITY);
    } else if (kind == 3) {                             // box edges +- a few ulp
      const long long k = (long long)((U(rng) * 2 - 1) * 40);
      r = k * L;
      const int s = (int)(U(rng) * 9) - 4;
      for (int t = 0; t < abs(s); ++t) r = nextafter(r, s > 0 ? INFINITY : -INFINITY);
    } else if (kind == 4) {                             // fp32 cell edges
      const long long k = (long long)((U(rng) * 2 - 1) * 20 * n);
      float f = (float)(k * a.d);
      const int s = (int)(U(rng) * 5) - 2;
      for (int t = 0; t < abs(s); ++t) f = nextafterf(f, s > 0 ? INFINITY : -INFINITY);
      r = f;
    } else if (kind == 5) {                             // old + displacement sums
      const float x = (float)((U(rng) * 2 - 1) * 4 * L);
      const float dx = (float)((U(rng) * 2 - 1) * 0.3);
      r = (double)x + (double)dx;
    } else if (kind == 6) {                             // far away and huge
      r = (U(rng) * 2 - 1) * pow(10.0, U(rng) * 12) * L;
    } else if (kind == 8) {                             // in-box cell edges +- a few ulp
      const long long k = (long long)(U(rng) * (n + 1));
      r = k * a.d;
      const int s = (int)(U(rng) * 9) - 4;
      for (int t = 0; t < abs(s); ++t) r = nextafter(r, s > 0 ? INFINITY : -INFINITY);
    } else if (kind == 9) {                             // around L and L (1 - 2^-50)
      r = U(rng) < 0.5 ? L : a.lhi;
      const int s = (int)(U(rng) * 41) - 20;
      for (int t = 0; t < abs(s); ++t) r = nextafter(r, s > 0 ? INFINITY : -INFINITY);
      if (U(rng) < 0.5) r = (float)r;
    } else if (kind == 10) {                            // in-box fp32 cell edges
      const long long k = (long long)(U(rng) * (n + 1));
      float f = (float)(k * a.d);
      const int s = (int)(U(rng) * 7) - 3;
      for (int t = 0; t < abs(s); ++t) f = nextafterf(f, s > 0 ? INFINITY : -INFINITY);
      r = f;
    } else if (kind == 11) {                            // negative fp32 cell edge + tiny fp32 step
      const long long k = (long long)(U(rng) * 3 * n) + 1;
      const float x = (float)(-k * a.d);
      const float dx = (float)((U(rng) * 2 - 1) * a.d * ldexp(1.0, -(int)(20 + U(rng) * 45)));
      r = (double)x + (double)dx;
    } else if (kind == 12) {                            // fp32 around -L 2^-29 and tiny negatives
      if (U(rng) < 0.5) {
        float f = (float)(-L * ldexp(1.0, -29));
        const int s = (int)(U(rng) * 9) - 4;
        for (int t = 0; t < abs(s); ++t) f = nextafterf(f, s > 0 ? INFINITY : -INFINITY);
        r = f;
      } else {
        r = (float)(-L * ldexp(U(rng) + 0.5, -(int)(25 + U(rng) * 120)));
      }
    } else if (kind == 13) {                            // fp64 just below a negative cell edge
      const long long k = (long long)(U(rng) * 3 * n);
      r = -k * a.d - a.d * ldexp(U(rng), -(int)(38 + U(rng) * 20));
    } else {                                            // tiny and signed zeros
      const double v[6] = {0.0, -0.0, 1e-300, -1e-300, 5e-324, -5e-324};
      r = v[i % 6] * (U(rng) < 0.5 ? 1 : 1e10);
    }
    check(r);
  }
  const double special[5] = {INFINITY, -INFINITY, NAN, 1e300, -1e300};
  for (double r : special) check(r);
  if (argc > 5 && atoi(argv[5]) != a.pow2) {
    fprintf(stderr, "pow2 flag %d, expected %s\n", a.pow2, argv[5]);
    ++bad;
  }
  printf("%lld %lld\n", checked, bad);
  return bad != 0;
}
