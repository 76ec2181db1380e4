// FP64 host geometry: Gauss-Legendre quadrature and the normalized associated
// Legendre basis (Pmn) with its latitude derivative (Hmn = (1-mu^2) dPmn/dmu).
//
// The reference builds these with scipy (roots_legendre, assoc_legendre_p_all)
// and an exact big-integer factorial scaling (spectral_transform.py:80-146,
// :219-287).  Here they are computed from the mathematics directly:
//   * nodes by Newton iteration on P_J from the Tricomi initial guess, weights
//     2 / ((1-x^2) P_J'(x)^2);
//   * the orthonormal, Condon-Shortley-free basis by the standard stable
//     three-term recurrences for \bar P_n^m (no factorials, so no overflow at
//     n+m = 212), which equals the reference's scaled scipy values with the
//     odd-m sign flip applied.
#include "geometry.h"

#include <algorithm>
#include <cmath>

namespace swx {

bool resolution(int M, int* J, int* I, int* dt_default) {
  int j = 0, i = 0, d = 0;
  switch (M) {
    case 42: j = 64; i = 128; d = 1200; break;
    case 63: j = 96; i = 192; d = 900; break;
    case 106: j = 160; i = 320; d = 600; break;
    default: return false;
  }
  if (J) *J = j;
  if (I) *I = i;
  if (dt_default) *dt_default = d;
  return true;
}

void gauss_legendre(int J, double* x, double* w) {
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < J; ++k) {
    // Root k (descending from near +1).
    double z = std::cos(pi * (k + 0.75) / (J + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int n = 2; n <= J; ++n) {
        double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
        p0 = p1;
        p1 = p2;
      }
      // p1 = P_J(z), p0 = P_{J-1}(z)
      dp = J * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) <= 1e-17 * std::max(1.0, std::fabs(z))) break;
    }
    // Final derivative at the converged root.
    double p0 = 1.0, p1 = z;
    for (int n = 2; n <= J; ++n) {
      double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
      p0 = p1;
      p1 = p2;
    }
    dp = J * (z * p1 - p0) / (z * z - 1.0);
    const int idx = J - 1 - k;  // ascending order
    x[idx] = z;
    w[idx] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// Reference scale sqrt(r) with r = (2n+1)(n-m)! / (2 (n+m)!) rounded to a
// double the way spectral_transform.py:143-145 does it (exact rational ->
// nearest double, so r becomes subnormal or 0 once (n+m)! / (n-m)! exceeds
// ~1e308).  Returns the ratio sqrt(r_double) / sqrt(r_exact) that maps the
// exactly normalized basis onto the reference's basis: 1 where r is a normal
// double, < 1 or 0 where it under/overflowed (only at T106: n + m >= ~170).
double reference_scale_factor(int m, int n) {
  long double r = (2.0L * n + 1.0L) / 2.0L;
  for (int k = n - m + 1; k <= n + m; ++k) r /= (long double)k;
  const double rd = (double)r;
  if (rd >= 2.2250738585072014e-308) return 1.0;
  if (rd == 0.0) return 0.0;
  return (double)(std::sqrt((long double)rd) / std::sqrt(r));
}

void legendre_basis(int M, int N, int J, const double* mus, double* P, double* H, bool reference_scaling) {
  const int NM = M + 1, NN = N + 1;
  std::fill(P, P + (size_t)J * NM * NN, 0.0);
  std::fill(H, H + (size_t)J * NM * NN, 0.0);
  for (int j = 0; j < J; ++j) {
    const double x = mus[j];
    const double s = std::sqrt(1.0 - x * x);
    double* Pj = P + (size_t)j * NM * NN;
    double* Hj = H + (size_t)j * NM * NN;
    double pmm = std::sqrt(0.5);  // \bar P_0^0
    for (int m = 0; m <= M; ++m) {
      if (m > 0) pmm *= std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * s;
      if (m > N) break;
      double* row = Pj + (size_t)m * NN;
      row[m] = pmm;
      if (m + 1 <= N) row[m + 1] = std::sqrt(2.0 * m + 3.0) * x * pmm;
      for (int n = m + 2; n <= N; ++n) {
        const double nn = n, mm = m;
        const double a = std::sqrt((4.0 * nn * nn - 1.0) / (nn * nn - mm * mm));
        const double b = std::sqrt(((nn - 1.0) * (nn - 1.0) - mm * mm) / (4.0 * (nn - 1.0) * (nn - 1.0) - 1.0));
        row[n] = a * (x * row[n - 1] - b * row[n - 2]);
      }
      double* hrow = Hj + (size_t)m * NN;
      for (int n = m; n <= N; ++n) {
        const double nn = n, mm = m;
        double h = -nn * x * row[n];
        if (n > m) h += std::sqrt((2.0 * nn + 1.0) / (2.0 * nn - 1.0) * (nn * nn - mm * mm)) * row[n - 1];
        hrow[n] = h;
      }
    }
  }
  if (!reference_scaling) return;
  for (int m = 0; m <= M && m <= N; ++m)
    for (int n = m; n <= N; ++n) {
      const double f = reference_scale_factor(m, n);
      if (f == 1.0) continue;
      for (int j = 0; j < J; ++j) {
        P[((size_t)j * NM + m) * NN + n] *= f;
        H[((size_t)j * NM + m) * NN + n] *= f;
      }
    }
}

void lambdas(int I, double* lam) {
  // linspace(-pi, pi, I, endpoint=False): start + k * step with step = 2 pi / I.
  const double pi = 3.14159265358979323846;
  const double step = (2.0 * pi) / I;
  for (int i = 0; i < I; ++i) lam[i] = -pi + i * step;
}

}  // namespace swx
