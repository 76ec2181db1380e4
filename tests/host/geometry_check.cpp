// Host geometry under AddressSanitizer / UndefinedBehaviorSanitizer (built and run by
// tests/test_host_sanitizers.py with g++): every entry point of geometry.h at T42, T63 and T106,
// with checks the mathematics fixes (node order and mirror symmetry, weight sum 2, quadrature
// orthonormality of the basis, the longitudes), and the unsupported truncation refused.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../my_swampe_amd/csrc/geometry.h"

static int failures = 0;
static void expect(bool ok, const char* what, int M) {
  if (!ok) {
    std::printf("FAIL M=%d: %s\n", M, what);
    ++failures;
  }
}

int main() {
  for (int M : {42, 63, 106}) {
    int J = 0, I = 0, dt = 0;
    expect(swx::resolution(M, &J, &I, &dt), "resolution", M);
    const int N = M, Mp = M + 1, Np = N + 1;
    std::vector<double> mus(J), w(J), lam(I);
    swx::gauss_legendre(J, mus.data(), w.data());
    double wsum = 0.0, asym = 0.0;
    bool ascending = true;
    for (int j = 0; j < J; ++j) {
      wsum += w[j];
      asym = std::fmax(asym, std::fabs(mus[j] + mus[J - 1 - j]) + std::fabs(w[j] - w[J - 1 - j]));
      if (j > 0 && !(mus[j] > mus[j - 1])) ascending = false;
    }
    expect(ascending, "nodes ascending", M);
    expect(std::fabs(wsum - 2.0) < 1e-13, "weights sum to 2", M);
    expect(asym < 1e-14, "nodes and weights mirror", M);
    std::vector<double> P((size_t)J * Mp * Np), H(P.size());
    for (int scaled = 0; scaled < 2; ++scaled) {
      swx::legendre_basis(M, N, J, mus.data(), P.data(), H.data(), scaled != 0);
      bool finite = true;
      for (size_t e = 0; e < P.size(); ++e) finite = finite && std::isfinite(P[e]) && std::isfinite(H[e]);
      expect(finite, "basis finite", M);
    }
    // Orthonormality of the exactly normalized basis under the Gauss quadrature.
    swx::legendre_basis(M, N, J, mus.data(), P.data(), H.data(), false);
    double err = 0.0;
    for (int m = 0; m < Mp; ++m)
      for (int n = m; n < Np; ++n)
        for (int k = m; k < Np; ++k) {
          double s = 0.0;
          for (int j = 0; j < J; ++j) s += w[j] * P[((size_t)j * Mp + m) * Np + n] * P[((size_t)j * Mp + m) * Np + k];
          err = std::fmax(err, std::fabs(s - (n == k ? 1.0 : 0.0)));
        }
    expect(err < 1e-11, "quadrature orthonormality", M);
    double scale = swx::reference_scale_factor(0, 0);
    expect(std::isfinite(scale) && scale >= 0.0 && scale <= 1.0, "reference scale factor", M);
    swx::lambdas(I, lam.data());
    double lerr = 0.0;
    for (int i = 0; i < I; ++i) lerr = std::fmax(lerr, std::fabs(lam[i] - (-M_PI + 2.0 * M_PI * i / I)));
    expect(lerr < 1e-14, "longitudes", M);
    std::printf("M=%d J=%d I=%d orthonormality %.2e\n", M, J, I, err);
  }
  int J = 0, I = 0, dt = 0;
  expect(!swx::resolution(50, &J, &I, &dt), "M=50 refused", 50);
  return failures == 0 ? 0 : 1;
}
