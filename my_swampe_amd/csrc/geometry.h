// FP64 host geometry for the spectral transforms (see geometry.cpp).
#pragma once

namespace swx {

// J (Gauss latitudes), I (longitudes) and the default dt for truncation M.
// Only M in {42, 63, 106} is supported (initial_conditions.py:66-81).
bool resolution(int M, int* J, int* I, int* dt_default);

// Gauss-Legendre nodes (ascending) and weights of order J.
void gauss_legendre(int J, double* x, double* w);

// Orthonormal associated Legendre basis without Condon-Shortley phase and its
// (1-mu^2) d/dmu, both [J][M+1][N+1], zero for n < m.  With
// reference_scaling, entries whose factorial normalization underflows in the
// reference are scaled the same way (zeroed or precision-reduced).
void legendre_basis(int M, int N, int J, const double* mus, double* P, double* H, bool reference_scaling);

// sqrt(r_double) / sqrt(r_exact) of the reference normalization of (m, n):
// 1 where it is exact, < 1 or 0 where it under/overflows.
double reference_scale_factor(int m, int n);

// Longitudes -pi + 2 pi i / I.
void lambdas(int I, double* lam);

}  // namespace swx
