// Spectral-space pieces of the adjoint Legendre step shared by the plain and the
// matrix-core kernels (FP64 arithmetic in both).  Reference functions differentiated:
// modEuler_tdiff.py:30-268 (the FP64 update and diffusion) and
// spectral_transform.py:577-594 (invrsUV).
#pragma once

#include "swx_kernels.h"

namespace swx {

// Mode A at one (member, m, n): mu = sigma lam, then in place lam <- f2 C^T(mu) (the
// transpose of the carry terms of the modified-Euler update), and the cotangents of the
// Legendre inputs before the transposed contraction, in slot order: P slots 0..4 =
// (mu_Phi, n(n+1) mu_Phi, mu_delta, n(n+1) mu_delta, mu_eta), H slots 5..8 =
// (mu_Phi, mu_delta, n(n+1) mu_delta, mu_eta).
__device__ inline void leg_adj_carry(const double* cf, const CoefLayout& cl, double f2, int n, int diffflag,
                                     double* Lam, long long so, long long fst, double slot[kNT][2]) {
  const double dt = cf[C_DT], aa = cf[C_A], Phibar = cf[C_PHIBAR];
  const double t2 = (2.0 * dt / (aa * aa)) / 4.0;
  const double nn = (double)n * (double)(n + 1);
  const double sg = diffflag ? cf[cl.n_sigma + n] : 1.0;
  const double sp = diffflag ? cf[cl.n_sigmaphi + n] : 1.0;
  for (int c = 0; c < 2; ++c) {
    const double me = sg * Lam[0 * fst + so + c];
    const double md = sg * Lam[1 * fst + so + c];
    const double mp = sp * Lam[2 * fst + so + c];
    const double xp = mp * (1.0 - Phibar * 0.5 * (nn / (aa * aa))) + nn * t2 * md;
    const double xd = -dt * Phibar * mp + md * (1.0 - Phibar * nn / (aa * aa));
    Lam[0 * fst + so + c] = f2 * me;
    Lam[1 * fst + so + c] = f2 * xd;
    Lam[2 * fst + so + c] = f2 * xp;
    slot[0][c] = mp;
    slot[1][c] = nn * mp;
    slot[2][c] = md;
    slot[3][c] = nn * md;
    slot[4][c] = me;
    slot[5][c] = mp;
    slot[6][c] = md;
    slot[7][c] = nn * md;
    slot[8][c] = me;
  }
}

// Inputs of the recomputed inverse transform of S_t at (member, m, n): eta, delta, Phi and
// the wind potentials delta tcmn, (eta - f) tcmn (spectral_transform.py:577-583).
__device__ inline void leg_adj_recompute_inputs(const double* cf, const CoefLayout& cl, int m, int n,
                                                const double* St, long long so, long long fst, double x[5][2]) {
  const double tcmn = cf[cl.n_tcmn + n];
  const double f = (m == 0 && n == 1) ? cf[C_FMN01] : 0.0;
  for (int c = 0; c < 2; ++c) {
    const double e = St[0 * fst + so + c], d = St[1 * fst + so + c], p = St[2 * fst + so + c];
    x[0][c] = e;
    x[1][c] = d;
    x[2][c] = p;
    x[3][c] = n == 0 ? 0.0 : d * tcmn;
    x[4][c] = n == 0 ? 0.0 : (e - (c == 0 ? f : 0.0)) * tcmn;
  }
}

// Modes B / TERM at one (member, m, n): lam (+)= invLeg^T of the Fourier cotangents, from the
// transposed contractions p = P^T(eta, delta, Phi, U, V bar) (re, im pairs) and
// h = H^T(U bar, V bar), with the wind-inversion adjoint
//   dhat = i m P^T(U bar) - H^T(V bar),  ehat = H^T(U bar) + i m P^T(V bar).
__device__ inline void leg_adj_lambda(const double* cf, const CoefLayout& cl, int m, int n, const double p[10],
                                      const double h[4], bool assign, double* Lam, long long so, long long fst) {
  const double mm = (double)m;
  const double dh[2] = {-mm * p[7] - h[2], mm * p[6] - h[3]};
  const double eh[2] = {h[0] - mm * p[9], h[1] + mm * p[8]};
  const double tcmn = n == 0 ? 0.0 : cf[cl.n_tcmn + n];
  for (int c = 0; c < 2; ++c) {
    const double le = p[0 + c] + tcmn * eh[c];
    const double ld = p[2 + c] + tcmn * dh[c];
    const double lp = p[4 + c];
    if (assign) {
      Lam[0 * fst + so + c] = le;
      Lam[1 * fst + so + c] = ld;
      Lam[2 * fst + so + c] = lp;
    } else {
      Lam[0 * fst + so + c] += le;
      Lam[1 * fst + so + c] += ld;
      Lam[2 * fst + so + c] += lp;
    }
  }
}

}  // namespace swx
