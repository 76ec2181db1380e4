// Grid-point physics and Legendre-input assembly shared by the plain and the
// MFMA grid kernels (FP64 arithmetic in both).
#pragma once

#include <cfloat>

#include "swx_kernels.h"

namespace swx {

// Forcing Q/F/G (forcing.py:65-155, computed whenever test is None:
// model.py:560-603) and the nonlinear products A..E
// (initial_conditions.py:283-292) at one grid point.
// out: A, B, C, D, E, PhiF, F, G.
__device__ inline void grid_products(const double* cf, const double* geo, const GeoLayout& gl, int j, int i, int test,
                                     double eta, double Phi, double U, double V, double out[8]) {
  const double Phibar = cf[C_PHIBAR];
  double PhiF = 0.0, F = 0.0, G = 0.0;
  if (test < 0) {
    const double Phieq =
        geo[gl.day + i] != 0.0 ? Phibar + (cf[C_DPHIEQ] * geo[gl.cosl + i]) * geo[gl.s1m + j] : Phibar;
    const double Q = cf[C_INVTAURAD] * (Phieq - (Phi + Phibar));
    const double Qc = Q < 0.0 ? 0.0 : Q;
    const double tot = Phi + Phibar;
    const double safe = fabs(tot) > 0.0 ? tot : DBL_MIN;
    const double Ru = -U * Qc / safe;
    const double Rv = -V * Qc / safe;
    const double td = cf[C_TAUDRAG];
    PhiF = Q;
    if (td == -1.0) {
      F = Ru;
      G = Rv;
    } else {
      F = Ru - (U / td);
      G = Rv - (V / td);
    }
  }
  out[0] = U * eta;
  out[1] = V * eta;
  out[2] = U * Phi;
  out[3] = V * Phi;
  out[4] = (U * U + V * V) / geo[gl.denomE + j];
  out[5] = PhiF;
  out[6] = F;
  out[7] = G;
}

// w-weighted Legendre inputs of one (member, j, m) from the truncated Fourier
// coefficients f[8][2] (A, B, C, D, E, PhiF, F, G) and, for the explicit
// scheme, the wind coefficients Um/Vm.  P-basis slots 0..4, H-basis 5..8:
//   modified Euler (modEuler_tdiff.py:72-110, :162-200, :253-259) regrouped by
//   basis and by the n(n+1) weight applied after the transform:
//     P0 (Phi)          -t1 i m C - Phibar/2 t1 i m B* + dt PhiF
//     P1 (Phi, n(n+1))  -Phibar/2 t2 E
//     P2 (delta)        t1 i m Bf
//     P3 (delta, n(n+1)) t2 E + t1 i m C / (2 a^2) + (dt/2) PhiF / a^2
//     P4 (eta)          -te i m A*
//     H5 (Phi)          t1 D - Phibar/2 t1 A*
//     H6 (delta)        t1 Af
//     H7 (delta, n(n+1)) t1 D / (2 a^2)
//     H8 (eta)          te B*
//   with t1 = tstepcoeff/4, t2 = tstepcoeff2/4, te = tstepcoeff (forced) or
//   tstepcoeff/2, Bf = B + F, Af = A - G, B* / A* = Bf / Af when forced else B / A.
//   explicit (explicit_tdiff.py:63-75, :134-153, :207-224), t1 = tstepcoeff:
//     P0 -t1 i m C + 2 dt PhiF;  H5 t1 D;  P2 forcing (-t1 i m U + t1 i m taudrag F)/taudrag
//     H6 (t1 V - taudrag t1 G)/taudrag;  P4 -t1 i m A (+ forcing);  H8 t1 B (+ forcing).
__device__ inline void assemble_slots(const double* cf, double tc, double wj, int m, int expflag, int forcflag,
                                      const double f[8][2], const double Um[2], const double Vm[2],
                                      double out[kNT][2]) {
  const double dt = cf[C_DT], aa = cf[C_A], Phibar = cf[C_PHIBAR];
  const double mm = (double)m;
  for (int k = 0; k < kNT; ++k) out[k][0] = out[k][1] = 0.0;
  const double* Am = f[0];
  const double* Bm = f[1];
  const double* Cm = f[2];
  const double* Dm = f[3];
  const double* Em = f[4];
  const double* PFm = f[5];
  const double* Fm = f[6];
  const double* Gm = f[7];
  // i m z = (-m z.im, m z.re)
  if (!expflag) {
    const double t1 = tc / 4.0;
    const double t2 = (2.0 * dt / (aa * aa)) / 4.0;
    const double te = forcflag ? tc : tc / 2.0;
    const double ia2 = 1.0 / (aa * aa);
    double Bs[2], As[2], Bf[2], Af[2];
    for (int c = 0; c < 2; ++c) {
      Bf[c] = Bm[c] + Fm[c];
      Af[c] = Am[c] - Gm[c];
      Bs[c] = forcflag ? Bf[c] : Bm[c];
      As[c] = forcflag ? Af[c] : Am[c];
    }
    const double fPF = forcflag ? 1.0 : 0.0;
    const double imC[2] = {-mm * Cm[1], mm * Cm[0]};
    const double imBs[2] = {-mm * Bs[1], mm * Bs[0]};
    const double imBf[2] = {-mm * Bf[1], mm * Bf[0]};
    const double imAs[2] = {-mm * As[1], mm * As[0]};
    for (int c = 0; c < 2; ++c) {
      out[0][c] = -(t1 * imC[c]) - 0.5 * Phibar * (t1 * imBs[c]) + fPF * (dt * PFm[c]);
      out[1][c] = -0.5 * Phibar * (t2 * Em[c]);
      out[2][c] = t1 * imBf[c];
      out[3][c] = t2 * Em[c] + (t1 * imC[c]) / (2.0 * aa * aa) + fPF * ((dt / 2.0) * PFm[c]) * ia2;
      out[4][c] = -(te * imAs[c]);
      out[5][c] = t1 * Dm[c] - 0.5 * Phibar * (t1 * As[c]);
      out[6][c] = t1 * Af[c];
      out[7][c] = (t1 * Dm[c]) / (2.0 * aa * aa);
      out[8][c] = te * Bs[c];
    }
  } else {
    const double t1 = tc;
    const double tau = cf[C_TAUDRAG];
    const double imC[2] = {-mm * Cm[1], mm * Cm[0]};
    const double imA[2] = {-mm * Am[1], mm * Am[0]};
    const double imU[2] = {-mm * Um[1], mm * Um[0]};
    const double imV[2] = {-mm * Vm[1], mm * Vm[0]};
    const double imF[2] = {-mm * Fm[1], mm * Fm[0]};
    const double imG[2] = {-mm * Gm[1], mm * Gm[0]};
    for (int c = 0; c < 2; ++c) {
      out[0][c] = -(t1 * imC[c]) + (forcflag ? (2.0 * dt) * PFm[c] : 0.0);
      out[5][c] = t1 * Dm[c];
      out[4][c] = -(t1 * imA[c]);
      out[8][c] = t1 * Bm[c];
      if (forcflag) {
        out[2][c] = -((t1 * imU[c]) / tau) + t1 * imF[c];
        out[6][c] = (t1 * Vm[c]) / tau - t1 * Gm[c];
        out[4][c] += -((t1 * imV[c]) / tau) + t1 * imG[c];
        out[8][c] += (t1 * Um[c]) / tau + t1 * Fm[c];
      }
    }
  }
  for (int k = 0; k < kNT; ++k) {
    out[k][0] *= wj;
    out[k][1] *= wj;
  }
}

}  // namespace swx
