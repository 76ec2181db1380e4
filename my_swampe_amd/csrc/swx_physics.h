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

// 1/x from the hardware reciprocal estimate and one Newton step (a few ulps).
__device__ __forceinline__ double rcp_nr(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return r * fma(-x, r, 2.0);
}

// grid_products with reciprocals in place of the divisions (1/(Phi + Phibar) once per
// point by rcp_nr, 1/taudrag from the coefficient table, 1/(2(1-mu^2)) passed in by the
// caller, who holds one latitude per lane): for the FP32-transform kernels, whose products
// are rounded to FP32, where the last-bit differences of the FP64 intermediates do not show.
__device__ inline void grid_products_recip(const double* cf, const double* geo, const GeoLayout& gl, int j, int i,
                                           int test, double eta, double Phi, double U, double V, double inv_denE,
                                           double out[8]) {
  const double Phibar = cf[C_PHIBAR];
  double PhiF = 0.0, F = 0.0, G = 0.0;
  if (test < 0) {
    const double Phieq =
        geo[gl.day + i] != 0.0 ? Phibar + (cf[C_DPHIEQ] * geo[gl.cosl + i]) * geo[gl.s1m + j] : Phibar;
    const double Q = cf[C_INVTAURAD] * (Phieq - (Phi + Phibar));
    const double Qc = Q < 0.0 ? 0.0 : Q;
    const double tot = Phi + Phibar;
    const double r = -Qc * rcp_nr(fabs(tot) > 0.0 ? tot : DBL_MIN);
    const double itd = cf[C_INVTD];  // 0 for the taudrag == -1 sentinel
    PhiF = Q;
    F = U * r - U * itd;
    G = V * r - V * itd;
  }
  out[0] = U * eta;
  out[1] = V * eta;
  out[2] = U * Phi;
  out[3] = V * Phi;
  out[4] = (U * U + V * V) * inv_denE;
  out[5] = PhiF;
  out[6] = F;
  out[7] = G;
}

// Initial fields' non-band-limited residue in the current physical level (RAW): eta, delta,
// Phi += c_curr R0p.  c_curr shrinks by about alpha per step; once it is below 1e-20 the
// term is under the FP64 rounding of the fields and is skipped (no R0p read).
__device__ inline void add_residue(const double* resid_c, const double* R0p, int b, long long plane, long long pix,
                                   double& eta, double& delta, double& Phi) {
  const double c = resid_c[2 * b + 1];
  if (fabs(c) > 1e-20) {
    const double* r = R0p + (long long)b * 3 * plane + pix;
    eta += c * r[0];
    delta += c * r[plane];
    Phi += c * r[2 * plane];
  }
}

// Physical carry of one grid point (phys: eta p,c; delta p,c; Phi p,c; U; V planes) given
// the current level eta, delta, Phi of this step.  Filter of the prev level (model.py:991-1087):
//   classic RA  prev' = curr + alpha (prev - 2 curr + new), new = the current level;
//   RAW         prev' = curr + alpha w d with d = prev - 2 curr + new; the current level is
//               eff = new - alpha (1 - w) d, so d = (prev - 2 curr + eff) / (1 - alpha (1 - w)).
// Phi_mid = the filtered prev level of Phi (the diagnostics' Phi extrema).
__device__ inline void physical_carry(const double* cf, int do_ra, int raw, double alpha, double* phys,
                                      long long plane, long long pix, double eta, double delta, double Phi,
                                      double U, double V, double& Phi_mid) {
  const double ep = phys[0 * plane + pix], ec = phys[1 * plane + pix];
  const double dp = phys[2 * plane + pix], dc = phys[3 * plane + pix];
  const double pp = phys[4 * plane + pix], pc = phys[5 * plane + pix];
  double em = ec, dm = dc, pm = pc;
  if (do_ra) {
    if (raw) {
      const double w = cf[C_WILLIAMS];
      const double g = alpha * w / (1.0 - alpha * (1.0 - w));
      em = ec + g * (ep - 2.0 * ec + eta);
      dm = dc + g * (dp - 2.0 * dc + delta);
      pm = pc + g * (pp - 2.0 * pc + Phi);
    } else {
      em = ec + alpha * (ep - 2.0 * ec + eta);
      dm = dc + alpha * (dp - 2.0 * dc + delta);
      pm = pc + alpha * (pp - 2.0 * pc + Phi);
    }
  }
  Phi_mid = pm;
  phys[0 * plane + pix] = em;
  phys[1 * plane + pix] = eta;
  phys[2 * plane + pix] = dm;
  phys[3 * plane + pix] = delta;
  phys[4 * plane + pix] = pm;
  phys[5 * plane + pix] = Phi;
  phys[6 * plane + pix] = U;
  phys[7 * plane + pix] = V;
}

// Semi-implicit lagged momentum remainder of (member b, row j, wavenumber m)
// (model.py:775-792, :1135-1143).  lagR receives the remainder the next step's Legendre
// inputs use, that of the previous level (kept in a.lag), and a.lag receives the current
// one, F_m + U_m / taudrag (U_m, V_m: the step's winds, the pinned initial ones in test 1).
// At initialization (`init`) the current remainder is both used and stored.  a.lagx (the
// exported carry) receives the value used on initialization and export_raw steps.
template <typename R>
__device__ inline void si_lag_exchange(const GridArgs<R>& a, const double* cf, int b, int j, int m,
                                       const double fm[8][2], const double Um[2], const double Vm[2], bool init,
                                       double lagR[2][2]) {
  const int Mp = a.M + 1;
  const long long o0 = (((long long)b * 2 + 0) * a.J + j) * Mp * 2 + 2 * m;
  const long long o1 = o0 + (long long)a.J * Mp * 2;
  double u[2] = {Um[0], Um[1]}, v[2] = {Vm[0], Vm[1]};
  if (a.test == 1 && !init) {
    u[0] = a.g0uv[o0];
    u[1] = a.g0uv[o0 + 1];
    v[0] = a.g0uv[o1];
    v[1] = a.g0uv[o1 + 1];
  }
  const double itd = cf[C_INVTD];
  const double cur[2][2] = {{fm[6][0] + u[0] * itd, fm[6][1] + u[1] * itd},
                            {fm[7][0] + v[0] * itd, fm[7][1] + v[1] * itd}};
  if (init) {
    for (int f = 0; f < 2; ++f)
      for (int c = 0; c < 2; ++c) lagR[f][c] = (double)(R)cur[f][c];
  } else {
    lagR[0][0] = a.lag[o0];
    lagR[0][1] = a.lag[o0 + 1];
    lagR[1][0] = a.lag[o1];
    lagR[1][1] = a.lag[o1 + 1];
  }
  a.lag[o0] = (R)cur[0][0];
  a.lag[o0 + 1] = (R)cur[0][1];
  a.lag[o1] = (R)cur[1][0];
  a.lag[o1 + 1] = (R)cur[1][1];
  if (init || a.export_raw) {
    a.lagx[o0] = (R)lagR[0][0];
    a.lagx[o0 + 1] = (R)lagR[0][1];
    a.lagx[o1] = (R)lagR[1][0];
    a.lagx[o1 + 1] = (R)lagR[1][1];
  }
}

// Per-(member, row) constants of the slot assembly, computed once per row so that the
// per-wavenumber work has no divisions (reciprocals of a^2 and taudrag; the products with
// them differ from the divided forms by at most an ulp).
struct SlotConsts {
  double dt, Phibar, tc, wj;
  double t1, t2, te, ia2, i2a2, itau;
  int expflag, forcflag;
};

__device__ inline SlotConsts slot_consts(const double* cf, double tc, double wj, int expflag, int forcflag) {
  SlotConsts k;
  const double dt = cf[C_DT], aa = cf[C_A];
  k.dt = dt;
  k.Phibar = cf[C_PHIBAR];
  k.tc = tc;
  k.wj = wj;
  k.expflag = expflag;
  k.forcflag = forcflag;
  k.ia2 = 1.0 / (aa * aa);
  k.i2a2 = 1.0 / (2.0 * aa * aa);
  if (!expflag) {
    k.t1 = tc / 4.0;
    k.t2 = (2.0 * dt / (aa * aa)) / 4.0;
    k.te = forcflag ? tc : tc / 2.0;
  } else {
    k.t1 = tc;
    k.t2 = 0.0;
    k.te = tc;
  }
  k.itau = 1.0 / cf[C_TAUDRAG];
  return k;
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
__device__ inline void assemble_slots(const SlotConsts& k, int m, const double f[8][2], const double Um[2],
                                      const double Vm[2], double out[kNT][2]) {
  const double dt = k.dt, Phibar = k.Phibar;
  const double mm = (double)m;
  for (int s = 0; s < kNT; ++s) out[s][0] = out[s][1] = 0.0;
  const double* Am = f[0];
  const double* Bm = f[1];
  const double* Cm = f[2];
  const double* Dm = f[3];
  const double* Em = f[4];
  const double* PFm = f[5];
  const double* Fm = f[6];
  const double* Gm = f[7];
  // i m z = (-m z.im, m z.re)
  if (!k.expflag) {
    const double t1 = k.t1, t2 = k.t2, te = k.te;
    double Bs[2], As[2], Bf[2], Af[2];
    for (int c = 0; c < 2; ++c) {
      Bf[c] = Bm[c] + Fm[c];
      Af[c] = Am[c] - Gm[c];
      Bs[c] = k.forcflag ? Bf[c] : Bm[c];
      As[c] = k.forcflag ? Af[c] : Am[c];
    }
    const double fPF = k.forcflag ? 1.0 : 0.0;
    const double imC[2] = {-mm * Cm[1], mm * Cm[0]};
    const double imBs[2] = {-mm * Bs[1], mm * Bs[0]};
    const double imBf[2] = {-mm * Bf[1], mm * Bf[0]};
    const double imAs[2] = {-mm * As[1], mm * As[0]};
    for (int c = 0; c < 2; ++c) {
      out[0][c] = -(t1 * imC[c]) - 0.5 * Phibar * (t1 * imBs[c]) + fPF * (dt * PFm[c]);
      out[1][c] = -0.5 * Phibar * (t2 * Em[c]);
      out[2][c] = t1 * imBf[c];
      out[3][c] = t2 * Em[c] + (t1 * imC[c]) * k.i2a2 + fPF * ((dt / 2.0) * PFm[c]) * k.ia2;
      out[4][c] = -(te * imAs[c]);
      out[5][c] = t1 * Dm[c] - 0.5 * Phibar * (t1 * As[c]);
      out[6][c] = t1 * Af[c];
      out[7][c] = (t1 * Dm[c]) * k.i2a2;
      out[8][c] = te * Bs[c];
    }
  } else {
    const double t1 = k.t1;
    const double imC[2] = {-mm * Cm[1], mm * Cm[0]};
    const double imA[2] = {-mm * Am[1], mm * Am[0]};
    const double imU[2] = {-mm * Um[1], mm * Um[0]};
    const double imV[2] = {-mm * Vm[1], mm * Vm[0]};
    const double imF[2] = {-mm * Fm[1], mm * Fm[0]};
    const double imG[2] = {-mm * Gm[1], mm * Gm[0]};
    for (int c = 0; c < 2; ++c) {
      out[0][c] = -(t1 * imC[c]) + (k.forcflag ? (2.0 * dt) * PFm[c] : 0.0);
      out[5][c] = t1 * Dm[c];
      out[4][c] = -(t1 * imA[c]);
      out[8][c] = t1 * Bm[c];
      if (k.forcflag) {
        out[2][c] = -((t1 * imU[c]) * k.itau) + t1 * imF[c];
        out[6][c] = (t1 * Vm[c]) * k.itau - t1 * Gm[c];
        out[4][c] += -((t1 * imV[c]) * k.itau) + t1 * imG[c];
        out[8][c] += (t1 * Um[c]) * k.itau + t1 * Fm[c];
      }
    }
  }
  for (int s = 0; s < kNT; ++s) {
    out[s][0] *= k.wj;
    out[s][1] *= k.wj;
  }
}

// The modified-Euler slots of assemble_slots in FP32, for the matrix-core grid kernels whose
// inputs (forward-DFT outputs) and outputs (the Legendre inputs) are FP32: the same
// combinations with the quadrature weight folded into the coefficients.
struct SlotConstsF {
  float t1w, t2w, tew, hPb, i2a2, dtw, dt2ia2w;
  int forcflag;
};

__device__ inline SlotConstsF slot_consts_f32(const SlotConsts& k) {
  SlotConstsF c;
  c.t1w = (float)(k.t1 * k.wj);
  c.t2w = (float)(k.t2 * k.wj);
  c.tew = (float)(k.te * k.wj);
  c.hPb = (float)(0.5 * k.Phibar);
  c.i2a2 = (float)k.i2a2;
  c.dtw = k.forcflag ? (float)(k.dt * k.wj) : 0.f;
  c.dt2ia2w = k.forcflag ? (float)((k.dt / 2.0) * k.ia2 * k.wj) : 0.f;
  c.forcflag = k.forcflag;
  return c;
}

__device__ inline void assemble_slots_f32(const SlotConstsF& k, int m, const float f[8][2], float out[kNT][2]) {
  // Unfused arithmetic: every kernel variant that inlines this rounds identically.
#pragma clang fp contract(off)
  const float mm = (float)m;
  const float *Am = f[0], *Bm = f[1], *Cm = f[2], *Dm = f[3], *Em = f[4], *PFm = f[5], *Fm = f[6], *Gm = f[7];
  float Bf[2], Af[2], Bs[2], As[2];
  for (int c = 0; c < 2; ++c) {
    Bf[c] = Bm[c] + Fm[c];
    Af[c] = Am[c] - Gm[c];
    Bs[c] = k.forcflag ? Bf[c] : Bm[c];
    As[c] = k.forcflag ? Af[c] : Am[c];
  }
  // i m z = (-m z.im, m z.re)
  const float imC[2] = {-mm * Cm[1], mm * Cm[0]};
  const float imBs[2] = {-mm * Bs[1], mm * Bs[0]};
  const float imBf[2] = {-mm * Bf[1], mm * Bf[0]};
  const float imAs[2] = {-mm * As[1], mm * As[0]};
  for (int c = 0; c < 2; ++c) {
    out[0][c] = -(k.t1w * imC[c]) - k.hPb * (k.t1w * imBs[c]) + k.dtw * PFm[c];
    out[1][c] = -k.hPb * (k.t2w * Em[c]);
    out[2][c] = k.t1w * imBf[c];
    out[3][c] = k.t2w * Em[c] + (k.t1w * imC[c]) * k.i2a2 + k.dt2ia2w * PFm[c];
    out[4][c] = -(k.tew * imAs[c]);
    out[5][c] = k.t1w * Dm[c] - k.hPb * (k.t1w * As[c]);
    out[6][c] = k.t1w * Af[c];
    out[7][c] = (k.t1w * Dm[c]) * k.i2a2;
    out[8][c] = k.tew * Bs[c];
  }
}

// Semi-implicit Legendre inputs of one (member, j, m) (semi_implicit_tdiff.py:163-190),
// w-weighted, in the slot order spectral_update reads: P0 tc i m A*, P1 tc i m B*,
// P2 tc i m C, P3 E, P4 2 dt PhiF (forced); H5 tc A*, H6 tc B*, H7 tc D; H8 unused.
// A* = A - G~, B* = B + F~ when forced, with (F~, G~) the lagged momentum remainder
// (relax) or the current F, G.
__device__ inline void assemble_slots_si(const double* cf, double tc, double wj, int m, int forcflag, int relax,
                                         const double f[8][2], const double lagR[2][2], double out[kNT][2]) {
  const double dt = cf[C_DT], mm = (double)m;
  const double* Fe = relax ? lagR[0] : f[6];
  const double* Ge = relax ? lagR[1] : f[7];
  double Ae[2], Be[2];
  for (int c = 0; c < 2; ++c) {
    Ae[c] = forcflag ? f[0][c] - Ge[c] : f[0][c];
    Be[c] = forcflag ? f[1][c] + Fe[c] : f[1][c];
  }
  const double wt = wj * tc;
  // i m z = (-m z.im, m z.re)
  out[0][0] = -wt * mm * Ae[1];
  out[0][1] = wt * mm * Ae[0];
  out[1][0] = -wt * mm * Be[1];
  out[1][1] = wt * mm * Be[0];
  out[2][0] = -wt * mm * f[2][1];
  out[2][1] = wt * mm * f[2][0];
  out[3][0] = wj * f[4][0];
  out[3][1] = wj * f[4][1];
  out[4][0] = forcflag ? wj * 2.0 * dt * f[5][0] : 0.0;
  out[4][1] = forcflag ? wj * 2.0 * dt * f[5][1] : 0.0;
  for (int c = 0; c < 2; ++c) {
    out[5][c] = wt * Ae[c];
    out[6][c] = wt * Be[c];
    out[7][c] = wt * f[3][c];
    out[8][c] = 0.0;
  }
}

// FP64 spectral update of one (member, m, n) from the transformed Legendre inputs
// fp (P basis) / fh (H basis), followed by the time filter on the stored levels.
// The carry terms Pw(X) of the reference are f2 * S (Gauss quadrature).
//   modified Euler (modEuler_tdiff.py:30-268) or explicit (explicit_tdiff.py:20-233)
//   with the implicit sigma6 diffusion (filters.py:11-27, :95-151);
//   semi-implicit (fl.si; semi_implicit_tdiff.py:126-264): leapfrog base f2 * S_prev,
//     slots fp = (Pw of tc i m A*, tc i m B*, tc i m C, E, 2 dt PhiF), fh = (Hw of
//     tc A*, tc B*, tc D), implicit gravity waves and (fl.relax) relaxation/drag
//     solved per degree in closed form, exponential hyperdiffusion.
// Time filter (model.py:991-1110): RAW (fl.raw) moves d = S_prev - 2 S_curr + new
//   into both levels; semi-implicit alone filters the prev lineage (classic RA);
//   explicit moves current -> previous; modified Euler keeps only the new level, and with
//   fl.lineage also advances the classic RA filter of the prev level in spectral form.
// Returns eff (the stored current level: Fourier rows of eta, delta, Phi) and nw (the
// new level before the RAW adjustment: the winds are inverted from it).
// The coefficients of one (member, m, n) that the modified-Euler update and the wind potentials
// read, for callers that load them ahead of the update (the values of the coefficient table).
struct UpdCoef {
  double dt, aa, Phibar, alpha, sg, sph, tcmn, fmn;
};

__device__ inline UpdCoef load_upd_coef(const double* cf, const CoefLayout& cl, int m, int n) {
  UpdCoef c;
  c.dt = cf[C_DT];
  c.aa = cf[C_A];
  c.Phibar = cf[C_PHIBAR];
  c.alpha = cf[C_ALPHA];
  c.sg = cf[cl.n_sigma + n];
  c.sph = cf[cl.n_sigmaphi + n];
  c.tcmn = cf[cl.n_tcmn + n];
  c.fmn = (m == 0 && n == 1) ? cf[C_FMN01] : 0.0;
  return c;
}

// sc_pre: the current level of this (member, m, n) already loaded by the caller, or null.
// pc: its coefficients already loaded by the caller (load_upd_coef), or null.
__device__ inline void spectral_update(const double* cf, const CoefLayout& cl, double f2, int m, int n,
                                       const StepFlags& fl, const double fp[5][2], const double fh[4][2],
                                       double* Scur, double* Sprev, long long so, long long fstride, double eff[3][2],
                                       double nw[3][2], const double (*sc_pre)[2] = nullptr,
                                       const UpdCoef* pc = nullptr) {
  const double dt = pc ? pc->dt : cf[C_DT], aa = pc ? pc->aa : cf[C_A], Phibar = pc ? pc->Phibar : cf[C_PHIBAR];
  const double nn = (double)n * (double)(n + 1);
  // The prev level is read only by the schemes and filters that use it (the modified-Euler
  // step without RAW reads and writes the current level only).
  const bool need_prev = fl.si || fl.raw || fl.expflag || fl.lineage;
  double sc[3][2], sp[3][2];
  for (int f = 0; f < 3; ++f)
    for (int c = 0; c < 2; ++c) {
      sc[f][c] = sc_pre ? sc_pre[f][c] : Scur[f * fstride + so + c];
      sp[f][c] = need_prev ? Sprev[f * fstride + so + c] : 0.0;
    }
  if (fl.si) {
    const double two_dt = 2.0 * dt, sia = cf[C_SIALPHA];
    const double lam = nn / (aa * aa);
    const double itd = fl.relax ? cf[C_INVTD] : 0.0, itr = fl.relax ? cf[C_INVTAURAD] : 0.0;
    const double xi = two_dt * sia, expl = two_dt * (1.0 - sia);
    const double kd = xi * itd, kr = xi * itr, gd = expl * itd, gr = expl * itr;
    const double b = xi * lam, cc = xi * Phibar;
    const double fmn = (m == 0 && n == 1) ? cf[C_FMN01] : 0.0;
    for (int c = 0; c < 2; ++c) {
      const double e0 = f2 * sp[0][c], d0 = f2 * sp[1][c], p0 = f2 * sp[2][c];
      const double fr = c == 0 ? fmn : 0.0;
      const double eta_star = e0 - fp[0][c] + fh[1][c] - gd * (e0 - fr);
      const double delta_star =
          d0 + fp[1][c] + fh[0][c] + two_dt * lam * fp[3][c] + expl * lam * p0 - gd * d0;
      double Phi_star = p0 - fp[2][c] + fh[2][c] - expl * Phibar * d0 - gr * p0;
      // fp[4] = Pw(2 dt PhiF) is zero unless forced; relax implies forced.
      Phi_star += fp[4][c];
      if (fl.relax) Phi_star += two_dt * itr * (f2 * sc[2][c]);
      const double dn = (delta_star + b * Phi_star / (1.0 + kr)) / ((1.0 + kd) + b * cc / (1.0 + kr));
      nw[0][c] = (eta_star + kd * fr) / (1.0 + kd);
      nw[1][c] = dn;
      nw[2][c] = (Phi_star - cc * dn) / (1.0 + kr);
    }
    if (fl.diffflag) {
      const double sg = cf[cl.n_sigexp + n], sph = cf[cl.n_sigexpphi + n];
      for (int c = 0; c < 2; ++c) {
        nw[0][c] *= sg;
        nw[1][c] *= sg;
        nw[2][c] *= sph;
      }
    }
  } else {
    for (int c = 0; c < 2; ++c) {
      const double eo = f2 * sc[0][c];
      const double d_o = f2 * sc[1][c];
      const double po = f2 * sc[2][c];
      if (!fl.expflag) {
        const double t2 = (2.0 * dt / (aa * aa)) / 4.0;
        nw[2][c] = po - dt * Phibar * d_o - Phibar * 0.5 * ((1.0 / (aa * aa)) * (nn * po)) + fp[0][c] + nn * fp[1][c] +
                   fh[0][c];
        nw[1][c] = d_o + nn * (t2 * po) - Phibar * (nn * d_o) / (aa * aa) + fp[2][c] + nn * fp[3][c] + fh[1][c] +
                   nn * fh[2][c];
        nw[0][c] = eo + fp[4][c] + fh[3][c];
      } else {
        const double ep = f2 * sp[0][c];
        const double dp = f2 * sp[1][c];
        const double pp = f2 * sp[2][c];
        nw[2][c] = pp - 2.0 * dt * Phibar * d_o + fp[0][c] + fh[0][c];
        nw[1][c] = dp + fp[2][c] + fh[1][c];
        nw[0][c] = ep + fp[4][c] + fh[3][c];
      }
    }
    if (fl.diffflag) {
      const double sg = pc ? pc->sg : cf[cl.n_sigma + n], sph = pc ? pc->sph : cf[cl.n_sigmaphi + n];
      for (int c = 0; c < 2; ++c) {
        nw[2][c] *= sph;
        nw[1][c] *= sg;
        nw[0][c] *= sg;
      }
    }
  }
  const double alpha = pc ? pc->alpha : cf[C_ALPHA];
  for (int f = 0; f < 3; ++f)
    for (int c = 0; c < 2; ++c) {
      double cur = nw[f][c];
      if (fl.raw) {
        if (fl.do_ra) {
          const double w = cf[C_WILLIAMS];
          const double d = sp[f][c] - 2.0 * sc[f][c] + nw[f][c];
          Sprev[f * fstride + so + c] = sc[f][c] + alpha * w * d;
          cur = nw[f][c] - alpha * (1.0 - w) * d;
        } else {
          Sprev[f * fstride + so + c] = sc[f][c];
        }
      } else if (fl.si) {
        Sprev[f * fstride + so + c] =
            fl.do_ra ? sc[f][c] + alpha * (sp[f][c] - 2.0 * sc[f][c] + nw[f][c]) : sc[f][c];
      } else if (fl.expflag) {
        Sprev[f * fstride + so + c] = sc[f][c];
      } else if (fl.lineage) {
        // Modified Euler: the prev lineage of the physical fields (model.py:988-1005) in
        // spectral form; the step itself never reads it.
        Sprev[f * fstride + so + c] =
            fl.do_ra ? sc[f][c] + alpha * (sp[f][c] - 2.0 * sc[f][c] + nw[f][c]) : sc[f][c];
      }
      Scur[f * fstride + so + c] = cur;
      eff[f][c] = cur;
    }
}

// Per-member coefficients of the initial fields' non-band-limited residue in the filtered
// lineages (prev, curr): the new level is band-limited, so the residue only moves between the
// two filtered levels (same recurrences as spectral_update's time filter).
__device__ inline void residue_step(const double* cf, const StepFlags& fl, double* rc) {
  const double cp = rc[0], cc = rc[1], alpha = cf[C_ALPHA];
  if (fl.raw && fl.do_ra) {
    const double w = cf[C_WILLIAMS];
    const double d = cp - 2.0 * cc;
    rc[0] = cc + alpha * w * d;
    rc[1] = -alpha * (1.0 - w) * d;
  } else if (fl.si && fl.do_ra) {
    rc[0] = cc + alpha * (cp - 2.0 * cc);
    rc[1] = 0.0;
  } else {
    rc[0] = cc;
    rc[1] = 0.0;
  }
}

// Wind potentials of the inverse transform (spectral_transform.py:577-582):
// delta_hat = delta a/(n(n+1)), eta_hat = (eta - f) a/(n(n+1)), zero at n = 0.
__device__ inline void wind_potentials(const double* cf, const CoefLayout& cl, int m, int n, const double eta[2],
                                       const double delta[2], double dh[2], double eh[2],
                                       const UpdCoef* pc = nullptr) {
  const double tcmn = pc ? pc->tcmn : cf[cl.n_tcmn + n];
  const double f = pc ? pc->fmn : ((m == 0 && n == 1) ? cf[C_FMN01] : 0.0);
  for (int c = 0; c < 2; ++c) {
    dh[c] = (n == 0) ? 0.0 : delta[c] * tcmn;
    eh[c] = (n == 0) ? 0.0 : (eta[c] - (c == 0 ? f : 0.0)) * tcmn;
  }
}

// Adjoint of grid_products at one grid point (reverse mode, the Jacobian of
// forcing.py:65-155 and initial_conditions.py:283-292 transposed).
// pbar: cotangents of A, B, C, D, E, PhiF, F, G.  Returns the cotangents of
// eta, Phi, U, V and accumulates d/d(taurad, taudrag, DPhieq) into dpar[3].
// The Q < 0 clamp passes the gradient where Q >= 0 (jnp.where(Q < 0, 0, Q)).
__device__ inline void grid_products_adjoint(const double* cf, const double* geo, const GeoLayout& gl, int j, int i,
                                             int test, double eta, double Phi, double U, double V,
                                             const double pbar[8], double out[4], double dpar[3]) {
  const double Ab = pbar[0], Bb = pbar[1], Cb = pbar[2], Db = pbar[3], Eb = pbar[4];
  const double two_den = 2.0 / geo[gl.denomE + j];  // per row
  double etab = Ab * U + Bb * V;
  double Phib = Cb * U + Db * V;
  double Ub = Ab * eta + Cb * Phi + Eb * (two_den * U);
  double Vb = Bb * eta + Db * Phi + Eb * (two_den * V);
  if (test < 0) {
    const double Phibar = cf[C_PHIBAR];
    const double day = geo[gl.day + i];
    const double dphi_term = (geo[gl.cosl + i]) * geo[gl.s1m + j];
    const double Phieq = day != 0.0 ? Phibar + (cf[C_DPHIEQ] * geo[gl.cosl + i]) * geo[gl.s1m + j] : Phibar;
    const double itr = cf[C_INVTAURAD];
    const double Q = itr * (Phieq - (Phi + Phibar));
    const double pass = Q < 0.0 ? 0.0 : 1.0;
    const double Qc = Q < 0.0 ? 0.0 : Q;
    const double tot = Phi + Phibar;
    const double safe = fabs(tot) > 0.0 ? tot : DBL_MIN;
    const double Fb = pbar[6], Gb = pbar[7];
    const double td = cf[C_TAUDRAG];
    const double inv_td = td == -1.0 ? 0.0 : 1.0 / td;
    // F = -U Qc / tot - U / td, G = -V Qc / tot - V / td (one reciprocal of tot per point)
    const double inv = 1.0 / safe;
    const double FUGV = Fb * U + Gb * V;
    Ub += Fb * (-Qc * inv - inv_td);
    Vb += Gb * (-Qc * inv - inv_td);
    const double Qcb = -FUGV * inv;
    Phib += FUGV * Qc * (inv * inv);
    const double Qb = pbar[5] + Qcb * pass;
    Phib += Qb * (-itr);
    dpar[0] += Qb * (-Q * itr);                                   // d/dtaurad
    if (td != -1.0) dpar[1] += FUGV * (inv_td * inv_td);  // d/dtaudrag
    if (day != 0.0) dpar[2] += Qb * itr * dphi_term;              // d/dDPhieq
  }
  out[0] = etab;
  out[1] = Phib;
  out[2] = Ub;
  out[3] = Vb;
}

// Adjoint of assemble_slots for the modified-Euler scheme: slot cotangents
// tau[9][2] (without the w_j factor applied by the caller) -> cotangents of
// the Fourier coefficients of A, B, C, D, E, PhiF, F, G.  For a term
// c * (i m) * Z the cotangent of Z is c * (-i m) * tau.
__device__ inline void assemble_slots_adjoint(const double* cf, double tc, double wj, int m, int forcflag,
                                              const double tau_in[kNT][2], double fb[8][2]) {
  const double dt = cf[C_DT], aa = cf[C_A], Phibar = cf[C_PHIBAR];
  const double mm = (double)m;
  const double t1 = tc / 4.0;
  const double t2 = (2.0 * dt / (aa * aa)) / 4.0;
  const double te = forcflag ? tc : tc / 2.0;
  const double fPF = forcflag ? 1.0 : 0.0;
  double tau[kNT][2];
  for (int k = 0; k < kNT; ++k) {
    tau[k][0] = tau_in[k][0] * wj;
    tau[k][1] = tau_in[k][1] * wj;
  }
  // J(z) = (-i m) z = (m z.im, -m z.re)
  auto Jre = [&](int k) { return mm * tau[k][1]; };
  auto Jim = [&](int k) { return -mm * tau[k][0]; };
  double Cb[2], Bsb[2], PFb[2], Eb[2], Bfb[2], Asb[2], Db[2], Afb[2];
  Cb[0] = -t1 * Jre(0) + (t1 / (2.0 * aa * aa)) * Jre(3);
  Cb[1] = -t1 * Jim(0) + (t1 / (2.0 * aa * aa)) * Jim(3);
  Bsb[0] = -0.5 * Phibar * t1 * Jre(0) + te * tau[8][0];
  Bsb[1] = -0.5 * Phibar * t1 * Jim(0) + te * tau[8][1];
  Bfb[0] = t1 * Jre(2);
  Bfb[1] = t1 * Jim(2);
  Asb[0] = -te * Jre(4) - 0.5 * Phibar * t1 * tau[5][0];
  Asb[1] = -te * Jim(4) - 0.5 * Phibar * t1 * tau[5][1];
  Afb[0] = t1 * tau[6][0];
  Afb[1] = t1 * tau[6][1];
  for (int c = 0; c < 2; ++c) {
    PFb[c] = fPF * dt * tau[0][c] + fPF * ((dt / 2.0) / (aa * aa)) * tau[3][c];
    Eb[c] = -0.5 * Phibar * t2 * tau[1][c] + t2 * tau[3][c];
    Db[c] = t1 * tau[5][c] + (t1 / (2.0 * aa * aa)) * tau[7][c];
    // Bs = forcflag ? B + F : B ; Bf = B + F ; As = forcflag ? A - G : A ; Af = A - G
    fb[1][c] = Bsb[c] + Bfb[c];                         // B
    fb[6][c] = (forcflag ? Bsb[c] : 0.0) + Bfb[c];      // F
    fb[0][c] = Asb[c] + Afb[c];                         // A
    fb[7][c] = -((forcflag ? Asb[c] : 0.0) + Afb[c]);   // G
    fb[2][c] = Cb[c];
    fb[3][c] = Db[c];
    fb[4][c] = Eb[c];
    fb[5][c] = PFb[c];
  }
}

}  // namespace swx
