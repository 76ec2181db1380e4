// Launchers of the standalone FP64 transform kernels (swx_transforms.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace swx {

// Vorticity and divergence spectra from Fourier winds (spectral_transform.py:599-650,
// diagnostic_eta_delta): for B members, Um / Vm (B, J, M+1) complex, shared Pmn / Hmn
// (J, M+1, N+1), w (J), tstepcoeff and mJarray (J, M+1), fmn (M+1, N+1) complex, dt:
//   zeta_mn  = sum_j w_j c_jm (i m_jm V_jm P_jmn + U_jm H_jmn),  c = tstepcoeff / (2 dt)
//   eta_mn   = zeta_mn + fmn
//   delta_mn = sum_j w_j c_jm (i m_jm U_jm P_jmn - V_jm H_jmn)
// written to etamn / deltamn (B, M+1, N+1) complex.
hipError_t launch_diag_eta_delta_spec(const double* Um, const double* Vm, const double* fmn, int B, int J, int M,
                                      int N, const double* Pmn, const double* Hmn, const double* w,
                                      const double* tstepcoeff, const double* mJarray, double dt, double* etamn,
                                      double* deltamn, hipStream_t s);
// Grids of spectra S (B, M+1, N+1) complex: invrs_fft(invrs_leg(S)) (spectral_transform.py:
// 331-348, 438-481), (B, J, I) complex, for up to two fields at once (out1 may be null).
hipError_t launch_spec_to_grid(const double* S0, const double* S1, int B, int J, int I, int M, int N,
                               const double* Pmn, double* out0, double* out1, hipStream_t s);
// fwd_fft_trunc (spectral_transform.py:290-328) of K x J rows of C columns, real or interleaved
// complex: out (K, J, Mo) complex, Mo = min(M + 1, I), rows zero-padded or truncated to I.
// Needs I <= kMaxDftLength.
hipError_t launch_fwd_dft_trunc(const double* x, int is_complex, int K, int J, int C, int I, int Mo, double* out,
                                hipStream_t s);
// invrs_fft (spectral_transform.py:331-348): X (K, J, C) complex -> ifft(I X, n=I) (K, J, I)
// complex.  Needs I <= kMaxDftLength.
hipError_t launch_dft_synth(const double* X, int K, int J, int C, int I, double* out, hipStream_t s);
// fwd_leg / fwd_leg_w (spectral_transform.py:370-435): data (K, J, M+1) complex, basis
// (J, M+1, N+1), w (J) or null -> out (K, M+1, N+1) complex.
hipError_t launch_fwd_leg(const double* data, const double* basis, const double* w, int K, int J, int M, int N,
                          double* out, hipStream_t s);
// invrs_leg (spectral_transform.py:438-481): S (K, M+1, N+1) complex -> (K, J, I) complex in
// the reference's layout (+m at column m, conjugates at I - m).  Needs I >= M + 1.
hipError_t launch_invrs_leg(const double* S, const double* basis, int K, int J, int I, int M, int N, double* out,
                            hipStream_t s);
// invrsUV_with_coeffs (spectral_transform.py:535-596): U, V (K, J, I) and Um, Vm (K, J, M+1)
// complex, any of them null.  Needs I >= 2 M + 1 and I <= kMaxDftLength.
hipError_t launch_invrs_uv(const double* deltamn, const double* etamn, const double* fmn, int K, int J, int I, int M,
                           int N, const double* Pmn, const double* Hmn, const double* tstepcoeffmn,
                           const double* marray, double* U, double* V, double* Um, double* Vm, hipStream_t s);

// Longest longitude row the DFT kernels stage in LDS (twiddles and row, 64 KiB at this length).
constexpr int kMaxDftLength = 2048;

}  // namespace swx
