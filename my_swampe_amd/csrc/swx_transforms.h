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

}  // namespace swx
