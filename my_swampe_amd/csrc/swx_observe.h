// Launchers of the phase-curve observable kernels (swx_observe.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace swx {

// Largest longitude count and phase count one launch takes (LDS rows of one workgroup).
constexpr int kPhaseCurveMaxI = 1024;
constexpr int kPhaseCurveMaxK = 2048;

// flux[b][k] (and, when flux_raw is not null, the flux before normalization) of the grids
// Phi + b * mstride, b < B, each J x I row-major.  mus: the J Gauss nodes (device).
hipError_t launch_phase_curve(const double* Phi, long long mstride, int B, int J, int I, const double* mus,
                              const double* phases, int K, int normalize, double* flux, double* flux_raw,
                              hipStream_t s);
// g_Phi[b] (J x I, dense, overwritten) from the flux cotangent g_flux[b][k] and flux_raw.
hipError_t launch_phase_curve_vjp(const double* g_flux, const double* flux_raw, int B, int J, int I,
                                  const double* mus, const double* phases, int K, int normalize, double* g_Phi,
                                  hipStream_t s);
size_t phase_curve_lds_bytes(int J, int I, int K);
size_t phase_curve_vjp_lds_bytes(int J, int I, int K);

}  // namespace swx
