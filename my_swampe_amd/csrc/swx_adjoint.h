// Argument blocks and launchers of the adjoint kernels (swx_adjoint.hip).
#pragma once

#include "swx_kernels.h"

namespace swx {

enum LegAdjMode : int { LADJ_A = 0, LADJ_B = 1, LADJ_TERM = 2 };
enum GridAdjMode : int { GADJ_STEP = 0, GADJ_TERM = 1 };

template <typename R>
struct LegAdjArgs {
  int M, N, J, B;
  int mode, diffflag;
  const R* Pf;
  const R* Hf;
  const R* Pi;
  const R* Hi;
  const double* coef;
  int ncoef;
  const double* f2;
  double* Lam;        // [B][3][M+1][N+1] complex: dL/dS
  const double* St;   // [B][3][M+1][N+1] complex: forward state S_t (mode A)
  R* Tbar;            // [B][M+1][kNT][J] complex slot cotangents (written by mode A)
  R* G;               // [B][kNG][J][M+1] complex Fourier rows of S_t (written by mode A)
  const R* Gbar;      // [B][kNG][J][M+1] complex Fourier cotangents (modes B, TERM)
  // Fragment-ordered basis tables of the matrix-core kernels (LegArgs; FP32 plans).
  const float* fwdP;
  const float* fwdH;
  const float* invP;
  const float* invH;
};

template <typename R>
struct GridAdjArgs {
  int M, N, J, I, B;
  int mode, forcflag, test;
  const R* tw;
  const double* geo;
  const double* coef;
  int ncoef;
  const R* G;         // recomputed Fourier rows of S_t
  const R* Tbar;      // slot cotangents
  R* Gbar;            // out: Fourier cotangents
  const double* uvic; // pinned winds (test 1)
  double* pgrad;      // [B][J][3] accumulated d/d(taurad, taudrag, DPhieq)
  const double* g_eta;  // GADJ_TERM grid cotangents [B][J][I] (any may be null)
  const double* g_delta;
  const double* g_Phi;
  const double* g_U;
  const double* g_V;
  void* stc;            // folded MFMA adjoint at T106: cache of S_t's even / odd sums (FP32)
};

template <typename R>
hipError_t launch_leg_adj(const LegAdjArgs<R>& a, hipStream_t s);
template <typename R>
hipError_t launch_grid_adj(const GridAdjArgs<R>& a, hipStream_t s);
hipError_t launch_pgrad_reduce(const double* pg, double* out, int B, int J, hipStream_t s);
// acc[B][2] += the diffusion partials (dL/dK6, dL/dK6Phi) of the pair lam = dL/dS, S (both
// [B][3][M+1][N+1] complex128) under the coefficient table coef [B][ncoef].
hipError_t launch_diffusion_partial(const double* Lam, const double* S, const double* coef, int ncoef, int M, int N,
                                    int J, int B, double* acc, hipStream_t s);
// MFMA grid adjoint (GADJ_STEP, FP32; the MFMA grid kernel's DFT tables), swx_adjoint_mfma.hip.
// grid_adj_mfma_lds_bytes is 0 where it is not provided (T106).
hipError_t launch_grid_adj_mfma(const GridAdjArgs<float>& a, const float* TinvF, const float* TfwdF, hipStream_t s);
size_t grid_adj_mfma_lds_bytes(int M);
// Its half-period folded variant (the tables of the folded forward grid kernel).  At T106 it needs
// GridAdjArgs::stc of grid_adj_mfma_fold_cache_bytes (0 elsewhere).
hipError_t launch_grid_adj_mfma_fold(const GridAdjArgs<float>& a, const float* TinvH, const float* TfwdH,
                                     hipStream_t s);
size_t grid_adj_mfma_fold_cache_bytes(int M, int B, int J);
size_t grid_adj_mfma_fold_lds_bytes(int M);
// Matrix-core Legendre adjoint (all modes, FP32; swx_adjoint_leg_mfma.hip) and its largest
// per-workgroup LDS.
hipError_t launch_leg_adj_mfma(const LegAdjArgs<float>& a, hipStream_t s);
size_t leg_adj_mfma_lds_bytes(int M);

}  // namespace swx
