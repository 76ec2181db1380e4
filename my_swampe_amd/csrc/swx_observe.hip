// Phase-curve observable of the geopotential grid and its vector-Jacobian product.
//
//   flux[b][k] = sum_j sqrt(max(1 - mu_j^2, 0)) sum_i Phi[b][j][i] max(cos(lambda_i - phase_k), 0)
//   lambda_i   = -pi + i (2 pi / I)            (numpy linspace(-pi, pi, I, endpoint=False))
//   flux[b][k] /= mean_k |flux[b][k]|         (normalize != 0 and the mean > 0)
//
// Reference: unit_tests/test_parity_reference_regression.py:67-82 (_phase_curve_from_phi), the
// disc-integrated, visibility-weighted geopotential that the retrieval fits to a phase curve
// (retrieval/scripts/pipeline.py:870-925 feeds the same terminal or per-time Phi grids to its
// light-curve stage).
//
// One workgroup per member.  The only HBM traffic is the member's J x I grid (read once, rows
// coalesced) in the forward kernel and the J x I cotangent (written once) in the VJP: the
// projection factors as flux_k = sum_i c_i v_ik with the area-weighted column sums
// c_i = sum_j a_j Phi_ji, so the J-I-K triple sum becomes a J-I pass plus an I-K pass in LDS.
// In the I-K pass each wave takes one output (a phase, or in the VJP a longitude) at a time and
// its 64 lanes split the inner index, finishing with a cross-lane shuffle reduction, so any K
// from 1 to kPhaseCurveMaxK keeps the lanes busy.
#include "swx_observe.h"

#include <hip/hip_runtime.h>

// The reference forms 1 - mu*mu, i*step + start and the products of its sums with separate
// roundings; no fused multiply-adds in this file.
#pragma clang fp contract(off)

namespace swx {

namespace {

constexpr int kPcThreads = 256;
constexpr int kPcWaves = kPcThreads / 64;
constexpr double kPcPi = 3.14159265358979323846;

__device__ inline double pc_lambda(int i, int I) { return (double)i * ((2.0 * kPcPi) / I) + (-kPcPi); }

__device__ inline double pc_area(double mu) { return sqrt(fmax(1.0 - mu * mu, 0.0)); }

__device__ inline double pc_vis(double lam, double phase) {
  const double v = cos(lam - phase);
  return v > 0.0 ? v : 0.0;
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sum of one double per thread; every thread gets the total.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double tot = 0.0;
  for (int q = 0; q < kPcWaves; ++q) tot += red[q];
  __syncthreads();
  return tot;
}

// LDS: aw[J] lam[I] part[max(I, 256)] col[I] fl[K]
__global__ __launch_bounds__(kPcThreads) void k_phase_curve(const double* __restrict__ Phi, long long mstride,
                                                            int J, int I, const double* __restrict__ mus,
                                                            const double* __restrict__ phases, int K, int normalize,
                                                            double* __restrict__ flux,
                                                            double* __restrict__ flux_raw) {
  extern __shared__ double sm[];
  __shared__ double red[kPcWaves];
  double* aw = sm;
  double* lam = aw + J;
  double* part = lam + I;
  double* col = part + (I > kPcThreads ? I : kPcThreads);
  double* fl = col + I;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long b = blockIdx.x;

  for (int j = t; j < J; j += kPcThreads) aw[j] = pc_area(mus[j]);
  for (int i = t; i < I; i += kPcThreads) lam[i] = pc_lambda(i, I);
  __syncthreads();

  // Column sums: P threads share a longitude (rows j = q, q + P, ...), partials in LDS.
  const double* ph = Phi + b * mstride;
  const int P = I >= kPcThreads ? 1 : kPcThreads / I;
  for (int idx = t; idx < P * I; idx += kPcThreads) {
    const int i = idx % I, q = idx / I;
    double acc = 0.0;
    for (int j = q; j < J; j += P) acc += aw[j] * ph[(long long)j * I + i];
    part[idx] = acc;
  }
  __syncthreads();
  for (int i = t; i < I; i += kPcThreads) {
    double s = part[i];
    for (int q = 1; q < P; ++q) s += part[q * I + i];
    col[i] = s;
  }
  __syncthreads();

  for (int k = w; k < K; k += kPcWaves) {
    const double phk = phases[k];
    double acc = 0.0;
    for (int i = lane; i < I; i += 64) acc += col[i] * pc_vis(lam[i], phk);
    acc = wave_sum(acc);
    if (lane == 0) fl[k] = acc;
  }
  __syncthreads();

  double sa = 0.0;
  for (int k = t; k < K; k += kPcThreads) sa += fabs(fl[k]);
  const double norm = block_sum(sa, red) / K;
  const bool scale = normalize && norm > 0.0;
  for (int k = t; k < K; k += kPcThreads) {
    const double f = fl[k];
    if (flux_raw) flux_raw[b * K + k] = f;
    flux[b * K + k] = scale ? f / norm : f;
  }
}

// LDS: aw[J] lam[I] gc[I] gf[K] phs[K]
//
// With n = mean_k |f_k| and the normalized flux f_k / n, the cotangent of the raw flux is
//   gf_k = g_k / n - sign(f_k) (sum_l g_l f_l) / (K n^2),
// and the cotangent of the grid is g_Phi[j][i] = a_j sum_k gf_k v_ik.
__global__ __launch_bounds__(kPcThreads) void k_phase_curve_vjp(const double* __restrict__ g_flux,
                                                                const double* __restrict__ flux_raw, int J, int I,
                                                                const double* __restrict__ mus,
                                                                const double* __restrict__ phases, int K,
                                                                int normalize, double* __restrict__ g_Phi) {
  extern __shared__ double sm[];
  __shared__ double red[kPcWaves];
  double* aw = sm;
  double* lam = aw + J;
  double* gc = lam + I;
  double* gf = gc + I;
  double* phs = gf + K;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long b = blockIdx.x;

  for (int j = t; j < J; j += kPcThreads) aw[j] = pc_area(mus[j]);
  for (int i = t; i < I; i += kPcThreads) lam[i] = pc_lambda(i, I);
  double sa = 0.0, sg = 0.0;
  for (int k = t; k < K; k += kPcThreads) {
    const double f = flux_raw[b * K + k], g = g_flux[b * K + k];
    sa += fabs(f);
    sg += g * f;
    gf[k] = g;
    phs[k] = phases[k];
  }
  const double norm = block_sum(sa, red) / K;
  const double gdotf = block_sum(sg, red);
  if (normalize && norm > 0.0) {
    const double c2 = gdotf / ((double)K * norm * norm);
    for (int k = t; k < K; k += kPcThreads) {
      const double f = flux_raw[b * K + k];
      const double sgn = f > 0.0 ? 1.0 : (f < 0.0 ? -1.0 : 0.0);
      gf[k] = gf[k] / norm - sgn * c2;
    }
  }
  __syncthreads();

  for (int i = w; i < I; i += kPcWaves) {
    const double li = lam[i];
    double acc = 0.0;
    for (int k = lane; k < K; k += 64) acc += gf[k] * pc_vis(li, phs[k]);
    acc = wave_sum(acc);
    if (lane == 0) gc[i] = acc;
  }
  __syncthreads();

  double* out = g_Phi + b * (long long)J * I;
  for (int idx = t; idx < J * I; idx += kPcThreads) {
    const int j = idx / I, i = idx - j * I;
    out[idx] = aw[j] * gc[i];
  }
}

}  // namespace

size_t phase_curve_lds_bytes(int J, int I, int K) {
  return sizeof(double) * ((size_t)J + I + (I > kPcThreads ? I : kPcThreads) + I + K);
}

size_t phase_curve_vjp_lds_bytes(int J, int I, int K) { return sizeof(double) * ((size_t)J + 2 * I + 2 * K); }

hipError_t launch_phase_curve(const double* Phi, long long mstride, int B, int J, int I, const double* mus,
                              const double* phases, int K, int normalize, double* flux, double* flux_raw,
                              hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_phase_curve, dim3(B), dim3(kPcThreads), phase_curve_lds_bytes(J, I, K), s, Phi, mstride, J,
                     I, mus, phases, K, normalize, flux, flux_raw);
  return hipGetLastError();
}

hipError_t launch_phase_curve_vjp(const double* g_flux, const double* flux_raw, int B, int J, int I,
                                  const double* mus, const double* phases, int K, int normalize, double* g_Phi,
                                  hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_phase_curve_vjp, dim3(B), dim3(kPcThreads), phase_curve_vjp_lds_bytes(J, I, K), s, g_flux,
                     flux_raw, J, I, mus, phases, K, normalize, g_Phi);
  return hipGetLastError();
}

}  // namespace swx
