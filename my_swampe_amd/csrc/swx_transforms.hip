// Standalone FP64 transforms on the GPU: the diagnostic vorticity/divergence transform of the
// reference (spectral_transform.py:599-650, diagnostic_eta_delta) and the spectral-to-grid
// synthesis it ends with (invrs_leg then invrs_fft, spectral_transform.py:438-481, :331-348).
// The reference uses them outside the time step (its transform round-trip tests,
// unit_tests/test_transform_stack.py:100-180); here they run on the reference's own arrays
// (basis, weights, coefficient arrays passed in), one workgroup per (m, member) for the
// latitude sums and one per (latitude, member) for the synthesis.
#include "swx_transforms.h"

namespace swx {

namespace {

constexpr int kSpecThreads = 64;
constexpr int kGridThreads = 256;

__global__ void __launch_bounds__(kSpecThreads)
    k_diag_eta_delta_spec(const double* __restrict__ Um, const double* __restrict__ Vm,
                          const double* __restrict__ fmn, int J, int M, int N, const double* __restrict__ Pmn,
                          const double* __restrict__ Hmn, const double* __restrict__ w,
                          const double* __restrict__ tstepcoeff, const double* __restrict__ mJarray, double dt,
                          double* __restrict__ etamn, double* __restrict__ deltamn) {
  const int m = blockIdx.x, b = blockIdx.y;
  const int Mp = M + 1, Np = N + 1;
  const double* U = Um + (long long)b * J * Mp * 2;
  const double* V = Vm + (long long)b * J * Mp * 2;
  for (int n = threadIdx.x; n < Np; n += kSpecThreads) {
    double zr = 0.0, zi = 0.0, dr = 0.0, di = 0.0;
    for (int j = 0; j < J; ++j) {
      const double c = tstepcoeff[j * Mp + m] / (2.0 * dt) * w[j];
      const double mj = mJarray[j * Mp + m];
      const double p = Pmn[((long long)j * Mp + m) * Np + n], h = Hmn[((long long)j * Mp + m) * Np + n];
      const double ur = U[(j * Mp + m) * 2], ui = U[(j * Mp + m) * 2 + 1];
      const double vr = V[(j * Mp + m) * 2], vi = V[(j * Mp + m) * 2 + 1];
      // i m V P + U H  and  i m U P - V H  (i (x + i y) = -y + i x)
      zr += c * (-mj * vi * p + ur * h);
      zi += c * (mj * vr * p + ui * h);
      dr += c * (-mj * ui * p - vr * h);
      di += c * (mj * ur * p - vi * h);
    }
    const long long o = (((long long)b * Mp + m) * Np + n) * 2;
    etamn[o] = zr + fmn[(m * Np + n) * 2];
    etamn[o + 1] = zi + fmn[(m * Np + n) * 2 + 1];
    deltamn[o] = dr;
    deltamn[o + 1] = di;
  }
}

// x(i) = X_0 + sum_{m=1..M} (X_m e^{2 pi i m i / I} + conj(X_m) e^{-2 pi i m i / I}),
// X_m = sum_n S_mn P_jmn: the Hermitian layout of invrs_leg (+m at m, conj at I - m) under
// ifft(I x).  Real part Re X_0 + 2 sum Re(X_m e^{...}); imaginary part Im X_0.
__global__ void __launch_bounds__(kGridThreads)
    k_spec_to_grid(const double* __restrict__ S0, const double* __restrict__ S1, int J, int I, int M, int N,
                   const double* __restrict__ Pmn, double* __restrict__ out0, double* __restrict__ out1) {
  extern __shared__ double sX[];  // [field][m][re, im]
  const int j = blockIdx.x, b = blockIdx.y;
  const int Mp = M + 1, Np = N + 1;
  const int nf = S1 ? 2 : 1;
  for (int e = threadIdx.x; e < nf * Mp; e += kGridThreads) {
    const int f = e / Mp, m = e - f * Mp;
    const double* S = (f == 0 ? S0 : S1) + (long long)b * Mp * Np * 2;
    double xr = 0.0, xi = 0.0;
    for (int n = 0; n < Np; ++n) {
      const double p = Pmn[((long long)j * Mp + m) * Np + n];
      xr += S[(m * Np + n) * 2] * p;
      xi += S[(m * Np + n) * 2 + 1] * p;
    }
    sX[(f * Mp + m) * 2] = xr;
    sX[(f * Mp + m) * 2 + 1] = xi;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < I; i += kGridThreads) {
    for (int f = 0; f < nf; ++f) {
      const double* X = sX + f * Mp * 2;
      double re = 0.0;
      for (int m = 1; m < Mp; ++m) {
        double sn, cs;
        sincospi(2.0 * (double)(((long long)m * i) % I) / (double)I, &sn, &cs);
        re += X[m * 2] * cs - X[m * 2 + 1] * sn;
      }
      double* out = (f == 0 ? out0 : out1) + (((long long)b * J + j) * I + i) * 2;
      out[0] = X[0] + 2.0 * re;
      out[1] = X[1];
    }
  }
}

}  // namespace

hipError_t launch_diag_eta_delta_spec(const double* Um, const double* Vm, const double* fmn, int B, int J, int M,
                                      int N, const double* Pmn, const double* Hmn, const double* w,
                                      const double* tstepcoeff, const double* mJarray, double dt, double* etamn,
                                      double* deltamn, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_diag_eta_delta_spec, dim3(M + 1, B), dim3(kSpecThreads), 0, s, Um, Vm, fmn, J, M, N, Pmn, Hmn,
                     w, tstepcoeff, mJarray, dt, etamn, deltamn);
  return hipGetLastError();
}

hipError_t launch_spec_to_grid(const double* S0, const double* S1, int B, int J, int I, int M, int N,
                               const double* Pmn, double* out0, double* out1, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * 2 * 2 * (size_t)(M + 1);
  hipLaunchKernelGGL(k_spec_to_grid, dim3(J, B), dim3(kGridThreads), lds, s, S0, S1, J, I, M, N, Pmn, out0, out1);
  return hipGetLastError();
}

}  // namespace swx
