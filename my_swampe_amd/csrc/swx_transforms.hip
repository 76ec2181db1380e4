// Standalone FP64 transforms on the GPU, with the reference's spectral_transform interface
// (spectral_transform.py:290-650): the truncated forward longitude DFT (fwd_fft_trunc), the
// inverse DFT (invrs_fft), the forward and inverse Legendre transforms (fwd_leg, fwd_leg_w,
// invrs_leg), the wind inversion (invrsUV, invrsUV_with_coeffs) and the diagnostic
// vorticity/divergence transform (diagnostic_eta_delta) with the spectral-to-grid synthesis it
// ends with.  The reference uses them outside the time step (its transform round-trip tests,
// unit_tests/test_transform_stack.py:51-180); here they run on the caller's arrays (basis,
// weights, coefficient arrays passed in).  Longitude work runs one workgroup per (latitude row,
// member) with the row and a twiddle table e^{+-2 pi i r / I} in LDS; latitude sums run one
// workgroup per (m, member) with the n index across lanes, so every basis read is coalesced.
#include "swx_transforms.h"

namespace swx {

namespace {

constexpr int kSpecThreads = 64;
constexpr int kGridThreads = 256;
constexpr int kDftThreads = 128;

// tw[r] = (cos, sign * sin)(2 pi r / I), r in [0, I).  The argument 2 r / I is the one the
// direct synthesis would pass to sincospi for m i = r (mod I), so the table is bit-identical to
// evaluating each term.
__device__ void stage_twiddles(double* tw, int I, double sign) {
  for (int r = threadIdx.x; r < I; r += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * (double)r / (double)I, &sn, &cs);
    tw[2 * r] = cs;
    tw[2 * r + 1] = sign * sn;
  }
}

// One latitude row of the grid of Hermitian Fourier coefficients X (LDS, [m][re, im],
// m = 0..M): the reference layout of invrs_leg (+m at column m, conj at column I - m) under
// ifft(I x) is x(i) = X_0 + sum_{m>=1} (X_m e^{2 pi i m i / I} + conj(X_m e^{2 pi i m i / I})),
// real part Re X_0 + 2 sum Re(X_m e^{...}), imaginary part Im X_0.  tw holds e^{+2 pi i r / I}.
__device__ void synth_hermitian_row(const double* X, const double* tw, int Mp, int I, double* out) {
  for (int i = threadIdx.x; i < I; i += blockDim.x) {
    double re = 0.0;
    int r = 0;
    for (int m = 1; m < Mp; ++m) {
      r += i;
      if (r >= I) r -= I;
      re += X[m * 2] * tw[2 * r] - X[m * 2 + 1] * tw[2 * r + 1];
    }
    out[i * 2] = X[0] + 2.0 * re;
    out[i * 2 + 1] = X[1];
  }
}

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

// Grids of up to two spectra: X_m = sum_n S_mn P_jmn, then the Hermitian synthesis.
__global__ void __launch_bounds__(kGridThreads)
    k_spec_to_grid(const double* __restrict__ S0, const double* __restrict__ S1, int J, int I, int M, int N,
                   const double* __restrict__ Pmn, double* __restrict__ out0, double* __restrict__ out1) {
  extern __shared__ double lds[];  // twiddles [I][cos, sin], then [field][m][re, im]
  double* tw = lds;
  double* sX = lds + 2 * I;
  const int j = blockIdx.x, b = blockIdx.y;
  const int Mp = M + 1, Np = N + 1;
  const int nf = S1 ? 2 : 1;
  stage_twiddles(tw, I, 1.0);
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
  for (int f = 0; f < nf; ++f)
    synth_hermitian_row(sX + f * Mp * 2, tw, Mp, I, (f == 0 ? out0 : out1) + ((long long)b * J + j) * I * 2);
}

// fwd_fft_trunc (spectral_transform.py:290-328): X[k, j, m] = sum_{i < min(C, I)} (x_i / I)
// e^{-2 pi i m i / I} for m < Mo = min(M + 1, I); a row of C < I columns is zero-padded and
// one of C > I columns truncated, as fft(x / I, n=I) does.  x is real or interleaved complex.
__global__ void __launch_bounds__(kDftThreads)
    k_fwd_dft_trunc(const double* __restrict__ x, int is_complex, int J, int C, int I, int Mo,
                    double* __restrict__ out) {
  extern __shared__ double lds[];  // twiddles [I][cos, -sin], then the row [i][re, im] / I
  double* tw = lds;
  double* xs = lds + 2 * I;
  const int j = blockIdx.x, k = blockIdx.y;
  const int Cu = C < I ? C : I;
  const long long row = (long long)k * J + j;
  stage_twiddles(tw, I, -1.0);
  const double invI = 1.0 / (double)I;
  for (int i = threadIdx.x; i < Cu; i += kDftThreads) {
    if (is_complex) {
      xs[2 * i] = x[(row * C + i) * 2] * invI;
      xs[2 * i + 1] = x[(row * C + i) * 2 + 1] * invI;
    } else {
      xs[2 * i] = x[row * C + i] * invI;
      xs[2 * i + 1] = 0.0;
    }
  }
  __syncthreads();
  for (int m = threadIdx.x; m < Mo; m += kDftThreads) {
    double sr = 0.0, si = 0.0;
    int r = 0;  // m i mod I
    for (int i = 0; i < Cu; ++i) {
      const double c = tw[2 * r], s = tw[2 * r + 1];
      sr += xs[2 * i] * c - xs[2 * i + 1] * s;
      si += xs[2 * i] * s + xs[2 * i + 1] * c;
      r += m;
      if (r >= I) r -= I;
    }
    out[(row * Mo + m) * 2] = sr;
    out[(row * Mo + m) * 2 + 1] = si;
  }
}

// invrs_fft (spectral_transform.py:331-348): ifft(I X, n=I) = sum_{c < min(C, I)} X_c
// e^{+2 pi i c i / I}, complex, for a row of C coefficients (zero-padded or truncated to I).
__global__ void __launch_bounds__(kGridThreads)
    k_dft_synth(const double* __restrict__ X, int J, int C, int I, double* __restrict__ out) {
  extern __shared__ double lds[];  // twiddles [I][cos, sin], then the row [c][re, im]
  double* tw = lds;
  double* xs = lds + 2 * I;
  const int j = blockIdx.x, k = blockIdx.y;
  const int Cu = C < I ? C : I;
  const long long row = (long long)k * J + j;
  stage_twiddles(tw, I, 1.0);
  for (int c = threadIdx.x; c < Cu; c += kGridThreads) {
    xs[2 * c] = X[(row * C + c) * 2];
    xs[2 * c + 1] = X[(row * C + c) * 2 + 1];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < I; i += kGridThreads) {
    double sr = 0.0, si = 0.0;
    int r = 0;  // c i mod I
    for (int c = 0; c < Cu; ++c) {
      const double tc = tw[2 * r], ts = tw[2 * r + 1];
      sr += xs[2 * c] * tc - xs[2 * c + 1] * ts;
      si += xs[2 * c] * ts + xs[2 * c + 1] * tc;
      r += i;
      if (r >= I) r -= I;
    }
    out[(row * I + i) * 2] = sr;
    out[(row * I + i) * 2 + 1] = si;
  }
}

// fwd_leg / fwd_leg_w (spectral_transform.py:370-435): S[k, m, n] = sum_j data[k, j, m]
// (w_j B[j, m, n]), with w = null for a basis already weighted (fwd_leg_w).
__global__ void __launch_bounds__(kSpecThreads)
    k_fwd_leg(const double* __restrict__ data, const double* __restrict__ basis, const double* __restrict__ w, int J,
              int M, int N, double* __restrict__ out) {
  const int m = blockIdx.x, k = blockIdx.y;
  const int Mp = M + 1, Np = N + 1;
  const double* D = data + (long long)k * J * Mp * 2;
  for (int n = threadIdx.x; n < Np; n += kSpecThreads) {
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < J; ++j) {
      double b = basis[((long long)j * Mp + m) * Np + n];
      if (w) b = w[j] * b;
      sr += D[(j * Mp + m) * 2] * b;
      si += D[(j * Mp + m) * 2 + 1] * b;
    }
    out[(((long long)k * Mp + m) * Np + n) * 2] = sr;
    out[(((long long)k * Mp + m) * Np + n) * 2 + 1] = si;
  }
}

// invrs_leg (spectral_transform.py:438-481): X_m = sum_n B[j, m, n] S[m, n] at column m
// (m = 0..M); conj(X_m) at column I - m (m = 1..M), which the reference writes last, so it
// wins where the two ranges meet; zero elsewhere.  sum_n B conj(S_n) = conj(sum_n B S_n)
// exactly (negation commutes with rounding), so the conjugate columns reuse X_m.
__global__ void __launch_bounds__(kGridThreads)
    k_invrs_leg(const double* __restrict__ S, const double* __restrict__ basis, int J, int I, int M, int N,
                double* __restrict__ out) {
  extern __shared__ double sX[];  // [m][re, im]
  const int j = blockIdx.x, k = blockIdx.y;
  const int Mp = M + 1, Np = N + 1;
  const double* Sk = S + (long long)k * Mp * Np * 2;
  for (int m = threadIdx.x; m < Mp; m += kGridThreads) {
    double xr = 0.0, xi = 0.0;
    for (int n = 0; n < Np; ++n) {
      const double b = basis[((long long)j * Mp + m) * Np + n];
      xr += b * Sk[(m * Np + n) * 2];
      xi += b * Sk[(m * Np + n) * 2 + 1];
    }
    sX[2 * m] = xr;
    sX[2 * m + 1] = xi;
  }
  __syncthreads();
  double* row = out + ((long long)k * J + j) * I * 2;
  for (int c = threadIdx.x; c < I; c += kGridThreads) {
    double re = 0.0, im = 0.0;
    if (M > 0 && c >= I - M) {
      re = sX[2 * (I - c)];
      im = -sX[2 * (I - c) + 1];
    } else if (c <= M) {
      re = sX[2 * c];
      im = sX[2 * c + 1];
    }
    row[2 * c] = re;
    row[2 * c + 1] = im;
  }
}

// invrsUV / invrsUV_with_coeffs (spectral_transform.py:484-596).  With column n = 0 of delta
// and eta zeroed, ds = delta tc and es = (eta - f) tc (tc = tstepcoeffmn):
//   U_m = -(sum_n P i m ds - sum_n H es),  V_m = -(sum_n P i m es + sum_n H ds)
// (m = marray[m, n]), then U, V = the Hermitian synthesis of U_m, V_m.  Each output may be
// null; Um / Vm are the (J, M+1) positive-m coefficients.
__global__ void __launch_bounds__(kGridThreads)
    k_invrs_uv(const double* __restrict__ deltamn, const double* __restrict__ etamn, const double* __restrict__ fmn,
               int J, int I, int M, int N, const double* __restrict__ Pmn, const double* __restrict__ Hmn,
               const double* __restrict__ tcmn, const double* __restrict__ marray, double* __restrict__ U,
               double* __restrict__ V, double* __restrict__ Um, double* __restrict__ Vm) {
  extern __shared__ double lds[];  // twiddles [I][cos, sin], then [U, V][m][re, im]
  double* tw = lds;
  double* sX = lds + 2 * I;
  const int j = blockIdx.x, k = blockIdx.y;
  const int Mp = M + 1, Np = N + 1;
  const double* D = deltamn + (long long)k * Mp * Np * 2;
  const double* E = etamn + (long long)k * Mp * Np * 2;
  stage_twiddles(tw, I, 1.0);
  for (int e = threadIdx.x; e < 2 * Mp; e += kGridThreads) {
    const int f = e / Mp, m = e - f * Mp;
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;  // the P sum and the H sum
    for (int n = 0; n < Np; ++n) {
      const int o = m * Np + n;
      const double tc = tcmn[o], mr = marray[o];
      const double dr = (n ? D[o * 2] : 0.0) * tc, di = (n ? D[o * 2 + 1] : 0.0) * tc;
      const double er = ((n ? E[o * 2] : 0.0) - fmn[o * 2]) * tc;
      const double ei = ((n ? E[o * 2 + 1] : 0.0) - fmn[o * 2 + 1]) * tc;
      const double p = Pmn[((long long)j * Mp + m) * Np + n], h = Hmn[((long long)j * Mp + m) * Np + n];
      // i (m x) = (-m x_i, m x_r); U pairs P with i m ds and H with es, V pairs P with i m es
      // and H with ds.
      const double qr = f == 0 ? -(mr * di) : -(mr * ei);
      const double qi = f == 0 ? mr * dr : mr * er;
      const double hr = f == 0 ? er : dr, hi = f == 0 ? ei : di;
      ar += p * qr;
      ai += p * qi;
      br += h * hr;
      bi += h * hi;
    }
    const double xr = f == 0 ? -(ar - br) : -(ar + br);
    const double xi = f == 0 ? -(ai - bi) : -(ai + bi);
    sX[(f * Mp + m) * 2] = xr;
    sX[(f * Mp + m) * 2 + 1] = xi;
    double* co = f == 0 ? Um : Vm;
    if (co) {
      co[(((long long)k * J + j) * Mp + m) * 2] = xr;
      co[(((long long)k * J + j) * Mp + m) * 2 + 1] = xi;
    }
  }
  __syncthreads();
  if (U) synth_hermitian_row(sX, tw, Mp, I, U + ((long long)k * J + j) * I * 2);
  if (V) synth_hermitian_row(sX + Mp * 2, tw, Mp, I, V + ((long long)k * J + j) * I * 2);
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
  const size_t lds = sizeof(double) * (2 * (size_t)I + 2 * 2 * (size_t)(M + 1));
  hipLaunchKernelGGL(k_spec_to_grid, dim3(J, B), dim3(kGridThreads), lds, s, S0, S1, J, I, M, N, Pmn, out0, out1);
  return hipGetLastError();
}

hipError_t launch_fwd_dft_trunc(const double* x, int is_complex, int K, int J, int C, int I, int Mo, double* out,
                                hipStream_t s) {
  if (K <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * 2 * ((size_t)I + (size_t)(C < I ? C : I));
  hipLaunchKernelGGL(k_fwd_dft_trunc, dim3(J, K), dim3(kDftThreads), lds, s, x, is_complex, J, C, I, Mo, out);
  return hipGetLastError();
}

hipError_t launch_dft_synth(const double* X, int K, int J, int C, int I, double* out, hipStream_t s) {
  if (K <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * 2 * ((size_t)I + (size_t)(C < I ? C : I));
  hipLaunchKernelGGL(k_dft_synth, dim3(J, K), dim3(kGridThreads), lds, s, X, J, C, I, out);
  return hipGetLastError();
}

hipError_t launch_fwd_leg(const double* data, const double* basis, const double* w, int K, int J, int M, int N,
                          double* out, hipStream_t s) {
  if (K <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fwd_leg, dim3(M + 1, K), dim3(kSpecThreads), 0, s, data, basis, w, J, M, N, out);
  return hipGetLastError();
}

hipError_t launch_invrs_leg(const double* S, const double* basis, int K, int J, int I, int M, int N, double* out,
                            hipStream_t s) {
  if (K <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * 2 * (size_t)(M + 1);
  hipLaunchKernelGGL(k_invrs_leg, dim3(J, K), dim3(kGridThreads), lds, s, S, basis, J, I, M, N, out);
  return hipGetLastError();
}

hipError_t launch_invrs_uv(const double* deltamn, const double* etamn, const double* fmn, int K, int J, int I, int M,
                           int N, const double* Pmn, const double* Hmn, const double* tstepcoeffmn,
                           const double* marray, double* U, double* V, double* Um, double* Vm, hipStream_t s) {
  if (K <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (2 * (size_t)I + 2 * 2 * (size_t)(M + 1));
  hipLaunchKernelGGL(k_invrs_uv, dim3(J, K), dim3(kGridThreads), lds, s, deltamn, etamn, fmn, J, I, M, N, Pmn, Hmn,
                     tstepcoeffmn, marray, U, V, Um, Vm);
  return hipGetLastError();
}

}  // namespace swx
