// Per-step kernels of the spectral shallow-water core (plain path, FP64 or
// FP32 transforms; the spectral state update and diffusion are FP64).
//
// One step is two launches:
//   k_leg  (grid = (M+1, members)):  forward Legendre contraction of the
//          Fourier tendency inputs for zonal wavenumber m, FP64 update of the
//          spectral state (modified-Euler or explicit, diffusion), then the
//          inverse Legendre contraction with the wind inversion fused in.
//          Reference: modEuler_tdiff.py:30-268, explicit_tdiff.py:20-233,
//          spectral_transform.py:370-596, filters.py:11-27.
//   k_grid (grid = (J, members)):  inverse longitude DFT of one latitude row,
//          grid-space physics (Robert-Asselin filter, forcing Q/F/G, nonlinear
//          products A..E, diagnostics), forward truncated DFT and assembly of
//          the w-weighted Legendre inputs for the next step.
//          Reference: spectral_transform.py:290-348, forcing.py:14-155,
//          initial_conditions.py:242-292, model.py:795-1194.
//
// The linear carry terms (Pw(etam1), Pw(deltam1), Pw(Phim1), Pw(*m0)) of the
// reference are applied directly to the FP64 spectral state: Gauss quadrature
// makes Pw(invrs_leg(X)) == X exactly, so only the nonlinear and forcing
// tendencies pass through the transforms.
#include "swx_kernels.h"

#include <cfloat>

#include "swx_physics.h"

namespace swx {

namespace {

__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }

// Block-wide reduction of NV doubles (sum / min / max per op code) through LDS.
// op: 0 = sum, 1 = min, 2 = max.  All threads must call it.
template <int NV>
__device__ void block_reduce(double (&v)[NV], const int (&op)[NV], double* scratch) {
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  for (int k = 0; k < NV; ++k) scratch[k * nt + tid] = v[k];
  __syncthreads();
  for (int s = nt / 2; s > 0; s >>= 1) {
    if (tid < s) {
      for (int k = 0; k < NV; ++k) {
        double a = scratch[k * nt + tid], b = scratch[k * nt + tid + s];
        scratch[k * nt + tid] = op[k] == 0 ? a + b : (op[k] == 1 ? dmin(a, b) : dmax(a, b));
      }
    }
    __syncthreads();
  }
  for (int k = 0; k < NV; ++k) v[k] = scratch[k * nt];
  __syncthreads();
}

}  // namespace

// ----------------------------------------------------------------------------
// Legendre kernel
// ----------------------------------------------------------------------------

template <typename R>
__global__ void __launch_bounds__(128) k_leg(LegArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int m, b;
  leg_item(1, m, b);
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int M = a.M, N = a.N, J = a.J;
  const int K = N + 1 - m;  // retained degrees n = m..N
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);

  R* sT = reinterpret_cast<R*>(smem_raw);         // [kNT][J][2]
  R* sX = sT + kNT * J * 2;                       // [5][N+1][2]
  double* sRed = reinterpret_cast<double*>(sX + 5 * (N + 1) * 2);

  if (a.mode == LEG_STEP && a.diagnostics) {
    // Blow-up gate on the current winds (model.py:838-845, :1192-1194).
    double v[2] = {0.0, DBL_MAX};
    const int op[2] = {0, 1};
    for (int j = tid; j < J; j += nt) {
      v[0] += a.part_uv[((long long)b * J + j) * 2 + 0];
      v[1] = dmin(v[1], a.part_uv[((long long)b * J + j) * 2 + 1]);
    }
    block_reduce<2>(v, op, sRed);
    const double rms = sqrt(v[0]);
    const bool dead_next = (a.dead[b] != 0) || (rms > cf[C_BLOWUP]);
    if (m == 0 && tid == 0) {
      a.dead[b] = dead_next ? 1 : 0;
      if (a.hist_scal) {
        double* h = a.hist_scal + ((long long)b * a.hist_T + a.hist_k) * 4;
        h[0] = rms;
        h[1] = v[1];
      }
      if (a.hist_dead) a.hist_dead[(long long)b * a.hist_T + a.hist_k] = dead_next ? 1 : 0;
    }
    if (dead_next) return;
  }
  if (a.mode == LEG_STEP && (a.si || a.raw_filter) && m == 0 && tid == 0) residue_step(cf, a.flags(), a.resid_c + 2 * b);

  // Stage the Legendre inputs of (b, m).
  const long long toff = (((long long)b * (M + 1) + m) * kNT) * J * 2;
  const int nslots = (a.mode == LEG_STEP) ? kNT : (a.mode == LEG_INIT ? 3 : 0);
  for (int e = tid; e < nslots * J * 2; e += nt) sT[e] = a.T[toff + e];
  __syncthreads();

  const long long boff = tri_off(m, N, J);
  const R* Pf = a.Pf + boff;
  const R* Hf = a.Hf + boff;
  double* S = a.S + (long long)b * 2 * 3 * (M + 1) * (N + 1) * 2;
  double* Sprev = S;
  double* Scur = S + 3 * (M + 1) * (N + 1) * 2;
  const long long fstride = (long long)(M + 1) * (N + 1) * 2;

  for (int np = tid; np < K; np += nt) {
    const int n = m + np;
    const long long so = ((long long)m * (N + 1) + n) * 2;
    double eta_n[2], delta_n[2], Phi_n[2];
    // Wind potentials come from the new level (differs from the stored one under RAW).
    double eta_w[2], delta_w[2];
    if (a.mode == LEG_PRIME) {
      // Fourier rows of a stored level (restart from a saved state, exports).
      const double* Sl = a.prime_prev ? Sprev : Scur;
      for (int c = 0; c < 2; ++c) {
        eta_n[c] = Sl[0 * fstride + so + c];
        delta_n[c] = Sl[1 * fstride + so + c];
        Phi_n[c] = Sl[2 * fstride + so + c];
      }
    } else if (a.mode == LEG_INIT) {
      R acc[6] = {0, 0, 0, 0, 0, 0};
      for (int j = 0; j < J; ++j) {
        const R p = Pf[(long long)j * K + np];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          acc[2 * k] += sT[(k * J + j) * 2] * p;
          acc[2 * k + 1] += sT[(k * J + j) * 2 + 1] * p;
        }
      }
      eta_n[0] = acc[0]; eta_n[1] = acc[1];
      delta_n[0] = acc[2]; delta_n[1] = acc[3];
      Phi_n[0] = acc[4]; Phi_n[1] = acc[5];
      // The stored level is divided by f2 so that the carry f2 * S of the
      // first step equals Pw(DFT(field0)) exactly as in the reference.
      const double f2 = a.f2[(long long)m * (N + 1) + n];
      const double inv = f2 == 0.0 ? 0.0 : 1.0 / f2;
      for (int c = 0; c < 2; ++c) {
        Sprev[0 * fstride + so + c] = Scur[0 * fstride + so + c] = eta_n[c] * inv;
        Sprev[1 * fstride + so + c] = Scur[1 * fstride + so + c] = delta_n[c] * inv;
        Sprev[2 * fstride + so + c] = Scur[2 * fstride + so + c] = Phi_n[c] * inv;
      }
    } else {
      R acc[2 * kNT];
#pragma unroll
      for (int k = 0; k < 2 * kNT; ++k) acc[k] = 0;
      for (int j = 0; j < J; ++j) {
        const R p = Pf[(long long)j * K + np];
        const R h = Hf[(long long)j * K + np];
#pragma unroll
        for (int k = 0; k < kNP; ++k) {
          acc[2 * k] += sT[(k * J + j) * 2] * p;
          acc[2 * k + 1] += sT[(k * J + j) * 2 + 1] * p;
        }
#pragma unroll
        for (int k = kNP; k < kNT; ++k) {
          acc[2 * k] += sT[(k * J + j) * 2] * h;
          acc[2 * k + 1] += sT[(k * J + j) * 2 + 1] * h;
        }
      }
      double fp[5][2], fh[4][2];
      for (int k = 0; k < 5; ++k) {
        fp[k][0] = acc[2 * k];
        fp[k][1] = acc[2 * k + 1];
      }
      for (int k = 0; k < 4; ++k) {
        fh[k][0] = acc[2 * (kNP + k)];
        fh[k][1] = acc[2 * (kNP + k) + 1];
      }
      double eff[3][2], nw[3][2];
      spectral_update(cf, cl, a.f2[(long long)m * (N + 1) + n], m, n, a.flags(), fp, fh, Scur, Sprev, so, fstride,
                      eff, nw);
      for (int c = 0; c < 2; ++c) {
        eta_n[c] = eff[0][c];
        delta_n[c] = eff[1][c];
        Phi_n[c] = eff[2][c];
        eta_w[c] = nw[0][c];
        delta_w[c] = nw[1][c];
      }
    }
    if (a.mode != LEG_STEP)
      for (int c = 0; c < 2; ++c) {
        eta_w[c] = eta_n[c];
        delta_w[c] = delta_n[c];
      }
    // Inputs of the inverse transform: eta, delta, Phi and the wind potentials.
    double dh[2], eh[2];
    wind_potentials(cf, cl, m, n, eta_w, delta_w, dh, eh);
    for (int c = 0; c < 2; ++c) {
      sX[(0 * (N + 1) + np) * 2 + c] = (R)eta_n[c];
      sX[(1 * (N + 1) + np) * 2 + c] = (R)delta_n[c];
      sX[(2 * (N + 1) + np) * 2 + c] = (R)Phi_n[c];
      sX[(3 * (N + 1) + np) * 2 + c] = (R)dh[c];
      sX[(4 * (N + 1) + np) * 2 + c] = (R)eh[c];
    }
  }
  if (a.mode == LEG_INIT && !a.diag_winds) return;
  __syncthreads();

  const R* Pi = a.Pi + boff;
  const R* Hi = a.Hi + boff;
  R* G = a.G + fidx(b, 0, m, 0, M + 1, J);
  const long long fs = (long long)(M + 1) * J * 2;  // field stride of the m-major rows
  const R mr = (R)m;
  for (int j = tid; j < J; j += nt) {
    R p[10];
    R h[4];
#pragma unroll
    for (int k = 0; k < 10; ++k) p[k] = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = 0;
    for (int np = 0; np < K; ++np) {
      const R pv = Pi[(long long)np * J + j];
      const R hv = Hi[(long long)np * J + j];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        p[2 * k] += pv * sX[(k * (N + 1) + np) * 2];
        p[2 * k + 1] += pv * sX[(k * (N + 1) + np) * 2 + 1];
      }
      // H applied to eta_hat (slot 4) and delta_hat (slot 3).
      h[0] += hv * sX[(4 * (N + 1) + np) * 2];
      h[1] += hv * sX[(4 * (N + 1) + np) * 2 + 1];
      h[2] += hv * sX[(3 * (N + 1) + np) * 2];
      h[3] += hv * sX[(3 * (N + 1) + np) * 2 + 1];
    }
    const long long ro = (long long)j * 2;
    if (a.mode != LEG_INIT) {
      G[0 * fs + ro] = p[0];
      G[0 * fs + ro + 1] = p[1];
      G[1 * fs + ro] = p[2];
      G[1 * fs + ro + 1] = p[3];
      G[2 * fs + ro] = p[4];
      G[2 * fs + ro + 1] = p[5];
    }
    // U_m = -i m P(delta_hat) + H(eta_hat);  V_m = -i m P(eta_hat) - H(delta_hat)
    // (spectral_transform.py:584-594).
    G[3 * fs + ro] = mr * p[7] + h[0];
    G[3 * fs + ro + 1] = -mr * p[6] + h[1];
    G[4 * fs + ro] = mr * p[9] - h[2];
    G[4 * fs + ro + 1] = -mr * p[8] - h[3];
  }
}

// Blow-up gate of a diagnostics step ahead of the matrix-core Legendre kernel, one block per
// member: the same reduction of the grid kernel's row partials as k_leg's (model.py:838-845,
// :1192-1194), dead flag, per-step rms / spin_min and dead history.  The Legendre kernel then
// leaves dead members' carry as it is.
__global__ void __launch_bounds__(128) k_blowup_gate(const double* part_uv, unsigned char* dead, const double* coef,
                                                     int ncoef, int J, double* hist_scal, unsigned char* hist_dead,
                                                     int hist_T, int hist_k) {
  __shared__ double sRed[2 * 128];
  const int b = blockIdx.x, tid = threadIdx.x;
  double v[2] = {0.0, DBL_MAX};
  const int op[2] = {0, 1};
  for (int j = tid; j < J; j += blockDim.x) {
    v[0] += part_uv[((long long)b * J + j) * 2 + 0];
    v[1] = dmin(v[1], part_uv[((long long)b * J + j) * 2 + 1]);
  }
  block_reduce<2>(v, op, sRed);
  if (tid == 0) {
    const double rms = sqrt(v[0]);
    const bool dead_next = (dead[b] != 0) || (rms > coef[(long long)b * ncoef + C_BLOWUP]);
    dead[b] = dead_next ? 1 : 0;
    if (hist_scal) {
      double* h = hist_scal + ((long long)b * hist_T + hist_k) * 4;
      h[0] = rms;
      h[1] = v[1];
    }
    if (hist_dead) hist_dead[(long long)b * hist_T + hist_k] = dead_next ? 1 : 0;
  }
}

template <typename R>
hipError_t launch_blowup_gate(const LegArgs<R>& a, hipStream_t s) {
  hipLaunchKernelGGL(k_blowup_gate, dim3(a.B), dim3(128), 0, s, a.part_uv, a.dead, a.coef, a.ncoef, a.J, a.hist_scal,
                     a.hist_dead, a.hist_T, a.hist_k);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Grid kernel
// ----------------------------------------------------------------------------

// Inverse DFT of Hermitian-symmetric rows at longitude i: x_f = Re c_f0 + 2 sum_{m >= 1}
// Re(c_fm e^{+2 pi i m i / I}) (spectral_transform.py:331-348 with the real part taken,
// model.py:973-977) for the fields f0 <= f < NF of sG ([f][Mp][2]); x_f of f < f0 is left
// unset.  k_grid and k_phys_synth share it, so a row inverted by either rounds the same.
template <typename R, int NF>
__device__ __forceinline__ void inv_dft_point(const R* sG, const R* sTc, const R* sTs, int Mp, int I, int i, int f0,
                                              double x[NF]) {
  R acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = 0;
  int idx = 0;
  for (int m = 1; m < Mp; ++m) {
    idx += i;
    if (idx >= I) idx -= I;
    const R c = sTc[idx], s = sTs[idx];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (f < f0) continue;
      acc[f] += sG[(f * Mp + m) * 2] * c - sG[(f * Mp + m) * 2 + 1] * s;
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f)
    if (f >= f0) x[f] = (double)sG[(f * Mp) * 2] + 2.0 * (double)acc[f];
}

template <typename R>
__global__ void __launch_bounds__(256) k_grid(GridArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int j = blockIdx.x;
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int M = a.M, N = a.N, J = a.J, I = a.I;
  const int Mp = M + 1;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;

  constexpr int kMaxF = 10;  // forward-DFT fields per row
  R* sTc = reinterpret_cast<R*>(smem_raw);  // [I] cos table
  R* sTs = sTc + I;                          // [I] sin table
  R* sG = sTs + I;                           // [kNG][Mp][2]
  R* sX = sG + kNG * Mp * 2;                 // [kMaxF][I]
  R* sF = sX + kMaxF * I;                    // [kMaxF][Mp][2]
  double* sRed = reinterpret_cast<double*>(sF + kMaxF * Mp * 2);

  const long long plane = (long long)J * I;
  const long long row = (long long)j * I;
  double* phys = a.phys + (long long)b * 8 * plane;

  if (a.mode == GRID_STEP && a.diagnostics && a.dead[b]) {
    // Frozen member: emit the current level (model.py:847-864).
    double v[2] = {DBL_MAX, -DBL_MAX};
    const int op[2] = {1, 2};
    for (int i = tid; i < I; i += nt) {
      const double phi = phys[5 * plane + row + i];
      v[0] = dmin(v[0], phi);
      v[1] = dmax(v[1], phi);
      if (a.hist_fields) {
        const int src[5] = {1, 3, 5, 6, 7};
        for (int f = 0; f < 5; ++f)
          a.hist_fields[(((long long)b * 5 + f) * a.hist_T + a.hist_k) * plane + row + i] = phys[src[f] * plane + row + i];
      }
    }
    block_reduce<2>(v, op, sRed);
    if (a.hist_phiext && tid == 0) {
      double* h = a.hist_phiext + (((long long)b * a.hist_T + a.hist_k) * J + j) * 2;
      h[0] = v[0];
      h[1] = v[1];
    }
    return;
  }

  for (int i = tid; i < I; i += nt) {
    sTc[i] = a.tw[i];
    sTs[i] = a.tw[I + i];
  }
  const bool need_g = (a.mode == GRID_STEP) || (a.mode == GRID_INIT && (a.diag_winds || a.init_from_g));
  if (need_g) {
    for (int e = tid; e < kNG * Mp * 2; e += nt) {
      const int f = e / (Mp * 2), r = e % (Mp * 2);
      sG[e] = a.G[fidx(b, f, r >> 1, j, Mp, J) + (r & 1)];
    }
  }
  __syncthreads();

  const double cosphi = geo[gl.cosphi + j];
  const double alpha = cf[C_ALPHA];

  int nfwd = 0;  // number of forward-DFT fields staged in sX
  double red[4] = {0.0, DBL_MAX, DBL_MAX, -DBL_MAX};

  if (a.mode == GRID_DFT) {
    nfwd = a.have_uv ? 5 : 3;
    const double* src[5] = {a.in_eta, a.in_delta, a.in_Phi, a.in_U, a.in_V};
    for (int i = tid; i < I; i += nt)
      for (int f = 0; f < nfwd; ++f) sX[f * I + i] = (R)src[f][(long long)b * plane + row + i];
  } else {
    // Fields 0..7: A, B, C, D, E, PhiF, F, G; 8, 9: U, V when their DFT is needed.
    const bool dft_uv = (a.mode == GRID_INIT) || (a.test == 1 && a.expflag);
    nfwd = dft_uv ? 10 : 8;
    for (int i = tid; i < I; i += nt) {
      double eta, delta, Phi, U, V;
      const bool all_from_g = (a.mode == GRID_STEP) || a.init_from_g;
      if (a.mode == GRID_INIT && !a.init_from_g) {
        eta = a.in_eta[(long long)b * plane + row + i];
        delta = a.in_delta[(long long)b * plane + row + i];
        Phi = a.in_Phi[(long long)b * plane + row + i];
      }
      if (all_from_g || (a.mode == GRID_INIT && a.diag_winds)) {
        // Inverse DFT of the Hermitian-symmetric row: x = Re c0 + 2 sum_m Re(c_m e^{+i theta})
        // (spectral_transform.py:331-348 with the real part taken, model.py:973-977).
        double x[kNG];
        inv_dft_point<R, kNG>(sG, sTc, sTs, Mp, I, i, all_from_g ? 0 : 3, x);
        if (all_from_g) {
          eta = x[0];
          delta = x[1];
          Phi = x[2];
          if (a.raw_filter && a.mode == GRID_STEP) add_residue(a.resid_c, a.R0p, b, plane, row + i, eta, delta, Phi);
        }
        U = x[3];
        V = x[4];
      }
      if (a.mode == GRID_INIT && !a.diag_winds && !a.init_from_g) {
        U = a.in_U[(long long)b * plane + row + i];
        V = a.in_V[(long long)b * plane + row + i];
      }
      if (a.test == 1) {
        if (a.mode == GRID_INIT) {
          a.uvic[((long long)b * 2 + 0) * plane + row + i] = U;
          a.uvic[((long long)b * 2 + 1) * plane + row + i] = V;
        } else {
          U = a.uvic[((long long)b * 2 + 0) * plane + row + i];
          V = a.uvic[((long long)b * 2 + 1) * plane + row + i];
        }
      }
      // Physical carry and the time filter of the prev level (model.py:988-1087).
      double Phi_mid = Phi;
      if (a.mode == GRID_INIT) {
        phys[0 * plane + row + i] = eta;
        phys[1 * plane + row + i] = eta;
        phys[2 * plane + row + i] = delta;
        phys[3 * plane + row + i] = delta;
        phys[4 * plane + row + i] = Phi;
        phys[5 * plane + row + i] = Phi;
        phys[6 * plane + row + i] = U;
        phys[7 * plane + row + i] = V;
      } else if (a.phys_on) {
        physical_carry(cf, a.do_ra, a.raw_filter, alpha, phys, plane, row + i, eta, delta, Phi, U, V, Phi_mid);
      }
      if (a.hist_fields) {
        const double outv[5] = {eta, delta, Phi, U, V};
        for (int f = 0; f < 5; ++f)
          a.hist_fields[(((long long)b * 5 + f) * a.hist_T + a.hist_k) * plane + row + i] = outv[f];
      }
      // Forcing and nonlinear products (shared with the MFMA kernel).
      double prod[8];
      grid_products(cf, geo, gl, j, i, a.test, eta, Phi, U, V, prod);
      for (int f = 0; f < 8; ++f) sX[f * I + i] = (R)prod[f];
      if (nfwd == 10) {
        sX[8 * I + i] = (R)U;
        sX[9 * I + i] = (R)V;
      }
      // Diagnostics partials: RMS integrand, min |wind|, min/max of filtered Phi
      // (time_stepping.py:537-576; model.py:838-845, :1112-1117).
      const double uc = U / cosphi, vc = V / cosphi;
      red[0] += cf[cl.n_areac + j] * (uc * uc + vc * vc) / cf[C_AREAPLANET];
      red[1] = dmin(red[1], sqrt(U * U + V * V));
      red[2] = dmin(red[2], Phi_mid);
      red[3] = dmax(red[3], Phi_mid);
    }
  }
  if (a.mode != GRID_DFT && (a.diagnostics || a.hist_phiext)) {
    const int op[4] = {0, 1, 1, 2};
    block_reduce<4>(red, op, sRed);
    if (tid == 0) {
      if (a.diagnostics) {
        a.part_uv[((long long)b * J + j) * 2 + 0] = red[0];
        a.part_uv[((long long)b * J + j) * 2 + 1] = red[1];
      }
      if (a.hist_phiext && a.mode == GRID_STEP) {
        double* h = a.hist_phiext + (((long long)b * a.hist_T + a.hist_k) * J + j) * 2;
        h[0] = red[2];
        h[1] = red[3];
      }
    }
  }
  __syncthreads();

  // Forward truncated DFT: X_m = (1/I) sum_i x_i e^{-2 pi i m i / I} with the
  // index-based phase (spectral_transform.py:290-328).
  const double invI = 1.0 / (double)I;
  for (int p = tid; p < nfwd * Mp; p += nt) {
    const int f = p / Mp, m = p % Mp;
    R re = 0, im = 0;
    int idx = 0;
    const R* xf = sX + f * I;
    for (int i = 0; i < I; ++i) {
      const R x = xf[i];
      re += x * sTc[idx];
      im -= x * sTs[idx];
      idx += m;
      if (idx >= I) idx -= I;
    }
    sF[(f * Mp + m) * 2] = (R)((double)re * invI);
    sF[(f * Mp + m) * 2 + 1] = (R)((double)im * invI);
  }
  __syncthreads();

  if (a.mode == GRID_DFT) {
    // Raw Fourier of the initial fields and the w-weighted Legendre inputs
    // of the spectral state initialization.
    const double wj = geo[gl.w + j];
    for (int e = tid; e < nfwd * Mp; e += nt) {
      const int f = e / Mp, m = e % Mp;
      a.Gout[fidx(b, f, m, j, Mp, J)] = sF[(f * Mp + m) * 2];
      a.Gout[fidx(b, f, m, j, Mp, J) + 1] = sF[(f * Mp + m) * 2 + 1];
      if (f < 3) {
        R* Tm = a.T + (((long long)b * Mp + m) * kNT + f) * J * 2;
        Tm[j * 2] = (R)((double)sF[(f * Mp + m) * 2] * wj);
        Tm[j * 2 + 1] = (R)((double)sF[(f * Mp + m) * 2 + 1] * wj);
      }
    }
    return;
  }

  if (a.export_raw || a.mode == GRID_INIT) {
    R* rb = a.raw + (long long)b * kNRaw * J * Mp * 2;
    for (int e = tid; e < kNRaw * Mp; e += nt) {
      const int f = e / Mp, m = e % Mp;
      rb[((long long)f * J + j) * Mp * 2 + m * 2] = sF[(f * Mp + m) * 2];
      rb[((long long)f * J + j) * Mp * 2 + m * 2 + 1] = sF[(f * Mp + m) * 2 + 1];
    }
  }
  if (a.mode == GRID_INIT) {
    // Um0/Vm0 (model.py:727-729), the pinned copy for test 1, and the lagged
    // momentum-forcing remainder Fm + Um/taudrag (model.py:775-792; the
    // semi-implicit assembly below stores its own).
    R* g0 = a.g0uv + (long long)b * 2 * J * Mp * 2;
    R* lg = a.lag + (long long)b * 2 * J * Mp * 2;
    const double itd = cf[C_INVTD];
    for (int e = tid; e < 2 * Mp * 2; e += nt) {
      const int f = e / (Mp * 2), r = e % (Mp * 2);
      const R v = sF[(8 + f) * Mp * 2 + r];
      a.Gout[fidx(b, 3 + f, r >> 1, j, Mp, J) + (r & 1)] = v;
      g0[((long long)f * J + j) * Mp * 2 + r] = v;
      if (!a.si) lg[((long long)f * J + j) * Mp * 2 + r] = (R)((double)sF[(6 + f) * Mp * 2 + r] + (double)v * itd);
    }
  }

  // Assemble the w-weighted Legendre inputs (FP64 math).
  const double wj = geo[gl.w + j];
  const double tc = cf[cl.n_tc + j];
  const SlotConsts sk = slot_consts(cf, tc, wj, a.expflag, a.forcflag);
  for (int m = tid; m < Mp; m += nt) {
    double fm[8][2];
    for (int f = 0; f < 8; ++f) {
      fm[f][0] = (double)sF[(f * Mp + m) * 2];
      fm[f][1] = (double)sF[(f * Mp + m) * 2 + 1];
    }
    double Um[2] = {0.0, 0.0}, Vm[2] = {0.0, 0.0};
    if (a.expflag) {
      // Winds: pinned (test 1) or initial DFT (slots 8, 9), else the invrsUV output.
      const bool from_dft = (a.test == 1) || (a.mode == GRID_INIT);
      const R* su = from_dft ? sF + (8 * Mp + m) * 2 : sG + (3 * Mp + m) * 2;
      const R* sv = from_dft ? sF + (9 * Mp + m) * 2 : sG + (4 * Mp + m) * 2;
      Um[0] = su[0]; Um[1] = su[1];
      Vm[0] = sv[0]; Vm[1] = sv[1];
    }
    double out[kNT][2];
    if (a.si) {
      // Winds of the remainder: the initial DFT at initialization, else the step's.
      const bool init = a.mode == GRID_INIT;
      const R* su = init ? sF + (8 * Mp + m) * 2 : sG + (3 * Mp + m) * 2;
      const R* sv = init ? sF + (9 * Mp + m) * 2 : sG + (4 * Mp + m) * 2;
      const double u2[2] = {(double)su[0], (double)su[1]}, v2[2] = {(double)sv[0], (double)sv[1]};
      double lagR[2][2];
      si_lag_exchange(a, cf, b, j, m, fm, u2, v2, init, lagR);
      assemble_slots_si(cf, tc, wj, m, a.forcflag, a.relax, fm, lagR, out);
    } else {
      assemble_slots(sk, m, fm, Um, Vm, out);
    }
    R* Tm = a.T + (((long long)b * Mp + m) * kNT) * J * 2;
    for (int k = 0; k < kNT; ++k) {
      Tm[(k * J + j) * 2] = (R)out[k][0];
      Tm[(k * J + j) * 2 + 1] = (R)out[k][1];
    }
  }
}

template <typename R>
__global__ void k_copy_real(double* dst, long long dstride, const R* src, long long sstride, long long n) {
  const int b = blockIdx.y;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    dst[b * dstride + e] = (double)src[b * sstride + e];
}

// Plain device copy of n16 16-byte words (the previous Fourier rows a call's last step saves; a
// kernel rather than a copy command, so that captured step sequences hold kernel nodes only).
__global__ void k_copy_words(uint4* __restrict__ dst, const uint4* __restrict__ src, long long n16) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n16; e += (long long)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

template <typename R>
__global__ void k_copy_fourier(double* dst, const R* src, int f, int J, int Mp, const R* res, const double* rc,
                               int lvl) {
  const int b = blockIdx.y;
  const long long n = (long long)J * Mp * 2;
  const double cr = res ? rc[2 * b + lvl] : 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(e & 1), m = (int)((e >> 1) % Mp), j = (int)((e >> 1) / Mp);
    const long long o = fidx(b, f, m, j, Mp, J) + c;
    dst[(long long)b * n + e] = res ? (double)src[o] + cr * (double)res[o] : (double)src[o];
  }
}

// Physical carry of a modified-Euler plan rebuilt from Fourier rows, one block per (row j,
// member b): the current levels and winds from the current rows G (what the step's grid kernel
// writes for them), the prev levels from Q, the rows of the spectral prev lineage that the
// Legendre kernel advances (swx_physics.h:spectral_update, StepFlags::lineage).
template <typename R>
__global__ void __launch_bounds__(256) k_phys_synth(const R* G, const R* Q, double* phys, const R* tw, int J, int I,
                                                    int Mp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R* sTc = reinterpret_cast<R*>(smem_raw);  // [I]
  R* sTs = sTc + I;                          // [I]
  R* sG = sTs + I;                           // [kNG][Mp][2]
  R* sQ = sG + kNG * Mp * 2;                 // [3][Mp][2]
  const int j = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  for (int i = tid; i < I; i += blockDim.x) {
    sTc[i] = tw[i];
    sTs[i] = tw[I + i];
  }
  for (int e = tid; e < kNG * Mp * 2; e += blockDim.x) {
    const int f = e / (Mp * 2), r = e % (Mp * 2);
    sG[e] = G[fidx(b, f, r >> 1, j, Mp, J) + (r & 1)];
    if (f < 3) sQ[e] = Q[fidx(b, f, r >> 1, j, Mp, J) + (r & 1)];
  }
  __syncthreads();
  const long long plane = (long long)J * I;
  double* ph = phys + (long long)b * 8 * plane + (long long)j * I;
  for (int i = tid; i < I; i += blockDim.x) {
    double x[kNG], q[3];
    inv_dft_point<R, kNG>(sG, sTc, sTs, Mp, I, i, 0, x);
    inv_dft_point<R, 3>(sQ, sTc, sTs, Mp, I, i, 0, q);
    ph[0 * plane + i] = q[0];
    ph[1 * plane + i] = x[0];
    ph[2 * plane + i] = q[1];
    ph[3 * plane + i] = x[1];
    ph[4 * plane + i] = q[2];
    ph[5 * plane + i] = x[2];
    ph[6 * plane + i] = x[3];
    ph[7 * plane + i] = x[4];
  }
}

// Initial fields' non-band-limited residue of a semi-implicit / RAW plan, one block per
// (row j, member b).  On entry G holds the raw Fourier rows of the initial fields and R0f
// the rows of their band-limited part (the inverse Legendre transform of the initial
// spectral state).  Writes R0p = field0 - invDFT(band) (when R0p is non-null), then
// R0f = raw - band (fields eta, delta, Phi), and sets the residue coefficients to 1.
template <typename R>
__global__ void __launch_bounds__(256) k_residue_init(const R* G, R* R0f, double* R0p, const double* in_eta,
                                                      const double* in_delta, const double* in_Phi, const R* tw,
                                                      double* resid_c, int J, int I, int Mp) {
  const int j = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const long long plane = (long long)J * I;
  if (R0p) {
    const double* src[3] = {in_eta, in_delta, in_Phi};
    for (int e = tid; e < 3 * I; e += blockDim.x) {
      const int f = e / I, i = e % I;
      // x_i = Re c_0 + 2 sum_m Re(c_m e^{+2 pi i m i / I}) (spectral_transform.py:331-348).
      double acc = 0.0;
      int idx = 0;
      for (int m = 1; m < Mp; ++m) {
        idx += i;
        if (idx >= I) idx -= I;
        const long long o = fidx(b, f, m, j, Mp, J);
        acc += (double)R0f[o] * (double)tw[idx] - (double)R0f[o + 1] * (double)tw[I + idx];
      }
      const double band = (double)R0f[fidx(b, f, 0, j, Mp, J)] + 2.0 * acc;
      R0p[((long long)b * 3 + f) * plane + (long long)j * I + i] = src[f][(long long)b * plane + (long long)j * I + i] - band;
    }
    __syncthreads();
  }
  for (int e = tid; e < 3 * Mp * 2; e += blockDim.x) {
    const int f = e / (Mp * 2), r = e % (Mp * 2);
    const long long o = fidx(b, f, r >> 1, j, Mp, J) + (r & 1);
    R0f[o] = (R)((double)G[o] - (double)R0f[o]);
  }
  if (j == 0 && tid == 0) {
    resid_c[2 * b] = 1.0;
    resid_c[2 * b + 1] = 1.0;
  }
}

// One wave that waits `ticks` of the constant-rate wall clock (s_memrealtime) and exits: the
// probe by which a pipeline checks that two of its streams run at the same time.
__global__ void __launch_bounds__(64) k_wait_ticks(unsigned long long ticks) {
  const unsigned long long t0 = wall_clock64();
  while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}

hipError_t launch_wait_ticks(unsigned long long ticks, hipStream_t s) {
  hipLaunchKernelGGL(k_wait_ticks, dim3(1), dim3(64), 0, s, ticks);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Launchers
// ----------------------------------------------------------------------------

template <typename R>
hipError_t launch_leg(const LegArgs<R>& a, hipStream_t s) {
  const size_t smem = sizeof(R) * (kNT * a.J * 2 + 5 * (a.N + 1) * 2) + sizeof(double) * 2 * 128 + 16;
  dim3 grid(a.M + 1, a.B);
  hipLaunchKernelGGL(k_leg<R>, grid, dim3(128), smem, s, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_grid(const GridArgs<R>& a, hipStream_t s) {
  const int Mp = a.M + 1;
  const size_t smem = sizeof(R) * (2 * a.I + kNG * Mp * 2 + 10 * a.I + 10 * Mp * 2) + sizeof(double) * 4 * 256 + 16;
  dim3 grid(a.J, a.B);
  hipLaunchKernelGGL(k_grid<R>, grid, dim3(256), smem, s, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_copy_real(double* dst, long long dstride, const R* src, long long sstride, long long n, int B,
                            hipStream_t s) {
  if (n <= 0 || B <= 0) return hipSuccess;
  int gx = (int)((n + 255) / 256);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_copy_real<R>, dim3(gx, B), dim3(256), 0, s, dst, dstride, src, sstride, n);
  return hipGetLastError();
}

hipError_t launch_copy_words(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if (bytes % 16 || (reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16)
    return hipErrorInvalidValue;
  const long long n16 = (long long)(bytes / 16);
  long long gx = (n16 + 255) / 256;
  if (gx > 2048) gx = 2048;
  hipLaunchKernelGGL(k_copy_words, dim3((unsigned)gx), dim3(256), 0, s, reinterpret_cast<uint4*>(dst),
                     reinterpret_cast<const uint4*>(src), n16);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_copy_fourier(double* dst, const R* src, int f, int B, int J, int Mp, hipStream_t s, const R* res,
                               const double* rc, int lvl) {
  int gx = (int)(((long long)J * Mp * 2 + 255) / 256);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_copy_fourier<R>, dim3(gx, B), dim3(256), 0, s, dst, src, f, J, Mp, res, rc, lvl);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_phys_synth(const R* G, const R* Q, double* phys, const R* tw, int B, int J, int I, int Mp,
                             hipStream_t s) {
  const size_t lds = (size_t)(2 * I + (kNG + 3) * Mp * 2) * sizeof(R);
  hipLaunchKernelGGL(k_phys_synth<R>, dim3(J, B), dim3(256), lds, s, G, Q, phys, tw, J, I, Mp);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_residue_init(const R* G, R* R0f, double* R0p, const double* in_eta, const double* in_delta,
                               const double* in_Phi, const R* tw, double* resid_c, int B, int J, int I, int Mp,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_residue_init<R>, dim3(J, B), dim3(256), 0, s, G, R0f, R0p, in_eta, in_delta, in_Phi, tw,
                     resid_c, J, I, Mp);
  return hipGetLastError();
}

template hipError_t launch_copy_fourier<double>(double*, const double*, int, int, int, int, hipStream_t,
                                                const double*, const double*, int);
template hipError_t launch_copy_fourier<float>(double*, const float*, int, int, int, int, hipStream_t, const float*,
                                               const double*, int);
template hipError_t launch_phys_synth<double>(const double*, const double*, double*, const double*, int, int, int,
                                              int, hipStream_t);
template hipError_t launch_phys_synth<float>(const float*, const float*, double*, const float*, int, int, int, int,
                                             hipStream_t);
template hipError_t launch_residue_init<double>(const double*, double*, double*, const double*, const double*,
                                                const double*, const double*, double*, int, int, int, int,
                                                hipStream_t);
template hipError_t launch_residue_init<float>(const float*, float*, double*, const double*, const double*,
                                               const double*, const float*, double*, int, int, int, int,
                                               hipStream_t);
template hipError_t launch_blowup_gate<float>(const LegArgs<float>&, hipStream_t);
template hipError_t launch_leg<double>(const LegArgs<double>&, hipStream_t);
template hipError_t launch_leg<float>(const LegArgs<float>&, hipStream_t);
template hipError_t launch_grid<double>(const GridArgs<double>&, hipStream_t);
template hipError_t launch_grid<float>(const GridArgs<float>&, hipStream_t);
template hipError_t launch_copy_real<double>(double*, long long, const double*, long long, long long, int, hipStream_t);
template hipError_t launch_copy_real<float>(double*, long long, const float*, long long, long long, int, hipStream_t);

}  // namespace swx
