// MFMA grid kernel (FP32 transforms): inverse DFT -> grid physics -> forward
// DFT for a group of 16 latitude rows, as two back-to-back matrix-core GEMMs
// with the grid-space physics applied in registers between them.
//
//   GEMM 1 (inverse DFT, spectral_transform.py:331-348 + real part):
//     x[f][i][n] = sum_k Tinv[i][k] X[f][k][n],  k = 2m + (re, im),
//     Tinv[i][2m] = w_m cos(2 pi m i / I), Tinv[i][2m+1] = -w_m sin(...),
//     w_0 = 1, w_m = 2 (Hermitian layout, index-based phase).
//   grid physics (forcing.py, initial_conditions.py:283-292, model.py:988-1006)
//   GEMM 2 (forward truncated DFT, spectral_transform.py:290-328):
//     Y[f][k'][n] = sum_i Tfwd[k'][i] y[f][i][n],  Tfwd[2m][i] = cos / I,
//     Tfwd[2m+1][i] = -sin / I.
//   then the w-weighted Legendre inputs of the next step (swx_physics.h).
//
// Work decomposition (one workgroup = 4 waves = 16 rows (member, j)):
//   i is processed in chunks of 64: phase 1 gives each wave one 16-wide
//   i-tile of GEMM 1 (all 5 fields, v_mfma_f32_16x16x4f32), the physics on
//   its 16x16 tile, and the products stored in LDS as the B operand of
//   GEMM 2; phase 2 accumulates GEMM 2 over the chunk, wave w owning fields
//   2w and 2w+1 for every 16-row tile of k' (balanced for any M).
//   After the last chunk the accumulators go tile by tile through LDS to the
//   threads that assemble the Legendre inputs.
#include <cfloat>

#include "swx_kernels.h"
#include "swx_physics.h"

namespace swx {

typedef float v4f __attribute__((ext_vector_type(4)));

template <int MT>
struct MfmaShape {
  static constexpr int Mp = MT + 1;
  static constexpr int I = MT == 42 ? 128 : (MT == 63 ? 192 : 320);
  static constexpr int K1 = 2 * Mp;                  // inverse-DFT inner dimension
  static constexpr int K1p = (K1 + 3) / 4 * 4;       // padded to the MFMA k-step
  static constexpr int K2 = 2 * Mp;                  // forward-DFT outputs (re, im)
  static constexpr int K2p = (K2 + 15) / 16 * 16;    // padded to 16-row tiles
  static constexpr int NT2 = K2p / 16;
  static constexpr int ICH = 64;                     // longitudes per chunk
  static constexpr int ROWS = 16;
  // LDS floats: X^T [5][K1p][16] + Y^T [8][ICH][16]; the final exchange
  // buffer [2][8][16][16] reuses the Y^T region.
  static constexpr int LDS_FLOATS = 5 * K1p * 16 + 8 * ICH * 16;
};

template <int MT>
__global__ void __launch_bounds__(256) k_grid_mfma(GridArgs<float> a, const float* __restrict__ TinvT,
                                                   const float* __restrict__ TfwdT) {
  using S = MfmaShape<MT>;
  constexpr int Mp = S::Mp, I = S::I, K1 = S::K1, K1p = S::K1p, K2p = S::K2p, NT2 = S::NT2, ICH = S::ICH;
  extern __shared__ __align__(16) float lds[];
  float* sXT = lds;                    // [5][K1p][16]
  float* sYT = lds + 5 * K1p * 16;     // [8][ICH][16]
  float* sC = sYT;                     // [2][8][16][16] after the last chunk

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ln = lane & 15, lg = lane >> 4;
  const int J = a.J, N = a.N;
  const int row0 = blockIdx.x * S::ROWS;
  const int b = row0 / J, j0 = row0 % J;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;
  const long long plane = (long long)J * I;
  double* phys = a.phys + (long long)b * 8 * plane;
  const double alpha = cf[C_ALPHA];

  // Phase 0: Fourier rows (m-major, 16 consecutive j per (f, m)) -> X^T (k-major, zero padded).
  for (int e = tid; e < 5 * (K1p / 2) * 32; e += 256) {
    const int c = e & 1, n = (e >> 1) & 15, fm = e >> 5;
    const int f = fm / (K1p / 2), m = fm % (K1p / 2);
    const float v = m < Mp ? a.G[fidx(b, f, m, j0 + n, Mp, J) + c] : 0.0f;
    sXT[(f * K1p + 2 * m + c) * 16 + n] = v;
  }
  __syncthreads();

  v4f acc[2 * NT2];
#pragma unroll
  for (int p = 0; p < 2 * NT2; ++p) acc[p] = v4f{0.f, 0.f, 0.f, 0.f};

  const int n = ln;
  const int j = j0 + n;
  for (int ic = 0; ic < I; ic += ICH) {
    // ---- Phase 1: inverse DFT of one 16-wide i-tile, all five fields.  The
    // wave's A fragments (its i-tile of Tinv) are loaded up front.
    const int ib = ic + 16 * wave;
    v4f c1[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) c1[f] = v4f{0.f, 0.f, 0.f, 0.f};
    float afr[K1p / 4];
    {
      const float* Ta = TinvT + (long long)(ib + ln) * K1p + lg;
#pragma unroll
      for (int ks = 0; ks < K1p / 4; ++ks) afr[ks] = Ta[4 * ks];
    }
#pragma unroll
    for (int ks = 0; ks < K1p / 4; ++ks) {
      const float* xb = sXT + (4 * ks + lg) * 16 + ln;
#pragma unroll
      for (int f = 0; f < 5; ++f)
        c1[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks], xb[f * K1p * 16], c1[f], 0, 0, 0);
    }
    // ---- Grid physics on this lane's 4 points (row n, i = ib + 4 lg + q).
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int il = 16 * wave + 4 * lg + q;
      const int i = ic + il;
      const double eta = c1[0][q], delta = c1[1][q], Phi = c1[2][q];
      double U = c1[3][q], V = c1[4][q];
      const long long pix = (long long)j * I + i;
      if (a.test == 1) {
        U = a.uvic[((long long)b * 2 + 0) * plane + pix];
        V = a.uvic[((long long)b * 2 + 1) * plane + pix];
      }
      if (a.phys_on) {
        const double ep = phys[0 * plane + pix], ec = phys[1 * plane + pix];
        const double dp = phys[2 * plane + pix], dc = phys[3 * plane + pix];
        const double pp = phys[4 * plane + pix], pc = phys[5 * plane + pix];
        double em = ec, dm = dc, pm = pc;
        if (a.do_ra) {
          em = ec + alpha * (ep - 2.0 * ec + eta);
          dm = dc + alpha * (dp - 2.0 * dc + delta);
          pm = pc + alpha * (pp - 2.0 * pc + Phi);
        }
        phys[0 * plane + pix] = em;
        phys[1 * plane + pix] = eta;
        phys[2 * plane + pix] = dm;
        phys[3 * plane + pix] = delta;
        phys[4 * plane + pix] = pm;
        phys[5 * plane + pix] = Phi;
        phys[6 * plane + pix] = U;
        phys[7 * plane + pix] = V;
      }
      double prod[8];
      grid_products(cf, geo, gl, j, i, a.test, eta, Phi, U, V, prod);
#pragma unroll
      for (int f = 0; f < 8; ++f) sYT[(f * ICH + il) * 16 + n] = (float)prod[f];
    }
    __syncthreads();
    // ---- Phase 2: forward DFT partial sums over this chunk; the A fragments
    // of step ks+1 are fetched while step ks runs.
    float an[NT2];
#pragma unroll
    for (int t = 0; t < NT2; ++t) an[t] = TfwdT[(long long)(16 * t + ln) * I + ic + lg];
#pragma unroll 2
    for (int ks = 0; ks < ICH / 4; ++ks) {
      const int il = 4 * ks + lg;
      float ac[NT2];
#pragma unroll
      for (int t = 0; t < NT2; ++t) ac[t] = an[t];
      if (ks + 1 < ICH / 4) {
#pragma unroll
        for (int t = 0; t < NT2; ++t) an[t] = TfwdT[(long long)(16 * t + ln) * I + ic + il + 4];
      }
      const float b0 = sYT[((2 * wave) * ICH + il) * 16 + ln];
      const float b1 = sYT[((2 * wave + 1) * ICH + il) * 16 + ln];
#pragma unroll
      for (int t = 0; t < NT2; ++t) {
        acc[2 * t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t], b0, acc[2 * t], 0, 0, 0);
        acc[2 * t + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t], b1, acc[2 * t + 1], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // ---- Phase 3: assemble the Legendre inputs, two k'-tiles (16 wavenumbers)
  // per round through LDS, one (m, row) pair per thread.
#pragma unroll
  for (int t0 = 0; t0 < NT2; t0 += 2) {
#pragma unroll
    for (int dt2 = 0; dt2 < 2; ++dt2) {
      const int t = t0 + dt2;
      if (t < NT2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int f = 2 * wave + h;
#pragma unroll
          for (int q = 0; q < 4; ++q) sC[((dt2 * 8 + f) * 16 + 4 * lg + q) * 16 + ln] = acc[2 * t + h][q];
        }
      }
    }
    __syncthreads();
    {
      const int dt2 = tid >> 7, ml = (tid >> 4) & 7, nn = tid & 15;
      const int m = 8 * (t0 + dt2) + ml;
      if (t0 + dt2 < NT2 && m < Mp) {
        const int jj = j0 + nn;
        double fm[8][2];
#pragma unroll
        for (int f = 0; f < 8; ++f) {
          fm[f][0] = sC[((dt2 * 8 + f) * 16 + 2 * ml) * 16 + nn];
          fm[f][1] = sC[((dt2 * 8 + f) * 16 + 2 * ml + 1) * 16 + nn];
        }
        double Um[2], Vm[2];
        Um[0] = sXT[(3 * K1p + 2 * m) * 16 + nn];
        Um[1] = sXT[(3 * K1p + 2 * m + 1) * 16 + nn];
        Vm[0] = sXT[(4 * K1p + 2 * m) * 16 + nn];
        Vm[1] = sXT[(4 * K1p + 2 * m + 1) * 16 + nn];
        if (a.export_raw) {
          float* rb = a.raw + (long long)b * kNRaw * J * Mp * 2;
#pragma unroll
          for (int f = 0; f < kNRaw; ++f) {
            rb[((long long)f * J + jj) * Mp * 2 + 2 * m] = (float)fm[f][0];
            rb[((long long)f * J + jj) * Mp * 2 + 2 * m + 1] = (float)fm[f][1];
          }
        }
        double out[kNT][2];
        assemble_slots(cf, cf[cl.n_tc + jj], geo[gl.w + jj], m, a.expflag, a.forcflag, fm, Um, Vm, out);
        float* Tm = a.T + (((long long)b * Mp + m) * kNT) * J * 2;
#pragma unroll
        for (int k = 0; k < kNT; ++k) {
          Tm[(k * J + jj) * 2] = (float)out[k][0];
          Tm[(k * J + jj) * 2 + 1] = (float)out[k][1];
        }
      }
    }
    __syncthreads();
  }
}

template <int MT>
static hipError_t launch_mfma_t(const GridArgs<float>& a, const float* TinvT, const float* TfwdT, hipStream_t s) {
  using S = MfmaShape<MT>;
  const size_t smem = sizeof(float) * S::LDS_FLOATS;
  dim3 grid((unsigned)((long long)a.B * a.J / S::ROWS));
  hipLaunchKernelGGL(k_grid_mfma<MT>, grid, dim3(256), smem, s, a, TinvT, TfwdT);
  return hipGetLastError();
}

size_t grid_mfma_lds_bytes(int M) {
  switch (M) {
    case 42: return sizeof(float) * MfmaShape<42>::LDS_FLOATS;
    case 63: return sizeof(float) * MfmaShape<63>::LDS_FLOATS;
    case 106: return sizeof(float) * MfmaShape<106>::LDS_FLOATS;
    default: return 0;
  }
}

int grid_mfma_k1p(int M) {
  switch (M) {
    case 42: return MfmaShape<42>::K1p;
    case 63: return MfmaShape<63>::K1p;
    case 106: return MfmaShape<106>::K1p;
    default: return 0;
  }
}

int grid_mfma_k2p(int M) {
  switch (M) {
    case 42: return MfmaShape<42>::K2p;
    case 63: return MfmaShape<63>::K2p;
    case 106: return MfmaShape<106>::K2p;
    default: return 0;
  }
}

hipError_t launch_grid_mfma(const GridArgs<float>& a, const float* TinvT, const float* TfwdT, hipStream_t s) {
  switch (a.M) {
    case 42: return launch_mfma_t<42>(a, TinvT, TfwdT, s);
    case 63: return launch_mfma_t<63>(a, TinvT, TfwdT, s);
    case 106: return launch_mfma_t<106>(a, TinvT, TfwdT, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace swx
