// MFMA grid adjoint kernel (FP32): the reverse of k_grid_mfma for one step,
// replacing k_grid_adj in GADJ_STEP mode where its LDS fits (T42, T63).
//
// Per 16 rows (member, latitude), with the forward kernel's DFT tables:
//   phase 0  X^T [12][k][row]: the Fourier rows of S_t (eta, Phi, U, V) and the
//            Fourier cotangents Fbar = assemble^T(Tbar) (swx_physics.h), the
//            latter scaled by 1 / (I w_m) so that DFT^T = Tinv X~:
//            xbar_i = (1/I) sum_m (Fbar.re cos - Fbar.im sin)
//                   = sum_k Tinv[i][k] Fbar[k] / (I w_m);
//   phase 1  GEMM 1 (v_mfma_f32_16x16x4f32): the grid fields of S_t and the
//            product cotangents on one 16-wide longitude tile per wave;
//   pointwise  grid_products_adjoint (the Q < 0 clamp included) and the
//            per-row partials of taurad, taudrag, DPhieq;
//   phase 2  GEMM 2: invDFT^T = Tinv^T = diag(I w_m) Tfwd, wave w owning the
//            cotangent of eta, Phi, U or V over all k' tiles;
//   output   Gbar (delta's cotangent is zero) and pgrad.
// Reference functions differentiated: spectral_transform.py:290-348,
// forcing.py:65-155, initial_conditions.py:242-292.
#include "swx_adjoint.h"
#include "swx_grid_common.h"

namespace swx {

namespace {

template <int MT>
struct AdjShape {
  static constexpr int Mp = MT + 1;
  static constexpr int I = MT == 42 ? 128 : (MT == 63 ? 192 : 320);
  static constexpr int K1 = 2 * Mp;
  static constexpr int K1p = (K1 + 3) / 4 * 4;
  static constexpr int KQ1 = (K1p / 4 + 3) / 4 * 4;  // per-lane run of the inverse table (its layout)
  static constexpr int K2p = (K1 + 15) / 16 * 16;
  static constexpr int NT2 = K2p / 16;
  static constexpr int ICH = 64;
  static constexpr int NX = 12;  // GEMM-1 fields: eta, Phi, U, V of S_t, then the 8 Fourier cotangents
  // LDS floats: X^T [NX][K1p][16], grid cotangents [4][ICH][16], row partials [4 waves][16][3] doubles.
  static constexpr int LDS_FLOATS = NX * K1p * 16 + 4 * ICH * 16 + 2 * 4 * 16 * 3;
};

}  // namespace

template <int MT>
__global__ void __launch_bounds__(256) k_grid_adj_mfma(GridAdjArgs<float> a, const float* __restrict__ TinvF,
                                                       const float* __restrict__ TfwdF) {
  using S = AdjShape<MT>;
  constexpr int Mp = S::Mp, I = S::I, K1p = S::K1p, KQ1 = S::KQ1, NT2 = S::NT2, ICH = S::ICH, NX = S::NX;
  extern __shared__ __align__(16) float lds[];
  float* sXT = lds;                                                 // [NX][K1p][16]
  float* sYT = lds + NX * K1p * 16;                                 // [4][ICH][16]
  double* sRed = reinterpret_cast<double*>(sYT + 4 * ICH * 16);     // [4][16][3]

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int J = a.J, N = a.N;
  const int row0 = blockIdx.x * 16;
  const int b = row0 / J, j0 = row0 % J;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;
  const long long plane = (long long)J * I;

  // Phase 0a: Fourier rows of S_t, G fields 0, 2, 3, 4 -> X^T fields 0..3 (float4 loads first).
  {
    constexpr int NE = 4 * (K1p / 2) * 8;
    constexpr int NQ = (NE + 255) / 256;
    v4f gv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      const int jq = e & 7, fm = e >> 3;
      const int fx = fm / (K1p / 2) < 4 ? fm / (K1p / 2) : 3, m = fm % (K1p / 2);
      const int fg = fx == 0 ? 0 : fx + 1;
      const v4f v = *reinterpret_cast<const v4f*>(a.G + fidx(b, fg, m < Mp ? m : Mp - 1, j0 + 2 * jq, Mp, J));
      gv[u] = (e < NE && m < Mp) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      if (e < NE) {
        const int jq = e & 7, fm = e >> 3;
        const int fx = fm / (K1p / 2), m = fm % (K1p / 2);
        float* xr = sXT + (fx * K1p + 2 * m) * 16 + 2 * jq;
        *reinterpret_cast<float2*>(xr) = make_float2(gv[u][0], gv[u][2]);
        *reinterpret_cast<float2*>(xr + 16) = make_float2(gv[u][1], gv[u][3]);
      }
    }
  }
  // Phase 0b: Fourier cotangents per (m, row) -> X^T fields 4..11, scaled by 1 / (I w_m).
  for (int p = tid; p < (K1p / 2) * 16; p += 256) {
    const int m = p >> 4, r = p & 15;
    float* xr = sXT + (4 * K1p + 2 * m) * 16 + r;
    if (m < Mp) {
      const int jj = j0 + r;
      const float* Tb = a.Tbar + (((long long)b * Mp + m) * kNT) * J * 2;
      double tau[kNT][2];
#pragma unroll
      for (int k = 0; k < kNT; ++k) {
        tau[k][0] = Tb[(k * J + jj) * 2];
        tau[k][1] = Tb[(k * J + jj) * 2 + 1];
      }
      double fb[8][2];
      assemble_slots_adjoint(cf, cf[cl.n_tc + jj], geo[gl.w + jj], m, a.forcflag, tau, fb);
      const double sc = 1.0 / ((double)I * (m == 0 ? 1.0 : 2.0));
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        xr[(f * K1p) * 16] = (float)(fb[f][0] * sc);
        xr[(f * K1p + 1) * 16] = (float)(fb[f][1] * sc);
      }
    } else {
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        xr[(f * K1p) * 16] = 0.0f;
        xr[(f * K1p + 1) * 16] = 0.0f;
      }
    }
  }
  __syncthreads();

  v4f acc[NT2];
#pragma unroll
  for (int t = 0; t < NT2; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  double dpar[3] = {0.0, 0.0, 0.0};
  const int j = j0 + ln;
  const v4f* Tf = reinterpret_cast<const v4f*>(TfwdF) + lane;

  for (int ic = 0; ic < I; ic += ICH) {
    // ---- Phase 1: GEMM 1 on this wave's longitude tile, 12 fields.
    const int ib = ic + 16 * wave;
    v4f afr[KQ1 / 4];
    {
      const v4f* Ta = reinterpret_cast<const v4f*>(TinvF + ((long long)(ib / 16) * 64 + lane) * KQ1);
#pragma unroll
      for (int q4 = 0; q4 < KQ1 / 4; ++q4) afr[q4] = Ta[q4];
    }
    v4f an[NT2];
#pragma unroll
    for (int t = 0; t < NT2; ++t) an[t] = Tf[((long long)t * (I / 16) + ic / 16) * 64];
    v4f c1[NX];
#pragma unroll
    for (int f = 0; f < NX; ++f) c1[f] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K1p / 4; ++ks) {
      const float* xb = sXT + (4 * ks + lg) * 16 + ln;
#pragma unroll
      for (int f = 0; f < NX; ++f)
        c1[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks / 4][ks % 4], xb[f * K1p * 16], c1[f], 0, 0, 0);
    }
    // ---- Pointwise adjoint on this lane's 4 points (row ln, i = ib + 4 lg + q).
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int il = 16 * wave + 4 * lg + q;
      const int i = ic + il;
      double U = c1[2][q], V = c1[3][q];
      if (a.test == 1) {
        U = a.uvic[((long long)b * 2 + 0) * plane + (long long)j * I + i];
        V = a.uvic[((long long)b * 2 + 1) * plane + (long long)j * I + i];
      }
      double pbar[8];
#pragma unroll
      for (int f = 0; f < 8; ++f) pbar[f] = c1[4 + f][q];
      double o[4];
      grid_products_adjoint(cf, geo, gl, j, i, a.test, c1[0][q], c1[1][q], U, V, pbar, o, dpar);
      sYT[(0 * ICH + il) * 16 + ln] = (float)o[0];
      sYT[(1 * ICH + il) * 16 + ln] = (float)o[1];
      // The pinned winds of test 1 do not depend on the state.
      sYT[(2 * ICH + il) * 16 + ln] = a.test == 1 ? 0.0f : (float)o[2];
      sYT[(3 * ICH + il) * 16 + ln] = a.test == 1 ? 0.0f : (float)o[3];
    }
    __syncthreads();
    // ---- Phase 2: GEMM 2 partial sums, wave w owning cotangent field w; the next
    // 16-longitude group's A fragments are in flight while the current one runs.
#pragma unroll 1
    for (int gg = 0; gg < ICH / 16; ++gg) {
      v4f ac[NT2];
#pragma unroll
      for (int t = 0; t < NT2; ++t) ac[t] = an[t];
      if (gg + 1 < ICH / 16) {
#pragma unroll
        for (int t = 0; t < NT2; ++t) an[t] = Tf[((long long)t * (I / 16) + ic / 16 + gg + 1) * 64];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float bv = sYT[(wave * ICH + 16 * gg + 4 * q + lg) * 16 + ln];
#pragma unroll
        for (int t = 0; t < NT2; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], bv, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // Per-row parameter partials: row ln is held by lanes ln + 16 g of every wave.
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = dpar[c];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    dpar[c] = v;
  }
  if (lg == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) sRed[(wave * 16 + ln) * 3 + c] = dpar[c];
  __syncthreads();
  if (tid < 16 * 3) {
    const int r = tid / 3, c = tid % 3;
    const double sum = sRed[(0 * 16 + r) * 3 + c] + sRed[(1 * 16 + r) * 3 + c] + sRed[(2 * 16 + r) * 3 + c] +
                       sRed[(3 * 16 + r) * 3 + c];
    a.pgrad[((long long)b * J + j0 + r) * 3 + c] += sum;
  }

  // Gbar = I w_m (Tfwd xbar): wave w -> field eta, Phi, U, V (G index 0, 2, 3, 4).
  const int fo = wave == 0 ? 0 : wave + 1;
#pragma unroll
  for (int t = 0; t < NT2; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kp = 16 * t + 4 * lg + q;
      const int m = kp >> 1, c = kp & 1;
      if (m < Mp) {
        const float v = (m == 0 && c == 1) ? 0.0f : (float)I * (m == 0 ? 1.0f : 2.0f) * acc[t][q];
        a.Gbar[fidx(b, fo, m, j, Mp, J) + c] = v;
      }
    }
  }
  // delta's cotangent is zero (the products do not depend on delta).
  for (int p = tid; p < Mp * 16 * 2; p += 256) {
    const int m = p >> 5, r = (p >> 1) & 15, c = p & 1;
    a.Gbar[fidx(b, 1, m, j0 + r, Mp, J) + c] = 0.0f;
  }
}

// ----------------------------------------------------------------------------
// Half-period folded variant (the forward grid kernel's fold, swx_mfma.hip): GEMM 1 gives
// the even / odd wavenumber sums E, O of the 12 fields on i < H = I/2 (x(i) = E + O,
// x(i + H) = E - O), the pointwise adjoint runs on both points of each pair, and GEMM 2 takes
// s = y(i) + y(i + H) for the even-m output tiles and d = y(i) - y(i + H) for the odd ones.
// Tables and row orders: those of k_grid_mfma_fold (swx_kernels.h FoldDims).  At T42 (one
// chunk) s / d reuse the X^T space, which keeps two workgroups per CU.
// SC = true: the shapes of the S_t-cached kernel below (k_grid_adj_mfma_fold_sc8), 8 staged fields.
template <int MT, bool SC = false>
struct AdjFoldShape {
  static constexpr int Mp = MT + 1;
  static constexpr int I = MT == 42 ? 128 : (MT == 63 ? 192 : 320);
  static constexpr int H = I / 2;
  static constexpr int NE = (Mp + 1) / 2, NO = Mp / 2;
  static constexpr int KE4 = (2 * NE + 3) / 4 * 4, KO4 = (2 * NO + 3) / 4 * 4;
  static constexpr int K1 = KE4 + KO4;
  static constexpr int KQ1 = (K1 / 4 + 3) / 4 * 4;
  static constexpr int TE = (2 * NE + 15) / 16, TO = (2 * NO + 15) / 16;
  static constexpr int NT2 = TE + TO;
  static constexpr int ICH = 64;
  static constexpr int NCH = (H + ICH - 1) / ICH;
  static constexpr int PB = SC ? 0 : 4;   // X^T field of the first product cotangent
  static constexpr int NX = PB + 8;       // staged fields in the main loop
  static constexpr int XT_FLOATS = NX * K1 * 16;
  static constexpr int SD_FLOATS = 2 * 4 * ICH * 16;
  static constexpr int MAIN_FLOATS = NCH == 1 ? (XT_FLOATS > SD_FLOATS ? XT_FLOATS : SD_FLOATS) : XT_FLOATS + SD_FLOATS;
  static constexpr int LDS_FLOATS = MAIN_FLOATS + 2 * 4 * 16 * 3;  // + row partials [4][16][3] doubles
  // S_t cache: per (workgroup, chunk, wave) 8 float4 (E, O of 4 fields) x 64 lanes.
  static constexpr long long cache_floats(long long nblk) { return SC ? nblk * NCH * 4 * 8 * 64 * 4 : 0; }
};

// GEMM 1 on one wave's 16-point tile: even / odd wavenumber sums of NF staged X^T fields
// (field stride K1 x 16 floats from xt).  SB > 0: a scheduling barrier every SB k-steps, which
// bounds how many LDS reads the compiler hoists ahead (registers at two waves per SIMD).
template <int K1, int KE4, int KQ1, int NF, int SB = 0>
__device__ __forceinline__ void adj_fold_gemm1(const float* __restrict__ TinvF, const float* xt, int ib, int lane,
                                               v4f (&cE)[NF], v4f (&cO)[NF]) {
  const int ln = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int f = 0; f < NF; ++f) cE[f] = cO[f] = v4f{0.f, 0.f, 0.f, 0.f};
  v4f afr[KQ1 / 4];
  const v4f* Ta = reinterpret_cast<const v4f*>(TinvF + ((long long)(ib / 16) * 64 + lane) * KQ1);
#pragma unroll
  for (int q4 = 0; q4 < KQ1 / 4; ++q4) afr[q4] = Ta[q4];
#pragma unroll
  for (int ks = 0; ks < K1 / 4; ++ks) {
    const float* xb = xt + (4 * ks + lg) * 16 + ln;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (ks < KE4 / 4)
        cE[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks / 4][ks % 4], xb[f * K1 * 16], cE[f], 0, 0, 0);
      else
        cO[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks / 4][ks % 4], xb[f * K1 * 16], cO[f], 0, 0, 0);
    }
    if constexpr (SB > 0) {
      if (ks % SB == SB - 1) __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int MT>
__device__ __forceinline__ void grid_adj_mfma_fold_body(const GridAdjArgs<float>& a, const float* __restrict__ TinvF,
                                                        const float* __restrict__ TfwdF) {
  using S = AdjFoldShape<MT>;
  constexpr int Mp = S::Mp, I = S::I, H = S::H, K1 = S::K1, KE4 = S::KE4, KQ1 = S::KQ1, NT2 = S::NT2, TE = S::TE;
  constexpr int ICH = S::ICH, NX = S::NX, PB = S::PB, NCH = S::NCH;
  // Threads exchange data through LDS only, so the barriers are LDS barriers and global loads
  // stay in flight across them.
  extern __shared__ __align__(16) float lds[];
  float* sXT = lds;                                                       // [NX][K1][16]
  float* sS = lds + (NCH == 1 ? 0 : S::XT_FLOATS);                        // [4][ICH][16] s
  float* sD = sS + 4 * ICH * 16;                                          // [4][ICH][16] d
  double* sRed = reinterpret_cast<double*>(lds + S::MAIN_FLOATS);         // [4][16][3]

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int J = a.J, N = a.N;
  const int row0 = blockIdx.x * 16;
  const int b = row0 / J, j0 = row0 % J;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;
  const long long plane = (long long)J * I;
  auto slot_m = [](int sl) { return sl < KE4 / 2 ? 2 * sl : 2 * (sl - KE4 / 2) + 1; };
  // Phase 0a: Fourier rows of S_t (G fields 0, 2, 3, 4) -> X^T fields 0..3, folded k order.
  {
    constexpr int NE4 = 4 * (K1 / 2) * 8;
    constexpr int NQ = (NE4 + 255) / 256;
    v4f gv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      const int jq = e & 7, fs = e >> 3;
      const int fx = fs / (K1 / 2) < 4 ? fs / (K1 / 2) : 3, m = slot_m(fs % (K1 / 2));
      const int fg = fx == 0 ? 0 : fx + 1;
      const v4f v = *reinterpret_cast<const v4f*>(a.G + fidx(b, fg, m < Mp ? m : Mp - 1, j0 + 2 * jq, Mp, J));
      gv[u] = (e < NE4 && m < Mp) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      if (e < NE4) {
        const int jq = e & 7, fs = e >> 3;
        const int fx = fs / (K1 / 2), sl = fs % (K1 / 2);
        float* xr = sXT + (fx * K1 + 2 * sl) * 16 + 2 * jq;
        *reinterpret_cast<float2*>(xr) = make_float2(gv[u][0], gv[u][2]);
        *reinterpret_cast<float2*>(xr + 16) = make_float2(gv[u][1], gv[u][3]);
      }
    }
  }
  // The slot cotangents of the thread's phase-0b items.
  constexpr int NP0 = (K1 / 2) * 16;
  constexpr int NIT = (NP0 + 255) / 256;
  float tv[NIT][kNT][2];
#pragma unroll
  for (int u = 0; u < NIT; ++u) {
    const int p = tid + 256 * u;
    const int sl = (p < NP0 ? p : 0) >> 4, r = p & 15;
    const int m = slot_m(sl);
    const float* Tb = a.Tbar + (((long long)b * Mp + (m < Mp ? m : Mp - 1)) * kNT) * J * 2;
#pragma unroll
    for (int k = 0; k < kNT; ++k) {
      tv[u][k][0] = Tb[(k * J + j0 + r) * 2];
      tv[u][k][1] = Tb[(k * J + j0 + r) * 2 + 1];
    }
  }
  // Phase 0b: Fourier cotangents per (m, row) -> X^T fields PB.., scaled by 1 / (I w_m); the slot
  // cotangents of all the thread's items are loaded first.
  {
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
      const int p = tid + 256 * u;
      if (p >= NP0) continue;
      const int sl = p >> 4, r = p & 15;
      const int m = slot_m(sl);
      float* xr = sXT + (PB * K1 + 2 * sl) * 16 + r;
      if (m < Mp) {
        const int jj = j0 + r;
        double tau[kNT][2];
#pragma unroll
        for (int k = 0; k < kNT; ++k) {
          tau[k][0] = tv[u][k][0];
          tau[k][1] = tv[u][k][1];
        }
        double fb[8][2];
        assemble_slots_adjoint(cf, cf[cl.n_tc + jj], geo[gl.w + jj], m, a.forcflag, tau, fb);
        const double sc = 1.0 / ((double)I * (m == 0 ? 1.0 : 2.0));
#pragma unroll
        for (int f = 0; f < 8; ++f) {
          xr[(f * K1) * 16] = (float)(fb[f][0] * sc);
          xr[(f * K1 + 1) * 16] = (float)(fb[f][1] * sc);
        }
      } else {
#pragma unroll
        for (int f = 0; f < 8; ++f) {
          xr[(f * K1) * 16] = 0.0f;
          xr[(f * K1 + 1) * 16] = 0.0f;
        }
      }
    }
  }
  lds_barrier();

  v4f acc[NT2];
#pragma unroll
  for (int t = 0; t < NT2; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  double dpar[3] = {0.0, 0.0, 0.0};
  const int j = j0 + ln;
  const v4f* Tf = reinterpret_cast<const v4f*>(TfwdF) + lane;

  for (int ic = 0; ic < H; ic += ICH) {
    const int ntc = (H - ic) / 16 < ICH / 16 ? (H - ic) / 16 : ICH / 16;
    const bool active = wave < ntc;
    const int ib = ic + 16 * wave;
    // ---- Phase 1: even / odd sums of the 12 staged fields on this wave's tile.
    v4f cE[NX], cO[NX];
    if (active) {
      adj_fold_gemm1<K1, KE4, KQ1, NX>(TinvF, sXT, ib, lane, cE, cO);
    } else {
#pragma unroll
      for (int f = 0; f < NX; ++f) cE[f] = cO[f] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    v4f an[NT2];
#pragma unroll
    for (int t = 0; t < NT2; ++t) an[t] = Tf[((long long)t * (H / 16) + ic / 16) * 64];
    if constexpr (NCH == 1) lds_barrier();  // s / d overwrite X^T
    // ---- Pointwise adjoint on the lane's 4 point pairs (row ln, i = ib + 4 lg + q and i + H).
    if (active) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int il = 16 * wave + 4 * lg + q;
        const int i0 = ic + il;
        double o2[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = i0 + h * H;
          const float sg = h == 0 ? 1.0f : -1.0f;
          double x4[4];
#pragma unroll
          for (int f = 0; f < 4; ++f) x4[f] = cE[f][q] + sg * cO[f][q];
          if (a.test == 1) {
            x4[2] = a.uvic[((long long)b * 2 + 0) * plane + (long long)j * I + i];
            x4[3] = a.uvic[((long long)b * 2 + 1) * plane + (long long)j * I + i];
          }
          double pbar[8];
#pragma unroll
          for (int f = 0; f < 8; ++f) pbar[f] = cE[PB + f][q] + sg * cO[PB + f][q];
          grid_products_adjoint(cf, geo, gl, j, i, a.test, x4[0], x4[1], x4[2], x4[3], pbar, o2[h], dpar);
        }
        // The pinned winds of test 1 do not depend on the state.
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const bool zero = a.test == 1 && f >= 2;
          sS[(f * ICH + il) * 16 + ln] = zero ? 0.0f : (float)(o2[0][f] + o2[1][f]);
          sD[(f * ICH + il) * 16 + ln] = zero ? 0.0f : (float)(o2[0][f] - o2[1][f]);
        }
      }
    }
    lds_barrier();
    // ---- Phase 2: GEMM 2 partial sums, wave w owning cotangent field w; even-m tiles read s,
    // odd-m tiles d; the next group's A fragments are in flight while the current one runs.
#pragma unroll 1
    for (int gg = 0; gg < ntc; ++gg) {
      v4f ac[NT2];
#pragma unroll
      for (int t = 0; t < NT2; ++t) ac[t] = an[t];
      if (gg + 1 < ntc) {
#pragma unroll
        for (int t = 0; t < NT2; ++t) an[t] = Tf[((long long)t * (H / 16) + ic / 16 + gg + 1) * 64];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int il = 16 * gg + 4 * q + lg;
        const float bs = sS[(wave * ICH + il) * 16 + ln];
        const float bd = sD[(wave * ICH + il) * 16 + ln];
#pragma unroll
        for (int t = 0; t < NT2; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], t < TE ? bs : bd, acc[t], 0, 0, 0);
      }
    }
    lds_barrier();
  }

  // Per-row parameter partials: row ln is held by lanes ln + 16 g of every wave.
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = dpar[c];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    dpar[c] = v;
  }
  if (lg == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) sRed[(wave * 16 + ln) * 3 + c] = dpar[c];
  lds_barrier();
  if (tid < 16 * 3) {
    const int r = tid / 3, c = tid % 3;
    const double sum = sRed[(0 * 16 + r) * 3 + c] + sRed[(1 * 16 + r) * 3 + c] + sRed[(2 * 16 + r) * 3 + c] +
                       sRed[(3 * 16 + r) * 3 + c];
    a.pgrad[((long long)b * J + j0 + r) * 3 + c] += sum;
  }

  // Gbar = I w_m (Tfwd xbar) in the folded output rows: wave w -> field eta, Phi, U, V.
  const int fo = wave == 0 ? 0 : wave + 1;
#pragma unroll
  for (int t = 0; t < NT2; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = 16 * (t < TE ? t : t - TE) + 4 * lg + q;
      const int m = t < TE ? 2 * (rr >> 1) : 2 * (rr >> 1) + 1, c = rr & 1;
      if (m < Mp) {
        const float v = (m == 0 && c == 1) ? 0.0f : (float)I * (m == 0 ? 1.0f : 2.0f) * acc[t][q];
        a.Gbar[fidx(b, fo, m, j, Mp, J) + c] = v;
      }
    }
  }
  // delta's cotangent is zero (the products do not depend on delta).
  for (int p = tid; p < Mp * 16 * 2; p += 256) {
    const int m = p >> 5, r = (p >> 1) & 15, c = p & 1;
    a.Gbar[fidx(b, 1, m, j0 + r, Mp, J) + c] = 0.0f;
  }
}

template <int MT>
__global__ void __launch_bounds__(256) k_grid_adj_mfma_fold(GridAdjArgs<float> a, const float* __restrict__ TinvF,
                                                            const float* __restrict__ TfwdF) {
  grid_adj_mfma_fold_body<MT>(a, TinvF, TfwdF);
}

// T42: allocated for 2+ waves per SIMD (accumulators in VGPRs): 214 registers instead of 318,
// 2 workgroups per CU instead of 1, no spills (T63 would spill and keeps the kernel above).
template <int MT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_grid_adj_mfma_fold_w2(GridAdjArgs<float> a, const float* __restrict__ TinvF, const float* __restrict__ TfwdF) {
  grid_adj_mfma_fold_body<MT>(a, TinvF, TfwdF);
}

// Eight-wave form of the S_t-cached kernel (two waves per SIMD, 512 threads): phase S spreads the
// chunk tiles over all eight waves; in the main loop waves w and w + 4 share tile w & 3, wave
// w < 4 taking GEMM 1 of cotangent fields 0..3 and the pointwise adjoint of the tile's points
// q = 0, 1, wave w + 4 fields 4..7 and q = 2, 3.  Each hands the other its sums for the other's
// points through the s / d slots of those points (the slots the receiving lane overwrites with
// its own s / d after reading them), and GEMM 2 splits the output tiles between the halves.
template <int MT>
struct AdjFoldSC8Shape {
  using S = AdjFoldShape<MT, true>;
  static constexpr int TH = (S::NT2 + 1) / 2;  // GEMM-2 output tiles per half
  static constexpr int RED_FLOATS = 2 * 8 * 16 * 3;  // row partials [8][16][3] doubles
  static constexpr int LDS_FLOATS = S::XT_FLOATS + S::SD_FLOATS + RED_FLOATS;
};

template <int MT>
__device__ __forceinline__ void grid_adj_mfma_fold_sc8_body(const GridAdjArgs<float>& a,
                                                            const float* __restrict__ TinvF,
                                                            const float* __restrict__ TfwdF) {
  using S = AdjFoldShape<MT, true>;
  constexpr int Mp = S::Mp, I = S::I, H = S::H, K1 = S::K1, KE4 = S::KE4, KQ1 = S::KQ1, NT2 = S::NT2, TE = S::TE;
  constexpr int ICH = S::ICH, NCH = S::NCH, TH = AdjFoldSC8Shape<MT>::TH;
  constexpr int NW = 8, NTH = 64 * NW;
  constexpr int SB = 8;  // GEMM-1 k-steps between scheduling barriers
  static_assert(NCH > 1 && ICH == 64, "four 16-point tiles per chunk, s / d apart from X^T");
  extern __shared__ __align__(16) float lds[];
  float* sXT = lds;                                                // [8][K1][16]
  float* sS = lds + S::XT_FLOATS;                                  // [4][ICH][16] s
  float* sD = sS + 4 * ICH * 16;                                   // [4][ICH][16] d
  double* sRed = reinterpret_cast<double*>(sD + 4 * ICH * 16);     // [8][16][3]

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int wt = wave & 3;   // tile within the chunk
  const int hf = wave >> 2;  // half: GEMM-1 fields 4 hf .. 4 hf + 3, points q = 2 hf, 2 hf + 1
  const int J = a.J, N = a.N;
  const int row0 = blockIdx.x * 16;
  const int b = row0 / J, j0 = row0 % J;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;
  const long long plane = (long long)J * I;
  auto slot_m = [](int sl) { return sl < KE4 / 2 ? 2 * sl : 2 * (sl - KE4 / 2) + 1; };
  // S_t cache entries of (chunk c, tile t): 8 float4 (E, O of eta, Phi, U, V) at stride 64 lanes.
  auto cache_at = [&](int c, int t) {
    return reinterpret_cast<v4f*>(a.stc) + (((long long)blockIdx.x * NCH + c) * 4 + t) * 8 * 64 + lane;
  };
  // The 8 s / d slots of point il of the chunk, row ln: k < 4 -> s of field k, else d of field k - 4.
  auto slot = [&](int k, int il) -> float* {
    return k < 4 ? sS + (k * ICH + il) * 16 + ln : sD + ((k - 4) * ICH + il) * 16 + ln;
  };

  // Phase 0a: Fourier rows of S_t (G fields 0, 2, 3, 4) -> X^T fields 0..3, folded k order.
  {
    constexpr int NE4 = 4 * (K1 / 2) * 8;
    constexpr int NQ = (NE4 + NTH - 1) / NTH;
    v4f gv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + NTH * u;
      const int jq = e & 7, fs = e >> 3;
      const int fx = fs / (K1 / 2) < 4 ? fs / (K1 / 2) : 3, m = slot_m(fs % (K1 / 2));
      const int fg = fx == 0 ? 0 : fx + 1;
      const v4f v = *reinterpret_cast<const v4f*>(a.G + fidx(b, fg, m < Mp ? m : Mp - 1, j0 + 2 * jq, Mp, J));
      gv[u] = (e < NE4 && m < Mp) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + NTH * u;
      if (e < NE4) {
        const int jq = e & 7, fs = e >> 3;
        const int fx = fs / (K1 / 2), sl = fs % (K1 / 2);
        float* xr = sXT + (fx * K1 + 2 * sl) * 16 + 2 * jq;
        *reinterpret_cast<float2*>(xr) = make_float2(gv[u][0], gv[u][2]);
        *reinterpret_cast<float2*>(xr + 16) = make_float2(gv[u][1], gv[u][3]);
      }
    }
  }
  __syncthreads();
  // Phase S: GEMM 1 of the 4 fields of S_t, tile t of the H / 16 on wave t mod 8, into the cache.
#pragma unroll 1
  for (int t = wave; t < H / 16; t += NW) {
    v4f sE[4], sO[4];
    adj_fold_gemm1<K1, KE4, KQ1, 4, SB>(TinvF, sXT, 16 * t, lane, sE, sO);
    v4f* dst = cache_at(t / (ICH / 16), t % (ICH / 16));
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      dst[(2 * f) * 64] = sE[f];
      dst[(2 * f + 1) * 64] = sO[f];
    }
  }
  __syncthreads();  // the cache is read by other waves; the cotangents overwrite X^T next

  // The slot cotangents of the thread's phase-0b items (loaded after phase S: live across it they
  // would not fit the registers of two waves per SIMD).
  constexpr int NP0 = (K1 / 2) * 16;
  constexpr int NIT = (NP0 + NTH - 1) / NTH;
  float tv[NIT][kNT][2];
#pragma unroll
  for (int u = 0; u < NIT; ++u) {
    const int p = tid + NTH * u;
    const int sl = (p < NP0 ? p : 0) >> 4, r = p & 15;
    const int m = slot_m(sl);
    const float* Tb = a.Tbar + (((long long)b * Mp + (m < Mp ? m : Mp - 1)) * kNT) * J * 2;
#pragma unroll
    for (int k = 0; k < kNT; ++k) {
      tv[u][k][0] = Tb[(k * J + j0 + r) * 2];
      tv[u][k][1] = Tb[(k * J + j0 + r) * 2 + 1];
    }
  }

  // Phase 0b: Fourier cotangents per (m, row) -> X^T fields 0..7, scaled by 1 / (I w_m).
#pragma unroll
  for (int u = 0; u < NIT; ++u) {
    const int p = tid + NTH * u;
    if (p >= NP0) continue;
    const int sl = p >> 4, r = p & 15;
    const int m = slot_m(sl);
    float* xr = sXT + (2 * sl) * 16 + r;
    if (m < Mp) {
      const int jj = j0 + r;
      double tau[kNT][2];
#pragma unroll
      for (int k = 0; k < kNT; ++k) {
        tau[k][0] = tv[u][k][0];
        tau[k][1] = tv[u][k][1];
      }
      double fb[8][2];
      assemble_slots_adjoint(cf, cf[cl.n_tc + jj], geo[gl.w + jj], m, a.forcflag, tau, fb);
      const double sc = 1.0 / ((double)I * (m == 0 ? 1.0 : 2.0));
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        xr[(f * K1) * 16] = (float)(fb[f][0] * sc);
        xr[(f * K1 + 1) * 16] = (float)(fb[f][1] * sc);
      }
    } else {
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        xr[(f * K1) * 16] = 0.0f;
        xr[(f * K1 + 1) * 16] = 0.0f;
      }
    }
  }
  lds_barrier();

  v4f acc[TH];
#pragma unroll
  for (int t = 0; t < TH; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  double dpar[3] = {0.0, 0.0, 0.0};
  const int j = j0 + ln;
  const v4f* Tf = reinterpret_cast<const v4f*>(TfwdF) + lane;
  const int t2 = hf * TH;  // first GEMM-2 output tile of this half

#pragma unroll 1
  for (int ic = 0; ic < H; ic += ICH) {
    const int ntc = (H - ic) / 16 < ICH / 16 ? (H - ic) / 16 : ICH / 16;
    const bool active = wt < ntc;
    const int ib = ic + 16 * wt;
    // ---- Phase 1: S_t of this half's points from the cache; GEMM 1 of this half's 4 fields.
    float2 st[8];
    if (active) {
      const float2* src = reinterpret_cast<const float2*>(cache_at(ic / ICH, wt));
#pragma unroll
      for (int v = 0; v < 8; ++v) st[v] = src[2 * (v * 64) + hf];
    }
    v4f cE[4], cO[4];
    if (active) {
      adj_fold_gemm1<K1, KE4, KQ1, 4, SB>(TinvF, sXT + hf * 4 * K1 * 16, ib, lane, cE, cO);
      // Hand the other half its points' sums of this half's fields.
#pragma unroll
      for (int qq = 0; qq < 2; ++qq) {
        const int q = 2 * (1 - hf) + qq;
        const int il = 16 * wt + 4 * lg + q;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const float e = hf ? cE[f][qq] : cE[f][2 + qq];
          const float o = hf ? cO[f][qq] : cO[f][2 + qq];
          *slot(2 * f, il) = e;
          *slot(2 * f + 1, il) = o;
        }
      }
    }
    v4f an[TH];
#pragma unroll
    for (int t = 0; t < TH; ++t)
      an[t] = t2 + t < NT2 ? Tf[((long long)(t2 + t) * (H / 16) + ic / 16) * 64] : v4f{0.f, 0.f, 0.f, 0.f};
    lds_barrier();
    // ---- Pointwise adjoint on this half's 2 point pairs (row ln, i = ib + 4 lg + q and i + H).
    if (active) {
#pragma unroll
      for (int qq = 0; qq < 2; ++qq) {
        const int q = 2 * hf + qq;
        const int il = 16 * wt + 4 * lg + q;
        float xo[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) xo[k] = *slot(k, il);
        double o2[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = ic + il + h * H;
          const float sg = h == 0 ? 1.0f : -1.0f;
          double x4[4];
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            const float e = qq == 0 ? st[2 * f].x : st[2 * f].y;
            const float o = qq == 0 ? st[2 * f + 1].x : st[2 * f + 1].y;
            x4[f] = e + sg * o;
          }
          if (a.test == 1) {
            x4[2] = a.uvic[((long long)b * 2 + 0) * plane + (long long)j * I + i];
            x4[3] = a.uvic[((long long)b * 2 + 1) * plane + (long long)j * I + i];
          }
          double pbar[8];
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            const float e = hf ? cE[f][2 + qq] : cE[f][qq];
            const float o = hf ? cO[f][2 + qq] : cO[f][qq];
            const double own = e + sg * o;
            const double oth = xo[2 * f] + sg * xo[2 * f + 1];
            pbar[f] = hf ? oth : own;
            pbar[4 + f] = hf ? own : oth;
          }
          grid_products_adjoint(cf, geo, gl, j, i, a.test, x4[0], x4[1], x4[2], x4[3], pbar, o2[h], dpar);
        }
        // The pinned winds of test 1 do not depend on the state.
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const bool zero = a.test == 1 && f >= 2;
          *slot(f, il) = zero ? 0.0f : (float)(o2[0][f] + o2[1][f]);
          *slot(4 + f, il) = zero ? 0.0f : (float)(o2[0][f] - o2[1][f]);
        }
      }
    }
    lds_barrier();
    // ---- Phase 2: GEMM 2 partial sums, wave w owning cotangent field w & 3 and this half's
    // output tiles; even-m tiles read s, odd-m tiles d.
#pragma unroll 1
    for (int gg = 0; gg < ntc; ++gg) {
      v4f ac[TH];
#pragma unroll
      for (int t = 0; t < TH; ++t) ac[t] = an[t];
      if (gg + 1 < ntc) {
#pragma unroll
        for (int t = 0; t < TH; ++t)
          if (t2 + t < NT2) an[t] = Tf[((long long)(t2 + t) * (H / 16) + ic / 16 + gg + 1) * 64];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int il = 16 * gg + 4 * q + lg;
        const float bs = sS[(wt * ICH + il) * 16 + ln];
        const float bd = sD[(wt * ICH + il) * 16 + ln];
#pragma unroll
        for (int t = 0; t < TH; ++t)
          if (t2 + t < NT2)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], t2 + t < TE ? bs : bd, acc[t], 0, 0, 0);
      }
    }
    lds_barrier();
  }

  // Per-row parameter partials: row ln is held by lanes ln + 16 g of every wave.
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = dpar[c];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    dpar[c] = v;
  }
  if (lg == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) sRed[(wave * 16 + ln) * 3 + c] = dpar[c];
  lds_barrier();
  if (tid < 16 * 3) {
    const int r = tid / 3, c = tid % 3;
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sum += sRed[(w * 16 + r) * 3 + c];
    a.pgrad[((long long)b * J + j0 + r) * 3 + c] += sum;
  }

  // Gbar = I w_m (Tfwd xbar) in the folded output rows: wave w -> field eta, Phi, U, V of w & 3.
  const int fo = wt == 0 ? 0 : wt + 1;
#pragma unroll
  for (int t = 0; t < TH; ++t) {
    const int tg = t2 + t;
    if (tg >= NT2) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = 16 * (tg < TE ? tg : tg - TE) + 4 * lg + q;
      const int m = tg < TE ? 2 * (rr >> 1) : 2 * (rr >> 1) + 1, c = rr & 1;
      if (m < Mp) {
        const float v = (m == 0 && c == 1) ? 0.0f : (float)I * (m == 0 ? 1.0f : 2.0f) * acc[t][q];
        a.Gbar[fidx(b, fo, m, j, Mp, J) + c] = v;
      }
    }
  }
  // delta's cotangent is zero (the products do not depend on delta).
  for (int p = tid; p < Mp * 16 * 2; p += NTH) {
    const int m = p >> 5, r = (p >> 1) & 15, c = p & 1;
    a.Gbar[fidx(b, 1, m, j0 + r, Mp, J) + c] = 0.0f;
  }
}

template <int MT>
__global__ void __launch_bounds__(512) k_grid_adj_mfma_fold_sc8(GridAdjArgs<float> a, const float* __restrict__ TinvF,
                                                                const float* __restrict__ TfwdF) {
  grid_adj_mfma_fold_sc8_body<MT>(a, TinvF, TfwdF);
}

// The S_t-cached eight-wave form per truncation: T106 (whose 12-field staging would not fit the
// LDS) and T63 (one workgroup per CU either way; two waves per SIMD instead of one); T42 keeps
// two four-wave workgroups per CU.
template <int MT>
constexpr bool adj_fold_cached() { return MT != 42; }

template <int MT>
static hipError_t launch_adj_mfma_fold_t(const GridAdjArgs<float>& a, const float* TinvH, const float* TfwdH,
                                         hipStream_t s) {
  constexpr bool SC = adj_fold_cached<MT>();
  if (SC && a.stc == nullptr) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((long long)a.B * a.J / 16));
  if constexpr (SC) {
    const size_t smem = sizeof(float) * AdjFoldSC8Shape<MT>::LDS_FLOATS;
    hipLaunchKernelGGL(k_grid_adj_mfma_fold_sc8<MT>, grid, dim3(512), smem, s, a, TinvH, TfwdH);
  } else if constexpr (MT == 42) {
    hipLaunchKernelGGL(k_grid_adj_mfma_fold_w2<MT>, grid, dim3(256), sizeof(float) * AdjFoldShape<MT>::LDS_FLOATS, s,
                       a, TinvH, TfwdH);
  } else {
    hipLaunchKernelGGL(k_grid_adj_mfma_fold<MT>, grid, dim3(256), sizeof(float) * AdjFoldShape<MT>::LDS_FLOATS, s, a,
                       TinvH, TfwdH);
  }
  return hipGetLastError();
}

template <int MT>
constexpr size_t adj_fold_lds_bytes() {
  if constexpr (adj_fold_cached<MT>())
    return sizeof(float) * AdjFoldSC8Shape<MT>::LDS_FLOATS;
  else
    return sizeof(float) * AdjFoldShape<MT>::LDS_FLOATS;
}

size_t grid_adj_mfma_fold_lds_bytes(int M) {
  switch (M) {
    case 42: return adj_fold_lds_bytes<42>();
    case 63: return adj_fold_lds_bytes<63>();
    case 106: return adj_fold_lds_bytes<106>();
    default: return 0;
  }
}

size_t grid_adj_mfma_fold_cache_bytes(int M, int B, int J) {
  const long long nblk = (long long)B * J / 16;
  switch (M) {
    case 42: return sizeof(float) * (size_t)AdjFoldShape<42, adj_fold_cached<42>()>::cache_floats(nblk);
    case 63: return sizeof(float) * (size_t)AdjFoldShape<63, adj_fold_cached<63>()>::cache_floats(nblk);
    case 106: return sizeof(float) * (size_t)AdjFoldShape<106, adj_fold_cached<106>()>::cache_floats(nblk);
    default: return 0;
  }
}

hipError_t launch_grid_adj_mfma_fold(const GridAdjArgs<float>& a, const float* TinvH, const float* TfwdH,
                                     hipStream_t s) {
  switch (a.M) {
    case 42: return launch_adj_mfma_fold_t<42>(a, TinvH, TfwdH, s);
    case 63: return launch_adj_mfma_fold_t<63>(a, TinvH, TfwdH, s);
    case 106: return launch_adj_mfma_fold_t<106>(a, TinvH, TfwdH, s);
    default: return hipErrorInvalidValue;
  }
}

template <int MT>
static hipError_t launch_adj_mfma_t(const GridAdjArgs<float>& a, const float* TinvF, const float* TfwdF,
                                    hipStream_t s) {
  const size_t smem = sizeof(float) * AdjShape<MT>::LDS_FLOATS;
  hipLaunchKernelGGL(k_grid_adj_mfma<MT>, dim3((unsigned)((long long)a.B * a.J / 16)), dim3(256), smem, s, a, TinvF,
                     TfwdF);
  return hipGetLastError();
}

size_t grid_adj_mfma_lds_bytes(int M) {
  switch (M) {
    case 42: return sizeof(float) * AdjShape<42>::LDS_FLOATS;
    case 63: return sizeof(float) * AdjShape<63>::LDS_FLOATS;
    default: return 0;  // T106: the 12-field staging exceeds the workgroup's LDS
  }
}

hipError_t launch_grid_adj_mfma(const GridAdjArgs<float>& a, const float* TinvF, const float* TfwdF, hipStream_t s) {
  switch (a.M) {
    case 42: return launch_adj_mfma_t<42>(a, TinvF, TfwdF, s);
    case 63: return launch_adj_mfma_t<63>(a, TinvF, TfwdF, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace swx
