// Matrix-core Legendre adjoint (FP32 transforms, FP64 spectral arithmetic): the reverse of
// k_leg_mfma, with the spectral pieces of the plain k_leg_adj (swx_adjoint_common.h).
//
//   mode A (grid (M+1, member groups)):  per (member, n) the carry adjoint and the
//     cotangents of the Legendre inputs (FP64), and the inputs of the recomputed inverse
//     transform of S_t; then one inverse-shaped GEMM over n,
//       D[j][col] = sum_n B[j][n] X[n][col]  (B = P for the P columns, H for the H columns),
//     whose columns hold, per member, the 5 P-slot cotangents and eta, delta, Phi and the two
//     wind potentials (P; 20 reals), and the 4 H-slot cotangents and the two wind potentials
//     (H; 12 reals).  Writes the slot cotangents Tbar and the recomputed Fourier rows G_t.
//   modes B / TERM:  the Fourier cotangents of 5 fields (P columns, 10 reals per member) and
//     of U, V (H columns, 4 reals) through one forward-shaped GEMM over j,
//       C[n][col] = sum_j B^T[n][j] Gbar[j][col],
//     then lam (+)= the wind-inversion adjoint per (member, n) in FP64.
//
// The basis operands come from the forward kernel's fragment-ordered tables (LegArgs):
// inv[m][jt][lane][ks] for mode A, fwd[m][nt][lane][ks] for mode B.  MFMA tile
// v_mfma_f32_16x16x4f32: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16], and
// holds D[4 (l / 16) + q][l % 16].
#include "swx_adjoint.h"
#include "swx_adjoint_common.h"

namespace swx {

typedef float v4f __attribute__((ext_vector_type(4)));

namespace {
constexpr int c16(int x) { return (x + 15) / 16 * 16; }
constexpr int cmaxi(int a, int b) { return a > b ? a : b; }
}  // namespace

template <int MT, int MB>
struct LegAdjShape {
  static constexpr int Np = MT + 1;
  static constexpr int J = MT == 42 ? 64 : (MT == 63 ? 96 : 160);
  static constexpr int KNP = c16(Np);
  // mode A
  static constexpr int XP = c16(20 * MB), XH = c16(12 * MB), XT = XP + XH;
  static constexpr int XS = XT + 1;  // odd row stride: the per-degree rows start in different banks
  static constexpr int JO = J + 4;   // column stride of the GEMM results
  // Mode A: the GEMM results overwrite its B operand once every wave has read it.
  static constexpr int LDS_A = cmaxi(KNP * XS, XT * JO);
  // modes B / TERM
  static constexpr int RP = c16(10 * MB), RH = c16(4 * MB), RT = RP + RH;
  static constexpr int RS = RT + 1;
  static constexpr int LDS_B = cmaxi(J, KNP) * RS;
};

template <int MT, int MB>
__global__ void __launch_bounds__(256) k_leg_adj_a_mfma(LegAdjArgs<float> a) {
  using S = LegAdjShape<MT, MB>;
  constexpr int Np = S::Np, Mp = MT + 1, J = S::J, KNP = S::KNP, XP = S::XP, XT = S::XT, XS = S::XS, JO = S::JO;
  static_assert(KNP == leg_knp(MT), "fragment tables use leg_knp rows");
  extern __shared__ __align__(16) float lds[];
  float* sX = lds;                 // [KNP][XS] GEMM B operand
  float* sO = lds;                 // [XT][JO] results, column-major (after the GEMM, over sX)
  int m, grp;
  leg_item(1, m, grp);
  const int b0 = grp * MB;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int B = a.B;
  const int K = Np - m;
  const long long fst = (long long)Mp * Np * 2;
  const CoefLayout cl = CoefLayout::make(MT, J);

  static_assert(KNP % 4 == 0, "sX is zeroed as float4s");
  for (int e = tid; e < KNP * XS / 4; e += 256) reinterpret_cast<v4f*>(sX)[e] = v4f{0.f, 0.f, 0.f, 0.f};
  __syncthreads();
  // Spectral part (FP64) per (member, n).
  for (int p = tid; p < MB * K; p += 256) {
    const int mb = p / K, np = p % K, b = b0 + mb;
    if (b >= B) continue;
    const int n = m + np;
    const double* cf = a.coef + (long long)b * a.ncoef;
    const long long so = ((long long)m * Np + n) * 2;
    double slot[kNT][2], x[5][2];
    leg_adj_carry(cf, cl, a.f2[(long long)m * Np + n], n, a.diffflag, a.Lam + (long long)b * 3 * fst, so, fst, slot);
    leg_adj_recompute_inputs(cf, cl, m, n, a.St + (long long)b * 3 * fst, so, fst, x);
    float* xr = sX + np * XS;
    for (int c = 0; c < 2; ++c) {
      for (int k = 0; k < kNP; ++k) xr[mb * 20 + 2 * k + c] = (float)slot[k][c];
      for (int f = 0; f < 5; ++f) xr[mb * 20 + 10 + 2 * f + c] = (float)x[f][c];
      for (int k = 0; k < kNH; ++k) xr[XP + mb * 12 + 2 * k + c] = (float)slot[kNP + k][c];
      xr[XP + mb * 12 + 8 + c] = (float)x[4][c];   // eta potential (H part of U)
      xr[XP + mb * 12 + 10 + c] = (float)x[3][c];  // delta potential (H part of V)
    }
  }
  __syncthreads();

  // Inverse-shaped GEMM: wave w owns latitude tiles w, w+4, ... and every column tile; the
  // tiles' results stay in registers until every wave has finished reading sX.
  constexpr int CIP = XP / 16, CIH = (XT - XP) / 16;
  constexpr int KSM = KNP / 4;
  constexpr int NJT = (J / 16 + 3) / 4;  // latitude tiles per wave
  const int ksn = (K + 3) / 4;
  v4f dP[NJT][CIP], dH[NJT][CIH];
#pragma unroll
  for (int u = 0; u < NJT; ++u) {
    const int jt = wave + 4 * u;
#pragma unroll
    for (int c = 0; c < CIP; ++c) dP[u][c] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < CIH; ++c) dH[u][c] = v4f{0.f, 0.f, 0.f, 0.f};
    if (jt >= J / 16) continue;
    const long long fo = ((long long)(m * (J / 16) + jt) * 64 + lane) * KSM;
    const v4f* pr = reinterpret_cast<const v4f*>(a.invP + fo);
    const v4f* hr = reinterpret_cast<const v4f*>(a.invH + fo);
    v4f pa[KSM / 4], ha[KSM / 4];
#pragma unroll
    for (int q4 = 0; q4 < KSM / 4; ++q4) {
      const bool need = 4 * q4 < ksn;
      pa[q4] = need ? pr[q4] : v4f{0.f, 0.f, 0.f, 0.f};
      ha[q4] = need ? hr[q4] : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int ks = 0; ks < KSM; ++ks) {
      if (ks < ksn) {
        const float* xr = sX + (4 * ks + lg) * XS + ln;
#pragma unroll
        for (int c = 0; c < CIP; ++c)
          dP[u][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[ks / 4][ks % 4], xr[c * 16], dP[u][c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < CIH; ++c)
          dH[u][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ha[ks / 4][ks % 4], xr[XP + c * 16], dH[u][c], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // sX is no longer read: the results overwrite it
#pragma unroll
  for (int u = 0; u < NJT; ++u) {
    const int jt = wave + 4 * u;
    if (jt >= J / 16) continue;
    float* out = sO + ln * JO + jt * 16 + 4 * lg;
#pragma unroll
    for (int c = 0; c < CIP; ++c) *reinterpret_cast<v4f*>(out + c * 16 * JO) = dP[u][c];
#pragma unroll
    for (int c = 0; c < CIH; ++c) *reinterpret_cast<v4f*>(out + (XP + c * 16) * JO) = dH[u][c];
  }
  __syncthreads();

  // Slot cotangents Tbar [member][m][slot][j] and the recomputed Fourier rows G_t (m-major).
  const float mr = (float)m;
  const long long fs = (long long)Mp * J * 2;
  for (int p = tid; p < MB * J; p += 256) {
    const int mb = p / J, j = p % J, b = b0 + mb;
    if (b >= B) continue;
    const float* pc = sO + mb * 20 * JO + j;
    const float* hc = sO + (XP + mb * 12) * JO + j;
    float* Tb = a.Tbar + (((long long)b * Mp + m) * kNT) * J * 2;
#pragma unroll
    for (int k = 0; k < kNP; ++k) {
      Tb[(k * J + j) * 2] = pc[(2 * k) * JO];
      Tb[(k * J + j) * 2 + 1] = pc[(2 * k + 1) * JO];
    }
#pragma unroll
    for (int k = 0; k < kNH; ++k) {
      Tb[((kNP + k) * J + j) * 2] = hc[(2 * k) * JO];
      Tb[((kNP + k) * J + j) * 2 + 1] = hc[(2 * k + 1) * JO];
    }
    const float* q = pc + 10 * JO;  // eta, delta, Phi, delta potential, eta potential
    float* G = a.G + fidx(b, 0, m, j, Mp, J);
    G[0] = q[0];
    G[1] = q[JO];
    G[fs] = q[2 * JO];
    G[fs + 1] = q[3 * JO];
    G[2 * fs] = q[4 * JO];
    G[2 * fs + 1] = q[5 * JO];
    // U_m = -i m P(delta hat) + H(eta hat);  V_m = -i m P(eta hat) - H(delta hat)
    G[3 * fs] = mr * q[7 * JO] + hc[8 * JO];
    G[3 * fs + 1] = -mr * q[6 * JO] + hc[9 * JO];
    G[4 * fs] = mr * q[9 * JO] - hc[10 * JO];
    G[4 * fs + 1] = -mr * q[8 * JO] - hc[11 * JO];
  }
}

template <int MT, int MB>
__global__ void __launch_bounds__(256) k_leg_adj_b_mfma(LegAdjArgs<float> a) {
  using S = LegAdjShape<MT, MB>;
  constexpr int Np = S::Np, Mp = MT + 1, J = S::J, KNP = S::KNP, RP = S::RP, RT = S::RT, RS = S::RS;
  static_assert(KNP == leg_knp(MT), "fragment tables use leg_knp rows");
  extern __shared__ __align__(16) float lds[];
  float* sA = lds;  // [J][RS] Fourier cotangents, then [KNP][RS] contractions
  int m, grp;
  leg_item(1, m, grp);
  const int b0 = grp * MB;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int B = a.B;
  const int K = Np - m;
  const int Kt = (K + 15) / 16;
  const long long fst = (long long)Mp * Np * 2;
  const CoefLayout cl = CoefLayout::make(MT, J);

  // Stage Gbar rows [j][member, field, re/im] (P columns) and the U, V rows again (H columns).
  // Each (member, field) row is contiguous in j: float4 loads (two latitudes), issued first.
  {
    constexpr int QPM = kNG * J / 2;  // float4 per member
    constexpr int NQ = (MB * QPM + 255) / 256;
    v4f gv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int q = tid + 256 * u;
      const int mb = q / QPM < MB ? q / QPM : MB - 1;
      const int b = b0 + mb < B ? b0 + mb : B - 1;  // clamped address, zeroed below
      const int r = q % QPM, f = r / (J / 2), j = 2 * (r % (J / 2));
      const v4f v = *reinterpret_cast<const v4f*>(a.Gbar + fidx(b, f, m, j, Mp, J));
      gv[u] = (q < MB * QPM && b0 + q / QPM < B) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int q = tid + 256 * u;
      if (q < MB * QPM) {
        const int mb = q / QPM, r = q % QPM;
        const int f = r / (J / 2), j = 2 * (r % (J / 2));
        float* d0 = sA + j * RS + mb * 10 + 2 * f;
        d0[0] = gv[u][0];
        d0[1] = gv[u][1];
        d0[RS] = gv[u][2];
        d0[RS + 1] = gv[u][3];
        if (f >= 3) {
          float* d1 = sA + j * RS + RP + mb * 4 + 2 * (f - 3);
          d1[0] = gv[u][0];
          d1[1] = gv[u][1];
          d1[RS] = gv[u][2];
          d1[RS + 1] = gv[u][3];
        }
      }
    }
  }
  constexpr int padH = RT - RP - MB * 4, padP = RP - MB * 10;
  if constexpr (padH > 0)
    for (int e = tid; e < J * padH; e += 256) sA[(e / padH) * RS + RP + MB * 4 + e % padH] = 0.0f;
  if constexpr (padP > 0)
    for (int e = tid; e < J * padP; e += 256) sA[(e / padP) * RS + MB * 10 + e % padP] = 0.0f;
  __syncthreads();

  // Forward-shaped GEMM tiles (Kt row tiles x RT/16 column tiles).
  constexpr int CT = RT / 16;
  constexpr int MAXT = (KNP / 16 * CT + 3) / 4;
  v4f acc[MAXT];
#pragma unroll
  for (int q = 0; q < MAXT; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
  const int ntile = Kt * CT;
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    const int t = wave + 4 * q;
    if (t < ntile) {
      const int nt16 = t / CT, ct = t % CT;
      const bool isP = ct * 16 < RP;
      const v4f* ar = reinterpret_cast<const v4f*>((isP ? a.fwdP : a.fwdH) +
                                                   ((long long)(m * (KNP / 16) + nt16) * 64 + lane) * (J / 4));
      v4f av[J / 16];
#pragma unroll
      for (int q4 = 0; q4 < J / 16; ++q4) av[q4] = ar[q4];
#pragma unroll
      for (int ks = 0; ks < J / 4; ++ks)
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ks / 4][ks % 4], sA[(4 * ks + lg) * RS + ct * 16 + ln],
                                                      acc[q], 0, 0, 0);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    const int t = wave + 4 * q;
    if (t < ntile) {
      const int nt16 = t / CT, ct = t % CT;
#pragma unroll
      for (int r = 0; r < 4; ++r) sA[(nt16 * 16 + 4 * lg + r) * RS + ct * 16 + ln] = acc[q][r];
    }
  }
  __syncthreads();

  // lam (+)= wind-inversion adjoint (FP64) per (member, n).
  for (int p = tid; p < MB * K; p += 256) {
    const int mb = p / K, np = p % K, b = b0 + mb;
    if (b >= B) continue;
    const int n = m + np;
    const float* cr = sA + np * RS;
    double pd[10], hd[4];
    for (int k = 0; k < 10; ++k) pd[k] = (double)cr[mb * 10 + k];
    for (int k = 0; k < 4; ++k) hd[k] = (double)cr[RP + mb * 4 + k];
    leg_adj_lambda(a.coef + (long long)b * a.ncoef, cl, m, n, pd, hd, a.mode == LADJ_TERM,
                   a.Lam + (long long)b * 3 * fst, ((long long)m * Np + n) * 2, fst);
  }
}

// ----------------------------------------------------------------------------
// Launchers
// ----------------------------------------------------------------------------

namespace {

// Members per workgroup: mode A carries 32 GEMM columns per member and keeps both LDS
// buffers, so it takes fewer members than mode B (14 columns per member).
constexpr int mb_a(int M) { return M == 42 ? 4 : 2; }
constexpr int mb_b(int M) { return M == 42 ? 8 : 4; }

template <int MT>
hipError_t launch_mt(const LegAdjArgs<float>& a, hipStream_t s) {
  constexpr int MA = mb_a(MT), MBB = mb_b(MT);
  if (a.mode == LADJ_A) {
    const size_t lds = sizeof(float) * LegAdjShape<MT, MA>::LDS_A;
    dim3 grid(MT + 1, (a.B + MA - 1) / MA);
    hipLaunchKernelGGL((k_leg_adj_a_mfma<MT, MA>), grid, dim3(256), lds, s, a);
  } else {
    const size_t lds = sizeof(float) * LegAdjShape<MT, MBB>::LDS_B;
    dim3 grid(MT + 1, (a.B + MBB - 1) / MBB);
    hipLaunchKernelGGL((k_leg_adj_b_mfma<MT, MBB>), grid, dim3(256), lds, s, a);
  }
  return hipGetLastError();
}

template <int MT>
size_t lds_mt() {
  return sizeof(float) * cmaxi(LegAdjShape<MT, mb_a(MT)>::LDS_A, LegAdjShape<MT, mb_b(MT)>::LDS_B);
}

}  // namespace

size_t leg_adj_mfma_lds_bytes(int M) {
  switch (M) {
    case 42: return lds_mt<42>();
    case 63: return lds_mt<63>();
    case 106: return lds_mt<106>();
    default: return 0;
  }
}

hipError_t launch_leg_adj_mfma(const LegAdjArgs<float>& a, hipStream_t s) {
  if (a.J != (a.M == 42 ? 64 : (a.M == 63 ? 96 : 160)) || a.N != a.M) return hipErrorInvalidValue;
  switch (a.M) {
    case 42: return launch_mt<42>(a, s);
    case 63: return launch_mt<63>(a, s);
    case 106: return launch_mt<106>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace swx
