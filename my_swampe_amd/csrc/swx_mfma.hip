// MFMA grid kernel (FP32 transforms): inverse DFT -> grid physics -> forward
// DFT for a group of 16 latitude rows, as two back-to-back matrix-core GEMMs
// with the grid-space physics applied in registers between them.
//
//   GEMM 1 (inverse DFT, spectral_transform.py:331-348 + real part):
//     x[f][i][n] = sum_k Tinv[i][k] X[f][k][n],  k = 2m + (re, im),
//     Tinv[i][2m] = w_m cos(2 pi m i / I), Tinv[i][2m+1] = -w_m sin(...),
//     w_0 = 1, w_m = 2 (Hermitian layout, index-based phase).
//   grid physics (forcing.py, initial_conditions.py:283-292, model.py:988-1006):
//     the FP64 bookkeeping of the physical carry, residue and pinned winds where a
//     step needs it (swx_grid_common.h:grid_bookkeeping4), then the 8 products in
//     FP32 (grid_products_f32), the precision of the transforms around them
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

#include "swx_grid_common.h"

namespace swx {

template <int MT>
struct MfmaShape {
  static constexpr int Mp = MT + 1;
  static constexpr int I = MT == 42 ? 128 : (MT == 63 ? 192 : 320);
  static constexpr int K1 = 2 * Mp;                  // inverse-DFT inner dimension
  static constexpr int K1p = (K1 + 3) / 4 * 4;       // padded to the MFMA k-step
  static constexpr int KQ1 = (K1p / 4 + 3) / 4 * 4;  // per-lane inverse-table run (float4 loads)
  static constexpr int K2 = 2 * Mp;                  // forward-DFT outputs (re, im)
  static constexpr int K2p = (K2 + 15) / 16 * 16;    // padded to 16-row tiles
  static constexpr int NT2 = K2p / 16;
  static constexpr int ICH = 64;                     // longitudes per chunk
  static constexpr int ROWS = 16;
  // LDS floats: X^T [5][K1p][16] + Y^T [8][ICH][16]; the final exchange
  // buffer [2][8][16][16] reuses the Y^T region.
  static constexpr int LDS_FLOATS = 5 * K1p * 16 + 8 * ICH * 16;
};

// One workgroup's work; `blk` = its 16-row block (member * J / 16 + latitude block).
template <int MT>
__device__ __forceinline__ void grid_mfma_body(const GridArgs<float>& a, const float* __restrict__ TinvF,
                                               const float* __restrict__ TfwdF, const int blk, float* lds) {
  using S = MfmaShape<MT>;
  constexpr int Mp = S::Mp, I = S::I, K1p = S::K1p, NT2 = S::NT2, ICH = S::ICH, KQ1 = S::KQ1;
  static_assert(ICH % 16 == 0 && I % ICH == 0, "chunking");
  float* sXT = lds;                    // [5][K1p][16]
  float* sYT = lds + 5 * K1p * 16;     // [8][ICH][16]
  float* sC = sYT;                     // assembly tiles (kAsmFloats) after the last chunk
  static_assert(8 * S::ICH * 16 >= kAsmFloats, "assembly tiles fit the product region");

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ln = lane & 15, lg = lane >> 4;
  const int J = a.J, N = a.N;
  const int row0 = blk * S::ROWS;
  const int b = row0 / J, j0 = row0 % J;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;
  const long long plane = (long long)J * I;
  double* phys = a.phys + (long long)b * 8 * plane;
  const double alpha = cf[C_ALPHA];

  // Phase 0: Fourier rows (m-major, 16 consecutive j per (f, m)) -> X^T (k-major, zero padded).
  // One float4 = two latitudes x (re, im); all loads are issued before the LDS writes.
  {
    constexpr int NE = 5 * (K1p / 2) * 8;
    constexpr int NQ = (NE + 255) / 256;
    v4f gv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      const int jq = e & 7, fm = e >> 3;
      const int f = fm / (K1p / 2) < 5 ? fm / (K1p / 2) : 4, m = fm % (K1p / 2);
      const v4f v = *reinterpret_cast<const v4f*>(a.G + fidx(b, f, m < Mp ? m : Mp - 1, j0 + 2 * jq, Mp, J));
      gv[u] = (e < NE && m < Mp) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      if (e < NE) {
        const int jq = e & 7, fm = e >> 3;
        const int f = fm / (K1p / 2), m = fm % (K1p / 2);
        float* xr = sXT + (f * K1p + 2 * m) * 16 + 2 * jq;
        *reinterpret_cast<float2*>(xr) = make_float2(gv[u][0], gv[u][2]);
        *reinterpret_cast<float2*>(xr + 16) = make_float2(gv[u][1], gv[u][3]);
      }
    }
  }
  __syncthreads();

  v4f acc[2 * NT2];
  float2 uv[(NT2 + 1) / 2][2];
  assemble_prefetch_uv<Mp, NT2, 0>(a, b, j0, uv);
#pragma unroll
  for (int p = 0; p < 2 * NT2; ++p) acc[p] = v4f{0.f, 0.f, 0.f, 0.f};

  const int n = ln;
  const int j = j0 + n;
  const PointConsts pc = point_consts(cf, geo, gl, j);
  for (int ic = 0; ic < I; ic += ICH) {
    // ---- Phase 1: inverse DFT of one 16-wide i-tile, all five fields.  The
    // wave's A fragments (its i-tile of Tinv) are loaded up front.
    const int ib = ic + 16 * wave;
    v4f c1[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) c1[f] = v4f{0.f, 0.f, 0.f, 0.f};
    v4f afr[KQ1 / 4];
    {
      const v4f* Ta = reinterpret_cast<const v4f*>(TinvF + ((long long)(ib / 16) * 64 + lane) * KQ1);
#pragma unroll
      for (int q4 = 0; q4 < KQ1 / 4; ++q4) afr[q4] = Ta[q4];
    }
    // First k-group of this chunk's forward-DFT fragments, in flight during GEMM 1 and the physics.
    const v4f* Tf = reinterpret_cast<const v4f*>(TfwdF) + lane;
    v4f an[NT2];
#pragma unroll
    for (int t = 0; t < NT2; ++t) an[t] = Tf[((long long)t * (I / 16) + ic / 16) * 64];
#pragma unroll
    for (int ks = 0; ks < K1p / 4; ++ks) {
      const float* xb = sXT + (4 * ks + lg) * 16 + ln;
#pragma unroll
      for (int f = 0; f < 5; ++f)
        c1[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks / 4][ks % 4], xb[f * K1p * 16], c1[f], 0, 0, 0);
    }
    // ---- Grid physics on this lane's 4 points (row n, i = ib + 4 lg + q).
    {
      float fe[4], fd[4], fp[4], fu[4], fv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        fe[q] = c1[0][q];
        fd[q] = c1[1][q];
        fp[q] = c1[2][q];
        fu[q] = c1[3][q];
        fv[q] = c1[4][q];
      }
      grid_bookkeeping4(a, cf, phys, plane, alpha, b, j, ib + 4 * lg, fe, fd, fp, fu, fv);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int il = 16 * wave + 4 * lg + q;
        float prod[8];
        grid_point_step(a, geo, gl, pc, ic + il, fe[q], fp[q], fu[q], fv[q], prod);
#pragma unroll
        for (int f = 0; f < 8; ++f) sYT[(f * ICH + il) * 16 + n] = prod[f];
      }
    }
    __syncthreads();
    // ---- Phase 2: forward DFT partial sums over this chunk, one k-group (16
    // longitudes, one float4 of A per 16-row tile) at a time; the next group's
    // fragments are fetched while the current group's MFMAs run.
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
        const int il = 16 * gg + 4 * q + lg;
        const float b0 = sYT[((2 * wave) * ICH + il) * 16 + ln];
        const float b1 = sYT[((2 * wave + 1) * ICH + il) * 16 + ln];
#pragma unroll
        for (int t = 0; t < NT2; ++t) {
          acc[2 * t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], b0, acc[2 * t], 0, 0, 0);
          acc[2 * t + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], b1, acc[2 * t + 1], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }

  // ---- Phase 3: assemble the Legendre inputs (swx_grid_common.h).
  grid_assemble<Mp, NT2>(a, acc, sC, b, j0, cf, cl, geo, gl, uv);
}

// ----------------------------------------------------------------------------
// Half-period folded variant.  With H = I/2, e^{2 pi i m (i + H) / I} = (-1)^m e^{2 pi i m i / I},
// so both GEMMs need only the longitudes i < H:
//   inverse  x(i) = E(i) + O(i),  x(i + H) = E(i) - O(i),  E / O = the sums over even / odd m;
//   forward  Y[m] = sum_{i<H} F[m][i] (y(i) + (-1)^m y(i + H)).
// The k rows (m, re/im) are reordered even-m block first (KE4 rows), then the odd-m block, for
// both GEMMs; each lane runs the physics of the point pair (i, i + H) and stores
// s = y(i) + y(i + H) and d = y(i) - y(i + H), which the even / odd output tiles consume.
// Half the matrix-core work of grid_mfma_body.  Tables (fragment order as there):
//   inv[it][lane][ks] = Tinv(i = 16 it + ln, k'(4 ks + lg)),  ks < KQ1
//   fwd[t][g][lane][q] = Tfwd(k'(16 t + ln), i = 16 g + 4 q + lg),  g < H / 16
// with k' the folded order (grid_mfma_fold_row).
template <int MT>
struct MfmaFoldShape {
  static constexpr int Mp = MT + 1;
  static constexpr int I = MfmaShape<MT>::I;
  static constexpr int H = I / 2;
  static constexpr int NE = (Mp + 1) / 2, NO = Mp / 2;    // even / odd wavenumbers
  static constexpr int KE4 = (2 * NE + 3) / 4 * 4, KO4 = (2 * NO + 3) / 4 * 4;
  static constexpr int K1 = KE4 + KO4;                    // inverse inner dimension (folded order)
  static constexpr int KQ1 = (K1 / 4 + 3) / 4 * 4;
  static constexpr int TE = (2 * NE + 15) / 16, TO = (2 * NO + 15) / 16;
  static constexpr int NT2 = TE + TO;                     // forward output tiles
  static constexpr int ICH = 64;                          // half-period longitudes per chunk
  static constexpr int NCH = (H + ICH - 1) / ICH;
  static constexpr int ROWS = 16;
  static constexpr int XT_FLOATS = 5 * K1 * 16;
  static constexpr int SD_FLOATS = 2 * 8 * ICH * 16;
  // One chunk (T42): X^T is dead once GEMM 1 is done, so s / d reuse its space.
  static constexpr int LDS_FLOATS = NCH == 1 ? (XT_FLOATS > SD_FLOATS ? XT_FLOATS : SD_FLOATS)
                                             : XT_FLOATS + SD_FLOATS;
  // Fast variant: s goes through LDS first (even-m tiles), then d in the same buffer (odd-m
  // tiles) while the lanes hold d in registers: half the product region.
  static constexpr int S1_FLOATS = 8 * ICH * 16 > kAsmFloats ? 8 * ICH * 16 : kAsmFloats;
  static constexpr int LDS_FAST = NCH == 1 ? (XT_FLOATS > S1_FLOATS ? XT_FLOATS : S1_FLOATS) : XT_FLOATS + S1_FLOATS;
};

// Wave priority P over the folded kernel's MFMA loops (1 in, 0 after): the sequencer then
// issues for a wave in a GEMM ahead of waves in the physics or the staging, so the matrix cores
// stay fed while those catch up.  Measured per launch of the steady (FAST) variant: T42 x 256
// 37.7 -> 35.9 us, T106 x 64 148 -> 140 us; T63 x 128 (2 workgroups per CU) 57 -> 68 us, so T63
// keeps the default, and so does the bookkeeping variant (T42: 71.7 -> 73.8 us).
template <int MT, bool FAST, int P>
__device__ __forceinline__ void fold_gemm_prio() {
  if constexpr (MT != 63 && FAST) __builtin_amdgcn_s_setprio(P);
}

// FAST: a step without residue, pinned winds, physical carry, semi-implicit or explicit slots
// (the graph-replayed steps of the modified-Euler scheme): FP32-only point and slot code, and
// the two-round s / d exchange in half the LDS.
// DIAG (with FAST = false): per-step history and diagnostics (grid_bookkeeping4<true>).
template <int MT, bool FAST = false, bool DIAG = false>
__device__ __forceinline__ void grid_mfma_fold_body(const GridArgs<float>& a, const float* __restrict__ TinvF,
                                                    const float* __restrict__ TfwdF, const int blk, float* lds) {
  using S = MfmaFoldShape<MT>;
  constexpr int Mp = S::Mp, I = S::I, H = S::H, K1 = S::K1, KE4 = S::KE4, NT2 = S::NT2, TE = S::TE, ICH = S::ICH;
  constexpr int KQ1 = S::KQ1;
  float* sXT = lds;                                                  // [5][K1][16]
  float* sS = lds + (S::NCH == 1 ? 0 : S::XT_FLOATS);                // [8][ICH][16] s
  float* sD = sS + 8 * ICH * 16;                                     // [8][ICH][16] d
  float* sC = sS;                                                    // assembly tiles after the last chunk
  static_assert(S::SD_FLOATS >= kAsmFloats, "assembly tiles fit the s / d region");

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ln = lane & 15, lg = lane >> 4;
  const int J = a.J, N = a.N;
  const int row0 = blk * S::ROWS;
  const int b = row0 / J, j0 = row0 % J;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;
  const long long plane = (long long)J * I;
  double* phys = a.phys + (long long)b * 8 * plane;
  const double alpha = cf[C_ALPHA];
  // Per-step diagnostics and history (the bookkeeping variant only; model.py:838-864, :1175-1189).
  const bool want_diag = DIAG && (a.diagnostics || a.hist_phiext);
  if (DIAG && a.diagnostics && a.dead[b]) {
    // A frozen member (the whole workgroup: one member's 16 rows) emits its current level and
    // keeps its Legendre inputs (model.py:847-864), as the plain kernel does.
    grid_frozen_rows(a, phys, plane, b, j0, S::ROWS);
    return;
  }
  double diag[4] = {0.0, DBL_MAX, DBL_MAX, -DBL_MAX};

  // Phase 0: Fourier rows -> X^T in the folded k order.  Slot s < K1 / 2 is the row pair
  // (2s, 2s + 1) = (re, im) of m = 2s (s < KE4 / 2) or m = 2 (s - KE4 / 2) + 1.  The steady
  // variant stages eta, Phi, U, V only: its products do not read delta (field 1), whose grid
  // field only the physical carry and the per-step outputs of the bookkeeping variant need.
  constexpr int NF = FAST ? 4 : 5;
  {
    constexpr int NE = NF * (K1 / 2) * 8;
    constexpr int NQ = (NE + 255) / 256;
    v4f gv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      const int jq = e & 7, fs = e >> 3;
      const int fi = fs / (K1 / 2) < NF ? fs / (K1 / 2) : NF - 1, sl = fs % (K1 / 2);
      const int f = FAST && fi > 0 ? fi + 1 : fi;
      const int m = sl < KE4 / 2 ? 2 * sl : 2 * (sl - KE4 / 2) + 1;
      const v4f v = *reinterpret_cast<const v4f*>(a.G + fidx(b, f, m < Mp ? m : Mp - 1, j0 + 2 * jq, Mp, J));
      gv[u] = (e < NE && m < Mp) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      if (e < NE) {
        const int jq = e & 7, fs = e >> 3;
        const int fi = fs / (K1 / 2), sl = fs % (K1 / 2);
        const int f = FAST && fi > 0 ? fi + 1 : fi;
        float* xr = sXT + (f * K1 + 2 * sl) * 16 + 2 * jq;
        *reinterpret_cast<float2*>(xr) = make_float2(gv[u][0], gv[u][2]);
        *reinterpret_cast<float2*>(xr + 16) = make_float2(gv[u][1], gv[u][3]);
      }
    }
  }
  __syncthreads();

  v4f acc[2 * NT2];
  float2 uv[(NT2 + 1) / 2][2];
  assemble_prefetch_uv<Mp, NT2, TE>(a, b, j0, uv);
#pragma unroll
  for (int p = 0; p < 2 * NT2; ++p) acc[p] = v4f{0.f, 0.f, 0.f, 0.f};

  const int n = ln;
  const int j = j0 + n;
  const PointConsts pc = point_consts(cf, geo, gl, j);
  const v4f* Tf = reinterpret_cast<const v4f*>(TfwdF) + lane;
  for (int ic = 0; ic < H; ic += ICH) {
    const int ntc = (H - ic) / 16 < ICH / 16 ? (H - ic) / 16 : ICH / 16;  // 16-longitude tiles in this chunk
    // ---- Phase 1: even and odd inverse sums of this wave's tile, all five fields.
    v4f cE[5], cO[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) cE[f] = cO[f] = v4f{0.f, 0.f, 0.f, 0.f};
    const bool active = wave < ntc;
    const int ib = ic + 16 * wave;
    // Steady variant: the day-side cosines of the lane's 8 longitudes (i, i + H for its 4 points),
    // loaded now so that they arrive under GEMM 1.
    float dcp[FAST ? 4 : 1][2];
    if constexpr (FAST) {
      if (active) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = ic + 16 * wave + 4 * lg + q;
          dcp[q][0] = (float)geo[gl.dcos + i];
          dcp[q][1] = (float)geo[gl.dcos + i + H];
        }
      }
    }
    if (active) {
      v4f afr[KQ1 / 4];
      const v4f* Ta = reinterpret_cast<const v4f*>(TinvF + ((long long)(ib / 16) * 64 + lane) * KQ1);
#pragma unroll
      for (int q4 = 0; q4 < KQ1 / 4; ++q4) afr[q4] = Ta[q4];
      fold_gemm_prio<MT, FAST, 1>();
#pragma unroll
      for (int ks = 0; ks < K1 / 4; ++ks) {
        const float* xb = sXT + (4 * ks + lg) * 16 + ln;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
          if (FAST && f == 1) continue;  // delta: not staged, not used by the steady products
          if (ks < KE4 / 4)
            cE[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks / 4][ks % 4], xb[f * K1 * 16], cE[f], 0, 0, 0);
          else
            cO[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks / 4][ks % 4], xb[f * K1 * 16], cO[f], 0, 0, 0);
        }
      }
      fold_gemm_prio<MT, FAST, 0>();
    }
    // First k-group of this chunk's forward fragments, in flight during the physics.
    v4f an[NT2];
#pragma unroll
    for (int t = 0; t < NT2; ++t) an[t] = Tf[((long long)t * (H / 16) + ic / 16) * 64];
    if constexpr (FAST) {
      if constexpr (S::NCH == 1) __syncthreads();  // s overwrites X^T
      // ---- Physics (FP32 only), s to LDS, d kept in registers.
      float dv[8][4];
      if (active) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int il = 16 * wave + 4 * lg + q;
          float pa[8], pb[8];
          grid_products_f32(pc, a.test, dcp[q][0], cE[0][q] + cO[0][q], cE[2][q] + cO[2][q], cE[3][q] + cO[3][q],
                            cE[4][q] + cO[4][q], pa);
          grid_products_f32(pc, a.test, dcp[q][1], cE[0][q] - cO[0][q], cE[2][q] - cO[2][q], cE[3][q] - cO[3][q],
                            cE[4][q] - cO[4][q], pb);
#pragma unroll
          for (int f = 0; f < 8; ++f) {
            sS[(f * ICH + il) * 16 + n] = pa[f] + pb[f];
            dv[f][q] = pa[f] - pb[f];
          }
        }
      }
      __syncthreads();
      // ---- Phase 2 in two rounds: the even-m tiles over s, then the odd-m tiles over d.
#pragma unroll
      for (int round = 0; round < 2; ++round) {
        const int t0 = round == 0 ? 0 : TE, t1 = round == 0 ? TE : NT2;
        if (round == 1) {
          if (active) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
              for (int f = 0; f < 8; ++f) sS[(f * ICH + 16 * wave + 4 * lg + q) * 16 + n] = dv[f][q];
          }
          __syncthreads();  // (an[t] for the odd-m tiles still holds this chunk's first k-group)
        }
        fold_gemm_prio<MT, FAST, 1>();
#pragma unroll 1
        for (int gg = 0; gg < ntc; ++gg) {
          v4f ac[NT2];
#pragma unroll
          for (int t = 0; t < NT2; ++t)
            if (t >= t0 && t < t1) ac[t] = an[t];
          if (gg + 1 < ntc) {
#pragma unroll
            for (int t = 0; t < NT2; ++t)
              if (t >= t0 && t < t1) an[t] = Tf[((long long)t * (H / 16) + ic / 16 + gg + 1) * 64];
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int il = 16 * gg + 4 * q + lg;
            const float v0 = sS[((2 * wave) * ICH + il) * 16 + ln];
            const float v1 = sS[((2 * wave + 1) * ICH + il) * 16 + ln];
#pragma unroll
            for (int t = 0; t < NT2; ++t) {
              if (t >= t0 && t < t1) {
                acc[2 * t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], v0, acc[2 * t], 0, 0, 0);
                acc[2 * t + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], v1, acc[2 * t + 1], 0, 0, 0);
              }
            }
          }
        }
        fold_gemm_prio<MT, FAST, 0>();
        __syncthreads();
      }
    } else {
      if constexpr (S::NCH == 1) __syncthreads();  // s / d overwrite X^T
      // ---- Physics on the lane's 4 point pairs (row n, i = ib + 4 lg + q and i + H).
      if (active) {
        // Fields of the lane's 4 longitudes in each half period: h = 0 at i, h = 1 at i + H.
        float fe[2][4], fd[2][4], fp[2][4], fu[2][4], fv[2][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          fe[0][q] = cE[0][q] + cO[0][q];
          fd[0][q] = cE[1][q] + cO[1][q];
          fp[0][q] = cE[2][q] + cO[2][q];
          fu[0][q] = cE[3][q] + cO[3][q];
          fv[0][q] = cE[4][q] + cO[4][q];
          fe[1][q] = cE[0][q] - cO[0][q];
          fd[1][q] = cE[1][q] - cO[1][q];
          fp[1][q] = cE[2][q] - cO[2][q];
          fu[1][q] = cE[3][q] - cO[3][q];
          fv[1][q] = cE[4][q] - cO[4][q];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
          grid_bookkeeping4<DIAG>(a, cf, phys, plane, alpha, b, j, ib + 4 * lg + h * H, fe[h], fd[h], fp[h], fu[h],
                                  fv[h], want_diag ? diag : nullptr, cf[cl.n_areac + j], cf[C_AREAPLANET],
                                  geo[gl.cosphi + j]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int il = 16 * wave + 4 * lg + q;
          const int i = ic + il;
          float pa[8], pb[8];
          grid_point_step(a, geo, gl, pc, i, fe[0][q], fp[0][q], fu[0][q], fv[0][q], pa);
          grid_point_step(a, geo, gl, pc, i + H, fe[1][q], fp[1][q], fu[1][q], fv[1][q], pb);
#pragma unroll
          for (int f = 0; f < 8; ++f) {
            sS[(f * ICH + il) * 16 + n] = pa[f] + pb[f];
            sD[(f * ICH + il) * 16 + n] = pa[f] - pb[f];
          }
        }
      }
      __syncthreads();
      // ---- Phase 2: forward sums over this chunk's tiles; wave w owns fields 2w, 2w + 1; the
      // even-m output tiles read s, the odd-m ones d.
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
          const float s0 = sS[((2 * wave) * ICH + il) * 16 + ln];
          const float s1 = sS[((2 * wave + 1) * ICH + il) * 16 + ln];
          const float d0 = sD[((2 * wave) * ICH + il) * 16 + ln];
          const float d1 = sD[((2 * wave + 1) * ICH + il) * 16 + ln];
#pragma unroll
          for (int t = 0; t < NT2; ++t) {
            acc[2 * t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], t < TE ? s0 : d0, acc[2 * t], 0, 0, 0);
            acc[2 * t + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t][q], t < TE ? s1 : d1, acc[2 * t + 1], 0, 0, 0);
          }
        }
      }
      __syncthreads();
    }
  }

  // ---- Phase 3: assemble the Legendre inputs (swx_grid_common.h), folded row order.
  grid_assemble<Mp, NT2, TE, true, FAST>(a, acc, sC, b, j0, cf, cl, geo, gl, uv);
  if (DIAG && want_diag) {
    __syncthreads();
    grid_diag_rows(a, diag, b, j0, reinterpret_cast<double*>(lds));
  }
}

template <int MT, bool FAST, bool DIAG = false>
__global__ void __launch_bounds__(256) k_grid_mfma_fold(GridArgs<float> a, const float* __restrict__ TinvF,
                                                        const float* __restrict__ TfwdF) {
  extern __shared__ __align__(16) float lds[];
  grid_mfma_fold_body<MT, FAST, DIAG>(a, TinvF, TfwdF, blockIdx.x, lds);
}

// T42: the same kernel allocated for at least 2 waves per SIMD.  The compiler then keeps the
// forward-DFT accumulators in ordinary VGPRs instead of AGPRs: the FP64-bookkeeping variant fits
// 2 waves (220 registers instead of 274), the fast variant 4 (114 instead of 156).  A plan
// stepped beside another plan (swx_pipeline) keeps the fast variant at 3 workgroups per CU:
// measured at T42 x 2 x 128 members, the two plans together run 4 % faster that way, while
// one plan alone runs 4 % faster at 4.
template <int MT, bool FAST, bool DIAG = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_grid_mfma_fold_w2(GridArgs<float> a, const float* __restrict__ TinvF, const float* __restrict__ TfwdF) {
  extern __shared__ __align__(16) float lds[];
  grid_mfma_fold_body<MT, FAST, DIAG>(a, TinvF, TfwdF, blockIdx.x, lds);
}

template <int MT>
__global__ void __launch_bounds__(256) k_grid_mfma(GridArgs<float> a, const float* __restrict__ TinvF,
                                                   const float* __restrict__ TfwdF) {
  extern __shared__ __align__(16) float lds[];
  grid_mfma_body<MT>(a, TinvF, TfwdF, blockIdx.x, lds);
}

// Two plans of different truncations in one launch: blocks [0, nblk0) are plan 0's; each plan
// runs its own variant (FA / FB: half-period folded), so it reaches the state it reaches alone.
template <int MA, int MB, bool FA, bool FB>
__global__ void __launch_bounds__(256) k_grid_mfma_pair(GridArgs<float> a0, const float* __restrict__ ti0,
                                                        const float* __restrict__ tf0, GridArgs<float> a1,
                                                        const float* __restrict__ ti1, const float* __restrict__ tf1,
                                                        int nblk0) {
  extern __shared__ __align__(16) float lds[];
  if ((int)blockIdx.x < nblk0) {
    if constexpr (FA)
      grid_mfma_fold_body<MA>(a0, ti0, tf0, blockIdx.x, lds);
    else
      grid_mfma_body<MA>(a0, ti0, tf0, blockIdx.x, lds);
  } else {
    if constexpr (FB)
      grid_mfma_fold_body<MB>(a1, ti1, tf1, blockIdx.x - nblk0, lds);
    else
      grid_mfma_body<MB>(a1, ti1, tf1, blockIdx.x - nblk0, lds);
  }
}

// The steady form of the pair launch (both plans folded, a step without bookkeeping): the
// FP32-only bodies allocated for W waves per SIMD (two, or one with a T106 body, whose registers
// need a whole SIMD), plan 1's (the larger truncation's) workgroups first so that plan 0's
// shorter ones fill the tail.
template <int MA, int MB, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, 8)))
k_grid_mfma_pair_fast(GridArgs<float> a0, const float* __restrict__ ti0, const float* __restrict__ tf0,
                      GridArgs<float> a1, const float* __restrict__ ti1, const float* __restrict__ tf1, int nblk1) {
  extern __shared__ __align__(16) float lds[];
  if ((int)blockIdx.x < nblk1)
    grid_mfma_fold_body<MB, true>(a1, ti1, tf1, blockIdx.x, lds);
  else
    grid_mfma_fold_body<MA, true>(a0, ti0, tf0, blockIdx.x - nblk1, lds);
}

template <int MT>
static hipError_t launch_mfma_t(const GridArgs<float>& a, const float* TinvF, const float* TfwdF, hipStream_t s) {
  using S = MfmaShape<MT>;
  const size_t smem = sizeof(float) * S::LDS_FLOATS;
  dim3 grid((unsigned)((long long)a.B * a.J / S::ROWS));
  hipLaunchKernelGGL(k_grid_mfma<MT>, grid, dim3(256), smem, s, a, TinvF, TfwdF);
  return hipGetLastError();
}

// Whether a folded-grid launch can take the fast variant (no FP64 bookkeeping at this step).
static bool grid_fold_fast(const GridArgs<float>& a) {
  return !a.phys_on && !a.raw_filter && a.test != 1 && !a.si && !a.expflag;
}

template <int MT>
static hipError_t launch_mfma_fold_t(const GridArgs<float>& a, const float* TinvF, const float* TfwdF,
                                     hipStream_t s, bool shared) {
  using S = MfmaFoldShape<MT>;
  dim3 grid((unsigned)((long long)a.B * a.J / S::ROWS));
  const bool fast = grid_fold_fast(a);
  const bool diag = a.diagnostics || a.hist_fields || a.hist_phiext;
  const size_t smem = sizeof(float) * (fast ? S::LDS_FAST : S::LDS_FLOATS);
  if (diag) {
    // Steps with per-step outputs have the physical carry on (never the fast variant).
    if constexpr (MT == 42)
      hipLaunchKernelGGL((k_grid_mfma_fold_w2<MT, false, true>), grid, dim3(256), smem, s, a, TinvF, TfwdF);
    else
      hipLaunchKernelGGL((k_grid_mfma_fold<MT, false, true>), grid, dim3(256), smem, s, a, TinvF, TfwdF);
    return hipGetLastError();
  }
  if constexpr (MT == 42) {
    if (fast && shared)
      hipLaunchKernelGGL((k_grid_mfma_fold<MT, true>), grid, dim3(256), smem, s, a, TinvF, TfwdF);
    else if (fast)
      hipLaunchKernelGGL((k_grid_mfma_fold_w2<MT, true>), grid, dim3(256), smem, s, a, TinvF, TfwdF);
    else
      hipLaunchKernelGGL((k_grid_mfma_fold_w2<MT, false>), grid, dim3(256), smem, s, a, TinvF, TfwdF);
  } else {
    if (fast)
      hipLaunchKernelGGL((k_grid_mfma_fold<MT, true>), grid, dim3(256), smem, s, a, TinvF, TfwdF);
    else
      hipLaunchKernelGGL((k_grid_mfma_fold<MT, false>), grid, dim3(256), smem, s, a, TinvF, TfwdF);
  }
  return hipGetLastError();
}

size_t grid_mfma_fold_lds_bytes(int M) {
  switch (M) {
    case 42: return sizeof(float) * MfmaFoldShape<42>::LDS_FLOATS;
    case 63: return sizeof(float) * MfmaFoldShape<63>::LDS_FLOATS;
    case 106: return sizeof(float) * MfmaFoldShape<106>::LDS_FLOATS;
    default: return 0;
  }
}

template <int MT>
static void fold_dims_t(FoldDims* d) {
  using S = MfmaFoldShape<MT>;
  d->H = S::H;
  d->K1 = S::K1;
  d->KE4 = S::KE4;
  d->KQ1 = S::KQ1;
  d->TE = S::TE;
  d->NT2 = S::NT2;
}

bool grid_mfma_fold_dims(int M, FoldDims* d) {
  switch (M) {
    case 42: fold_dims_t<42>(d); return true;
    case 63: fold_dims_t<63>(d); return true;
    case 106: fold_dims_t<106>(d); return true;
    default: return false;
  }
}

hipError_t launch_grid_mfma_fold(const GridArgs<float>& a, const float* TinvF, const float* TfwdF, hipStream_t s,
                                 bool shared) {
  switch (a.M) {
    case 42: return launch_mfma_fold_t<42>(a, TinvF, TfwdF, s, shared);
    case 63: return launch_mfma_fold_t<63>(a, TinvF, TfwdF, s, shared);
    case 106: return launch_mfma_fold_t<106>(a, TinvF, TfwdF, s, shared);
    default: return hipErrorInvalidValue;
  }
}

size_t grid_mfma_lds_bytes(int M) {
  switch (M) {
    case 42: return sizeof(float) * MfmaShape<42>::LDS_FLOATS;
    case 63: return sizeof(float) * MfmaShape<63>::LDS_FLOATS;
    case 106: return sizeof(float) * MfmaShape<106>::LDS_FLOATS;
    default: return 0;
  }
}

int grid_mfma_kq1(int M) {
  switch (M) {
    case 42: return MfmaShape<42>::KQ1;
    case 63: return MfmaShape<63>::KQ1;
    case 106: return MfmaShape<106>::KQ1;
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

hipError_t launch_grid_mfma(const GridArgs<float>& a, const float* TinvF, const float* TfwdF, hipStream_t s) {
  switch (a.M) {
    case 42: return launch_mfma_t<42>(a, TinvF, TfwdF, s);
    case 63: return launch_mfma_t<63>(a, TinvF, TfwdF, s);
    case 106: return launch_mfma_t<106>(a, TinvF, TfwdF, s);
    default: return hipErrorInvalidValue;
  }
}



// ----------------------------------------------------------------------------
// MFMA Legendre kernel (FP32 transforms, FP64 spectral update).
//
// One workgroup = one zonal wavenumber m and MB members.  With K = N + 1 - m
// retained degrees:
//   phase F  forward contraction as GEMMs: C[n][r] = sum_j P^T[n][j] T[j][r]
//            (P slots) and H^T for the H slots, r = (member, slot, re/im);
//            B operands staged in LDS as [j][r]; A = basis rows read from the
//            [n][j] layout.  (modEuler_tdiff.py fwd_leg_w calls, regrouped.)
//   phase U  FP64 spectral update per (member, n) (swx_physics.h).
//   phase I  inverse contraction: D[j][r'] = sum_n P[j][n] X[n][r'] for
//            eta, delta, Phi and the two wind potentials, H[j][n] for the wind
//            potentials (spectral_transform.py:438-596).
//   phase W  U_m = -i m P d^ + H e^, V_m = -i m P e^ - H d^ and coalesced
//            writes of the m-major Fourier rows.
// Column groups are padded to 16 so that every MFMA tile is all-P or all-H.
// ----------------------------------------------------------------------------

namespace {
constexpr int ceil16(int x) { return (x + 15) / 16 * 16; }
constexpr int ceil4(int x) { return (x + 3) / 4 * 4; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
}  // namespace

template <int MT, int MB>
struct LegShape {
  static constexpr int Np = MT + 1;
  static constexpr int J = MT == 42 ? 64 : (MT == 63 ? 96 : 160);
  static constexpr int RP = ceil16(10 * MB), RH = ceil16(8 * MB), RT = RP + RH;  // forward columns
  static constexpr int XP = ceil16(10 * MB), XH = ceil16(4 * MB), XT = XP + XH;  // inverse columns
  static constexpr int KNP = ceil16(Np);  // padded degree rows
  static constexpr int JO = J + 4;         // column stride of the inverse results (bank spread)
  // Odd row strides (one unused column): consecutive rows start in different LDS banks, so
  // the per-row accesses of the staging and update phases do not conflict.
  static constexpr int RS = RT + 1, XS = XT + 1;
  static constexpr int SA = cmax(cmax(J * RS, KNP * RS), XT * JO);
  // Inverse inputs: rows n < K, zero to the k-step multiple (phase I reads rows < ceil4(K)).
  static constexpr int SX = ceil4(Np) * XS;
  static constexpr int LDS_FLOATS = SA + SX;
  // Folded kernel: the P columns unpadded (10 MB), the H columns right after them.  A P tile that
  // reaches past them reads H inputs; its outputs there are not stored (forward) or are
  // overwritten by the H tile's (inverse: the same wave stores it later).  4 KB less LDS at T42.
  static constexpr int RPF = 10 * MB, CTPF = (RPF + 15) / 16, RSF = RPF + RH + 1;
  static constexpr int XPF = 10 * MB, CIPF = (XPF + 15) / 16, XTF = XPF + XH, XSF = XTF + 1;
  static constexpr int SAF = cmax(cmax(J * RSF, 16 * leg_fold_ntf(MT) * RSF), XTF * JO);
  static constexpr int SXF = ceil4(Np) * XSF;
  static constexpr int LDS_FLOATS_F = SAF + SXF;
};

// Basis fragments of forward tile t (row tile t / CT, P or H by column tile t % CT) for this
// lane; a tile past the last (t >= ntile) loads tile 0's, which is never used.
template <int MT, int MB>
__device__ __forceinline__ void leg_fwd_frags(const LegArgs<float>& a, int m, int t, int ntile, int lane,
                                              v4f (&av)[LegShape<MT, MB>::J / 16]) {
  using S = LegShape<MT, MB>;
  constexpr int CT = S::RT / 16, J = S::J, KNP = S::KNP;
  const int tc = t < ntile ? t : 0;
  const int nt16 = tc / CT, ct = tc % CT;
  const v4f* ar = reinterpret_cast<const v4f*>((ct * 16 < S::RP ? a.fwdP : a.fwdH) +
                                               ((long long)(m * (KNP / 16) + nt16) * 64 + lane) * (J / 4));
#pragma unroll
  for (int q4 = 0; q4 < J / 16; ++q4) av[q4] = ar[q4];
}

// Folded tables: the fragments of forward tile t (row tile t / CT of the per-m list, P or H by
// column tile t % CT), J / 8 k-steps.
template <int MT, int MB>
__device__ __forceinline__ void leg_fwd_frags_fold(const LegArgs<float>& a, int m, int t, int ntile, int lane,
                                                   v4f (&av)[LegShape<MT, MB>::J / 32]) {
  using S = LegShape<MT, MB>;
  constexpr int CT = S::CTPF + S::RH / 16, J = S::J, NTF = leg_fold_ntf(MT);
  const int tc = t < ntile ? t : 0;
  const int r = tc / CT, ct = tc % CT;
  const v4f* ar = reinterpret_cast<const v4f*>((ct < S::CTPF ? a.ffwdP : a.ffwdH) +
                                               ((long long)(m * NTF + r) * 64 + lane) * (J / 8));
#pragma unroll
  for (int q4 = 0; q4 < J / 32; ++q4) av[q4] = ar[q4];
}

// Split-BF16x3 products on the matrix cores (v_mfma_f32_16x16x16_bf16): four consecutive k-steps
// of the FP32 16x16x4 form in one instruction.  A lane's 4 A values and 4 B values are those of
// the same k's (k-step 4 g + e, its lane group), so pairing element e of A with element e of B
// sums exactly the 16 products of the four k-steps.  x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi); a b is accumulated in FP32 as lo hi' + hi lo' + hi hi' (lo lo', about
// 2^-16 relative, dropped).
typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

struct SplitBF16 {
  v4s hi, lo;
};

__device__ __forceinline__ SplitBF16 split_bf16(v4f x) {
  v4bf h, l;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    h[e] = (__bf16)x[e];
    l[e] = (__bf16)(x[e] - (float)h[e]);
  }
  return SplitBF16{__builtin_bit_cast(v4s, h), __builtin_bit_cast(v4s, l)};
}

__device__ __forceinline__ v4f mfma_bf16x3(const SplitBF16& a, v4f b, v4f c) {
  const SplitBF16 bs = split_bf16(b);
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.lo, bs.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, bs.lo, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, bs.hi, c, 0, 0, 0);
}

// One workgroup's work: zonal wavenumber m, member group `grp` (members grp*MB ...).
// GEN = false: the modified-Euler scheme without the RAW filter only, whose spectral update
// the compiler then specializes (the semi-implicit, RAW and explicit paths of spectral_update
// hold many more live doubles; dropping them lowers the register count from 152 to 112 at
// T42, 4 members: 4 waves per SIMD instead of 3).
// LIN (with GEN = false): the step also advances the spectral prev lineage (StepFlags::lineage).
// FOLD: the equatorially folded contractions (swx_kernels.h, leg_fold_*): half the matrix-core
// work and half the basis fragments; the same FP64 update.
// BF3 (with FOLD): the forward and inverse contractions as split-BF16x3 products.
template <int MT, int MB, bool GEN = true, bool LIN = false, bool FOLD = true, bool BF3 = false>
__device__ __forceinline__ void leg_mfma_body(const LegArgs<float>& a, const int m, const int grp, float* lds) {
  static_assert(!BF3 || FOLD, "split-BF16x3 contractions are provided for the folded kernel");
  using S = LegShape<MT, MB>;
  constexpr int Np = S::Np, Mp = MT + 1, J = S::J, JO = S::JO;
  // Column layout (folded: the P groups unpadded, LegShape).  P column tiles 0 .. CTP-1 start at
  // 16 ct, H tiles at RP + 16 (ct - CTP).
  constexpr int RP = FOLD ? S::RPF : S::RP, RT = RP + S::RH, RS = FOLD ? S::RSF : S::RS;
  constexpr int CTP = FOLD ? S::CTPF : S::RP / 16;
  constexpr int XP = FOLD ? S::XPF : S::XP, XT = XP + S::XH, XS = XT + 1;
  constexpr int SA = FOLD ? S::SAF : S::SA;
  constexpr int KNP = S::KNP;
  static_assert(KNP == leg_knp(MT), "fragment tables use leg_knp rows");
  float* sA = lds;           // [J][RS] inputs, then [K][RS] forward results, then [XT][JO] inverse results
  float* sX = lds + SA;      // [ceil4(Np)][XS] inverse-transform inputs
  const int b0 = grp * MB;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int B = a.B, N = a.N;
  const int K = Np - m;
  const int Kt = (K + 15) / 16;
  const int K4 = ceil4(K);
  const CoefLayout cl = CoefLayout::make(N, J);
  // Folded layout of this m: even degrees np = 2 i (i < Ke), odd np = 2 i + 1 (i < Ko).
  const int Ke = (K + 1) / 2;
  const int nte = leg_fold_nte(K), nto = leg_fold_nto(K);

  // The current spectral level of this thread's phase-U pairs (p = tid + 256 u), loaded now so
  // that the loads complete under the staging and the forward GEMM (up to 2 pairs per thread;
  // larger groups load in phase U).
  constexpr int NPF = (MB * Np + 255) / 256;
  constexpr bool kPre = NPF <= 2;
  double scp[kPre ? NPF : 1][3][2];
  // With one pair per thread, its coefficients and f2 too (the update's other global reads).
  constexpr bool kPreC = NPF == 1;
  UpdCoef ucp[1];
  double f2p[1];
  if constexpr (kPre) {
    const long long fstride = (long long)Mp * Np * 2;
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
      const int p = tid + 256 * u;
      const int mb = p / K, np = p % K;
      const bool ok = p < MB * K && b0 + mb < B;
      const double* Sc = a.S + ((long long)(ok ? b0 + mb : 0) * 2 + 1) * 3 * Mp * Np * 2 +
                         ((long long)m * Np + (ok ? m + np : m)) * 2;
#pragma unroll
      for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int c = 0; c < 2; ++c) scp[u][f][c] = Sc[f * fstride + c];
      if constexpr (kPreC) {
        const int n = ok ? m + np : m;
        ucp[0] = load_upd_coef(a.coef + (long long)(ok ? b0 + mb : 0) * a.ncoef, cl, m, n);
        f2p[0] = a.f2[(long long)m * Np + n];
      }
    }
  }

  // The first forward tile's basis fragments, in flight during the staging.
  constexpr int CT = CTP + S::RH / 16;
  constexpr int NRT = FOLD ? leg_fold_ntf(MT) : ceil16(Np) / 16;  // row tiles per m, at most
  constexpr int MAXT = (NRT * CT + 3) / 4;
  constexpr int KSF = FOLD ? J / 8 : J / 4;                          // forward k-steps
  const int ntile = (FOLD ? nte + nto : Kt) * CT;
  v4f av0[KSF / 4];
  if constexpr (FOLD)
    leg_fwd_frags_fold<MT, MB>(a, m, wave, ntile, lane, av0);
  else
    leg_fwd_frags<MT, MB>(a, m, wave, ntile, lane, av0);

  // Phase 0: stage the Legendre inputs [j][member, slot, re/im]; zero sX.  Each member's
  // [slot][j][re/im] block is contiguous in T: float4 loads (two latitudes), all issued first.
  // Folded: rows j < J/2 hold s = T(j) + T(J-1-j), rows J/2 + j hold a = T(j) - T(J-1-j), from
  // the float4 of latitudes (j, j+1) and the one of (J-2-j, J-1-j).
  if constexpr (FOLD) {
    constexpr int QPM = kNT * J / 4;  // latitude-pair float4s of the southern half per member
    constexpr int NQ = (MB * QPM + 255) / 256;
    v4f tv[NQ], tw[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int q = tid + 256 * u;
      const int mb = q / QPM < MB ? q / QPM : MB - 1;
      const int b = b0 + mb < B ? b0 + mb : B - 1;  // clamped address, zeroed below
      const int r = q % QPM, k = r / (J / 4), j = 2 * (r % (J / 4));
      const float* tb = a.T + (((long long)b * Mp + m) * kNT + k) * J * 2;
      const v4f v = *reinterpret_cast<const v4f*>(tb + 2 * j);
      const v4f w = *reinterpret_cast<const v4f*>(tb + 2 * (J - 2 - j));
      const bool ok = q < MB * QPM && b0 + q / QPM < B;
      tv[u] = ok ? v : v4f{0.f, 0.f, 0.f, 0.f};
      tw[u] = ok ? w : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int q = tid + 256 * u;
      if (q < MB * QPM) {
        const int mb = q / QPM, r = q % QPM;
        const int k = r / (J / 4), j = 2 * (r % (J / 4));
        const int col = k < kNP ? mb * 10 + 2 * k : RP + mb * 8 + 2 * (k - kNP);
        float* d0 = sA + j * RS + col;
        float* d1 = sA + (J / 2 + j) * RS + col;
        // T(j) = tv.xy, T(j+1) = tv.zw, T(J-2-j) = tw.xy, T(J-1-j) = tw.zw
        d0[0] = tv[u][0] + tw[u][2];
        d0[1] = tv[u][1] + tw[u][3];
        d0[RS] = tv[u][2] + tw[u][0];
        d0[RS + 1] = tv[u][3] + tw[u][1];
        d1[0] = tv[u][0] - tw[u][2];
        d1[1] = tv[u][1] - tw[u][3];
        d1[RS] = tv[u][2] - tw[u][0];
        d1[RS + 1] = tv[u][3] - tw[u][1];
      }
    }
  } else {
    constexpr int QPM = kNT * J / 2;  // float4 per member
    constexpr int NQ = (MB * QPM + 255) / 256;
    v4f tv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int q = tid + 256 * u;
      const int mb = q / QPM < MB ? q / QPM : MB - 1;
      const int b = b0 + mb < B ? b0 + mb : B - 1;  // clamped address, zeroed below
      const v4f v = *reinterpret_cast<const v4f*>(a.T + (((long long)b * Mp + m) * kNT) * J * 2 + 4 * (q % QPM));
      tv[u] = (q < MB * QPM && b0 + q / QPM < B) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int q = tid + 256 * u;
      if (q < MB * QPM) {
        const int mb = q / QPM, r = q % QPM;
        const int k = r / (J / 2), j = 2 * (r % (J / 2));
        const int col = k < kNP ? mb * 10 + 2 * k : RP + mb * 8 + 2 * (k - kNP);
        float* d0 = sA + j * RS + col;
        d0[0] = tv[u][0];
        d0[1] = tv[u][1];
        d0[RS] = tv[u][2];
        d0[RS + 1] = tv[u][3];
      }
    }
  }
  constexpr int SX = FOLD ? S::SXF : S::SX;
  static_assert(SA % 4 == 0 && SX % 4 == 0, "sX is zeroed as float4s");
  for (int e = tid; e < SX / 4; e += 256) reinterpret_cast<v4f*>(sX)[e] = v4f{0.f, 0.f, 0.f, 0.f};
  constexpr int padH = RT - RP - MB * 8, padP = RP - MB * 10;
  if constexpr (padH > 0)
    for (int e = tid; e < J * padH; e += 256) sA[(e / padH) * RS + RP + MB * 8 + e % padH] = 0.0f;
  if constexpr (padP > 0)
    for (int e = tid; e < J * padP; e += 256) sA[(e / padP) * RS + MB * 10 + e % padP] = 0.0f;
  __syncthreads();

  // Phase F: forward contraction tiles (Kt row tiles x RT/16 column tiles).  The tiles' A
  // fragments (fragment-ordered table, contiguous per lane) are double-buffered: the next
  // tile's loads are in flight during the current tile's MFMAs.
  v4f acc[MAXT];
#pragma unroll
  for (int q = 0; q < MAXT; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
  v4f avn[KSF / 4];
#pragma unroll
  for (int q4 = 0; q4 < KSF / 4; ++q4) avn[q4] = av0[q4];
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    const int t = wave + 4 * q;
    v4f av[KSF / 4];
#pragma unroll
    for (int q4 = 0; q4 < KSF / 4; ++q4) av[q4] = avn[q4];
    if (q + 1 < MAXT) {
      if constexpr (FOLD)
        leg_fwd_frags_fold<MT, MB>(a, m, wave + 4 * (q + 1), ntile, lane, avn);
      else
        leg_fwd_frags<MT, MB>(a, m, wave + 4 * (q + 1), ntile, lane, avn);
    }
    if (t < ntile) {
      const int ct = t % CT;
      const int colb = ct < CTP ? 16 * ct : RP + 16 * (ct - CTP);
      // Folded: the B operand is s (rows 0 ..) for P tiles of even degrees and H tiles of odd
      // ones, else a (rows J/2 ..).
      int base = 0;
      if constexpr (FOLD) {
        const bool even = t / CT < nte, isP = ct < CTP;
        base = (even == isP) ? 0 : J / 2;
      }
      const float* br = sA + (base + lg) * RS + colb + ln;
      if constexpr (BF3) {
#pragma unroll
        for (int g = 0; g < KSF / 4; ++g) {
          const v4f bv = {br[16 * g * RS], br[(16 * g + 4) * RS], br[(16 * g + 8) * RS], br[(16 * g + 12) * RS]};
          acc[q] = mfma_bf16x3(split_bf16(av[g]), bv, acc[q]);
        }
      } else {
#pragma unroll
        for (int ks = 0; ks < KSF; ++ks)
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ks / 4][ks % 4], br[4 * ks * RS], acc[q], 0, 0, 0);
      }
    }
  }
  __syncthreads();
  // Forward results, row tile by row tile (folded: the even-degree tiles, then the odd ones).
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    const int t = wave + 4 * q;
    if (t < ntile) {
      const int nt16 = t / CT, ct = t % CT;
      const int col = (ct < CTP ? 16 * ct : RP + 16 * (ct - CTP)) + ln;
      if (ct >= CTP || col < RP) {  // (a folded P tile's columns past RP are the H tile's)
#pragma unroll
        for (int r = 0; r < 4; ++r) sA[(nt16 * 16 + 4 * lg + r) * RS + col] = acc[q][r];
      }
    }
  }
  __syncthreads();

  // Phase U: FP64 update per (member, n) and the inverse-transform inputs.
  StepFlags fl = a.flags();
  if constexpr (!GEN) {
    fl.si = 0;
    fl.raw = 0;
    fl.expflag = 0;
    fl.lineage = LIN ? 1 : 0;
  }
  if (GEN && (a.si || a.raw_filter) && m == 0 && tid < MB && b0 + tid < B)
    residue_step(a.coef + (long long)(b0 + tid) * a.ncoef, fl, a.resid_c + 2 * (b0 + tid));
#pragma unroll
  for (int u = 0; u < (MB * Np + 255) / 256; ++u) {
    const int p = tid + 256 * u;
    if (p >= MB * K) break;
    const int mb = p / K, np = p % K;
    const int b = b0 + mb;
    if (b >= B) continue;
    if (a.diagnostics && a.dead[b]) continue;  // frozen by the blow-up gate: carry kept as it is
    const int n = m + np;
    const double* cf = a.coef + (long long)b * a.ncoef;
    // Row of degree np in the forward results and in the inverse inputs.
    const int frow = FOLD ? ((np & 1) ? 16 * nte + (np >> 1) : (np >> 1)) : np;
    const int xrow = FOLD ? ((np & 1) ? Ke + (np >> 1) : (np >> 1)) : np;
    double fp[5][2], fh[4][2];
    for (int k = 0; k < 5; ++k)
      for (int c = 0; c < 2; ++c) fp[k][c] = sA[frow * RS + mb * 10 + 2 * k + c];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 2; ++c) fh[k][c] = sA[frow * RS + RP + mb * 8 + 2 * k + c];
    double* Sb = a.S + (long long)b * 2 * 3 * Mp * Np * 2;
    const long long fstride = (long long)Mp * Np * 2;
    const long long so = ((long long)m * Np + n) * 2;
    double eff[3][2], nw[3][2], dh[2], eh[2];
    spectral_update(cf, cl, kPreC ? f2p[0] : a.f2[(long long)m * Np + n], m, n, fl, fp, fh, Sb + 3 * fstride, Sb,
                    so, fstride, eff, nw, kPre ? scp[kPre ? u : 0] : nullptr, kPreC ? &ucp[0] : nullptr);
    wind_potentials(cf, cl, m, n, nw[0], nw[1], dh, eh, kPreC ? &ucp[0] : nullptr);
    float* xr = sX + xrow * XS;
    for (int c = 0; c < 2; ++c) {
      xr[mb * 10 + 0 + c] = (float)eff[0][c];
      xr[mb * 10 + 2 + c] = (float)eff[1][c];
      xr[mb * 10 + 4 + c] = (float)eff[2][c];
      xr[mb * 10 + 6 + c] = (float)dh[c];
      xr[mb * 10 + 8 + c] = (float)eh[c];
      xr[XP + mb * 4 + 0 + c] = (float)eh[c];
      xr[XP + mb * 4 + 2 + c] = (float)dh[c];
    }
  }
  __syncthreads();

  // Phase I: inverse contraction; wave w owns latitude tiles jt = w, w+4, ... and every
  // column tile, so its P and H fragments are loaded once and reused across them.
  constexpr int CIP = FOLD ? S::CIPF : XP / 16, CIH = (XT - XP) / 16;
  if constexpr (FOLD) {
    // Units (jt, parity) over the southern latitude tiles jt < J/32: the even-degree partial
    // sums E into rows jt*16 .. of the column-major output, the odd ones O into rows
    // J/2 + jt*16 ..; phase W combines them for both hemispheres.
    constexpr int KSE = leg_fold_kse(MT), KSO = leg_fold_kso(MT), KSL = KSE + KSO;
    const int ne = leg_fold_ne(K), kb = leg_fold_kb(K), no = leg_fold_no(K);
    for (int u = wave; u < J / 16; u += 4) {
      const int jt = u >> 1, odd = u & 1;
      const long long fo = ((long long)(m * (J / 32) + jt) * 64 + lane) * KSL + (odd ? KSE : 0);
      const v4f* pr = reinterpret_cast<const v4f*>(a.finvP + fo);
      const v4f* hr = reinterpret_cast<const v4f*>(a.finvH + fo);
      constexpr int KSX = KSE > KSO ? KSE : KSO;
      const int nks = odd ? no : ne;
      const int row0 = odd ? 4 * kb : 0;
      v4f pa[KSX / 4], ha[KSX / 4];
#pragma unroll
      for (int q4 = 0; q4 < KSX / 4; ++q4) {
        const bool need = 4 * q4 < nks;
        pa[q4] = need ? pr[q4] : v4f{0.f, 0.f, 0.f, 0.f};
        ha[q4] = need ? hr[q4] : v4f{0.f, 0.f, 0.f, 0.f};
      }
      v4f dP[CIP], dH[CIH];
#pragma unroll
      for (int c = 0; c < CIP; ++c) dP[c] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < CIH; ++c) dH[c] = v4f{0.f, 0.f, 0.f, 0.f};
      if constexpr (BF3) {
        // Groups of 4 k-steps; the steps past nks (and their staging rows) enter as zeros.
#pragma unroll
        for (int g = 0; g < KSX / 4; ++g) {
          if (4 * g < nks) {
            v4f pz, hz;
            int rw[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const bool in = 4 * g + e < nks;
              pz[e] = in ? pa[g][e] : 0.0f;
              hz[e] = in ? ha[g][e] : 0.0f;
              rw[e] = in ? row0 + 4 * (4 * g + e) + lg : 0;
            }
            const SplitBF16 ps = split_bf16(pz), hs = split_bf16(hz);
#pragma unroll
            for (int c = 0; c < CIP; ++c) {
              v4f bv;
#pragma unroll
              for (int e = 0; e < 4; ++e) bv[e] = 4 * g + e < nks ? sX[rw[e] * XS + ln + c * 16] : 0.0f;
              dP[c] = mfma_bf16x3(ps, bv, dP[c]);
            }
#pragma unroll
            for (int c = 0; c < CIH; ++c) {
              v4f bv;
#pragma unroll
              for (int e = 0; e < 4; ++e) bv[e] = 4 * g + e < nks ? sX[rw[e] * XS + ln + XP + c * 16] : 0.0f;
              dH[c] = mfma_bf16x3(hs, bv, dH[c]);
            }
          }
        }
      } else {
#pragma unroll
        for (int ks = 0; ks < KSX; ++ks) {
          if (ks < nks) {
            const float* xr = sX + (row0 + 4 * ks + lg) * XS + ln;
#pragma unroll
            for (int c = 0; c < CIP; ++c)
              dP[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[ks / 4][ks % 4], xr[c * 16], dP[c], 0, 0, 0);
#pragma unroll
            for (int c = 0; c < CIH; ++c)
              dH[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ha[ks / 4][ks % 4], xr[XP + c * 16], dH[c], 0, 0, 0);
          }
        }
      }
      float* out = sA + ln * JO + (odd ? J / 2 : 0) + jt * 16 + 4 * lg;
#pragma unroll
      for (int c = 0; c < CIP; ++c) *reinterpret_cast<v4f*>(out + c * 16 * JO) = dP[c];
#pragma unroll
      for (int c = 0; c < CIH; ++c) *reinterpret_cast<v4f*>(out + (XP + c * 16) * JO) = dH[c];
    }
  } else {
    constexpr int KSM = KNP / 4;  // k-steps at m = 0
    const int ksn = K4 / 4;
    for (int jt = wave; jt < J / 16; jt += 4) {
      const int j = jt * 16 + ln;
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
      v4f dP[CIP], dH[CIH];
#pragma unroll
      for (int c = 0; c < CIP; ++c) dP[c] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < CIH; ++c) dH[c] = v4f{0.f, 0.f, 0.f, 0.f};
      if constexpr (BF3) {
#pragma unroll
        for (int g = 0; g < KSM / 4; ++g) {
          if (4 * g < ksn) {
            v4f pz, hz;
            int rw[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const bool in = 4 * g + e < ksn;
              pz[e] = in ? pa[g][e] : 0.0f;
              hz[e] = in ? ha[g][e] : 0.0f;
              rw[e] = in ? 4 * (4 * g + e) + lg : 0;
            }
            const SplitBF16 ps = split_bf16(pz), hs = split_bf16(hz);
#pragma unroll
            for (int c = 0; c < CIP; ++c) {
              v4f bv;
#pragma unroll
              for (int e = 0; e < 4; ++e) bv[e] = 4 * g + e < ksn ? sX[rw[e] * XS + ln + c * 16] : 0.0f;
              dP[c] = mfma_bf16x3(ps, bv, dP[c]);
            }
#pragma unroll
            for (int c = 0; c < CIH; ++c) {
              v4f bv;
#pragma unroll
              for (int e = 0; e < 4; ++e) bv[e] = 4 * g + e < ksn ? sX[rw[e] * XS + ln + XP + c * 16] : 0.0f;
              dH[c] = mfma_bf16x3(hs, bv, dH[c]);
            }
          }
        }
      } else {
#pragma unroll
        for (int ks = 0; ks < KSM; ++ks) {
          if (ks < ksn) {
            const float* xr = sX + (4 * ks + lg) * XS + ln;
#pragma unroll
            for (int c = 0; c < CIP; ++c)
              dP[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[ks / 4][ks % 4], xr[c * 16], dP[c], 0, 0, 0);
#pragma unroll
            for (int c = 0; c < CIH; ++c)
              dH[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ha[ks / 4][ks % 4], xr[XP + c * 16], dH[c], 0, 0, 0);
          }
        }
      }
      // Column-major [col][JO] output: the 4 rows a lane holds are contiguous (one b128 store).
      float* out = sA + ln * JO + jt * 16 + 4 * lg;
#pragma unroll
      for (int c = 0; c < CIP; ++c) *reinterpret_cast<v4f*>(out + c * 16 * JO) = dP[c];
#pragma unroll
      for (int c = 0; c < CIH; ++c) *reinterpret_cast<v4f*>(out + (XP + c * 16) * JO) = dH[c];
    }
  }
  __syncthreads();

  // Phase W: Fourier rows of eta, delta, Phi, U, V (m-major, coalesced in j).
  const float mr = (float)m;
  for (int p = tid; p < MB * J; p += 256) {
    const int mb = p / J, j = p % J;
    const int b = b0 + mb;
    if (b >= B) continue;
    if (a.diagnostics && a.dead[b]) continue;
    float* G = a.G + fidx(b, 0, m, j, Mp, J);
    const long long fs = (long long)Mp * J * 2;
    float P[10], H[4];
    if constexpr (FOLD) {
      // X(j) = E + O on the southern half, X(J-1-j) = E - O on the northern (H: O - E).
      const int js = j < J / 2 ? j : J - 1 - j;
      const float sg = j < J / 2 ? 1.0f : -1.0f;
      const float* pe = sA + mb * 10 * JO + js;
      const float* he = sA + (XP + mb * 4) * JO + js;
#pragma unroll
      for (int c = 0; c < 10; ++c) P[c] = pe[c * JO] + sg * pe[c * JO + J / 2];
#pragma unroll
      for (int c = 0; c < 4; ++c) H[c] = sg * he[c * JO] + he[c * JO + J / 2];
    } else {
      const float* pr = sA + mb * 10 * JO + j;
      const float* hr = sA + (XP + mb * 4) * JO + j;
#pragma unroll
      for (int c = 0; c < 10; ++c) P[c] = pr[c * JO];
#pragma unroll
      for (int c = 0; c < 4; ++c) H[c] = hr[c * JO];
    }
    G[0] = P[0];
    G[1] = P[1];
    G[fs] = P[2];
    G[fs + 1] = P[3];
    G[2 * fs] = P[4];
    G[2 * fs + 1] = P[5];
    G[3 * fs] = mr * P[7] + H[0];
    G[3 * fs + 1] = -mr * P[6] + H[1];
    G[4 * fs] = mr * P[9] - H[2];
    G[4 * fs + 1] = -mr * P[8] - H[3];
  }
}

template <int MT, int MB, bool GEN, bool LIN, bool FOLD, bool BF3 = false>
__global__ void __launch_bounds__(256) k_leg_mfma(LegArgs<float> a, int heavy_first) {
  extern __shared__ __align__(16) float lds[];
  int m, grp;
  leg_item(heavy_first, m, grp);
  leg_mfma_body<MT, MB, GEN, LIN, FOLD, BF3>(a, m, grp, lds);
}

// The same kernel allocated for at least 4 waves per SIMD (MFMA accumulators in VGPRs), for the
// specialized modified-Euler update where that costs no spills: T42 (82 registers at 4
// members, 5 workgroups per CU with the LDS) and T63 with 2 or 4 members.
template <int MT, int MB, bool GEN, bool LIN, bool FOLD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_leg_mfma_w4(LegArgs<float> a, int heavy_first) {
  extern __shared__ __align__(16) float lds[];
  int m, grp;
  leg_item(heavy_first, m, grp);
  leg_mfma_body<MT, MB, GEN, LIN, FOLD>(a, m, grp, lds);
}

// The folded T42 kernel allocated for 5 waves per SIMD (at most 96 registers): with its 32 KB of
// LDS, 5 workgroups per CU, also for the steps that advance the prev lineage.
template <int MT, int MB, bool GEN, bool LIN, bool FOLD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_leg_mfma_w5(LegArgs<float> a, int heavy_first) {
  extern __shared__ __align__(16) float lds[];
  int m, grp;
  leg_item(heavy_first, m, grp);
  leg_mfma_body<MT, MB, GEN, LIN, FOLD>(a, m, grp, lds);
}

// Two plans of different truncations in one launch: block = grp * (M+1) + m, plan 0's first.
template <int MA, int MBA, int MB, int MBB, bool GEN, bool FOLD>
__global__ void __launch_bounds__(256) k_leg_mfma_pair(LegArgs<float> a0, LegArgs<float> a1, int nblk0) {
  extern __shared__ __align__(16) float lds[];
  const int blk = blockIdx.x;
  if (blk < nblk0)
    leg_mfma_body<MA, MBA, GEN, false, FOLD>(a0, blk % (MA + 1), blk / (MA + 1), lds);
  else
    leg_mfma_body<MB, MBB, GEN, false, FOLD>(a1, (blk - nblk0) % (MB + 1), (blk - nblk0) / (MB + 1), lds);
}

// The folded specialised-update pair (plain modified Euler in both plans; LIN0 / LIN1: the plan's
// step also advances its prev lineage) with member groups sized for one LDS budget (the larger
// of the two bodies' folded LDS) and allocated for W waves per SIMD; heavy-first order: plan 1's
// workgroups (m = L / g1, group L % g1) first, then plan 0's, so the light small-K items of the
// smaller truncation fill the tail.
template <int MA, int MBA, int MB, int MBB, int W, bool LIN0, bool LIN1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, 8)))
k_leg_mfma_pair_hf(LegArgs<float> a0, LegArgs<float> a1, int g0, int g1) {
  extern __shared__ __align__(16) float lds[];
  const int n1 = (MB + 1) * g1;
  const int L = blockIdx.x;
  if (L < n1) {
    leg_mfma_body<MB, MBB, false, LIN1, true>(a1, L / g1, L % g1, lds);
  } else {
    const int L0 = L - n1;
    leg_mfma_body<MA, MBA, false, LIN0, true>(a0, L0 / g0, L0 % g0, lds);
  }
}

// Whether a launch needs the general spectral update (any scheme other than plain modified Euler).
static bool leg_generic(const LegArgs<float>& a) { return a.si || a.raw_filter || a.expflag; }

template <int MT, int MB>
static hipError_t launch_leg_mfma_t(const LegArgs<float>& a, hipStream_t s, bool shared) {
  using S = LegShape<MT, MB>;
  const size_t smem = sizeof(float) * (a.leg_fold ? S::LDS_FLOATS_F : S::LDS_FLOATS);
  dim3 grid(MT + 1, (a.B + MB - 1) / MB);
  constexpr bool kW4 = MT == 42 || (MT == 63 && MB <= 4);
  constexpr bool kW5 = MT == 42 && MB == 4;
  const int heavy_first = shared ? 0 : 1;
  if (a.leg_fold && a.bf16x3) {
    if (leg_generic(a))
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, true, false, true, true>), grid, dim3(256), smem, s, a, heavy_first);
    else if (a.lineage)
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, false, true, true, true>), grid, dim3(256), smem, s, a, heavy_first);
    else
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, false, false, true, true>), grid, dim3(256), smem, s, a, heavy_first);
  } else if (a.leg_fold) {
    if (leg_generic(a))
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, true, false, true>), grid, dim3(256), smem, s, a, heavy_first);
    else if constexpr (kW5) {
      if (a.lineage)
        hipLaunchKernelGGL((k_leg_mfma_w5<MT, MB, false, true, true>), grid, dim3(256), smem, s, a, heavy_first);
      else
        hipLaunchKernelGGL((k_leg_mfma_w5<MT, MB, false, false, true>), grid, dim3(256), smem, s, a, heavy_first);
    } else if constexpr (kW4) {
      if (a.lineage)
        hipLaunchKernelGGL((k_leg_mfma_w4<MT, MB, false, true, true>), grid, dim3(256), smem, s, a, heavy_first);
      else
        hipLaunchKernelGGL((k_leg_mfma_w4<MT, MB, false, false, true>), grid, dim3(256), smem, s, a, heavy_first);
    } else if (a.lineage) {
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, false, true, true>), grid, dim3(256), smem, s, a, heavy_first);
    } else {
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, false, false, true>), grid, dim3(256), smem, s, a, heavy_first);
    }
  } else {
    if (leg_generic(a))
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, true, false, false>), grid, dim3(256), smem, s, a, heavy_first);
    else if constexpr (kW4) {
      if (a.lineage)
        hipLaunchKernelGGL((k_leg_mfma_w4<MT, MB, false, true, false>), grid, dim3(256), smem, s, a, heavy_first);
      else
        hipLaunchKernelGGL((k_leg_mfma_w4<MT, MB, false, false, false>), grid, dim3(256), smem, s, a,
                           heavy_first);
    } else if (a.lineage) {
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, false, true, false>), grid, dim3(256), smem, s, a, heavy_first);
    } else {
      hipLaunchKernelGGL((k_leg_mfma<MT, MB, false, false, false>), grid, dim3(256), smem, s, a, heavy_first);
    }
  }
  return hipGetLastError();
}

// Member-group sizes instantiated per truncation (the plan times them at creation).
int leg_mfma_candidates(int M, int* mbs) {
  switch (M) {
    case 42: mbs[0] = 4; mbs[1] = 8; return 2;
    case 63:
    case 106: mbs[0] = 2; mbs[1] = 4; mbs[2] = 8; return 3;
    default: return 0;
  }
}

int leg_mfma_default_mb(int M) { return M == 42 ? 8 : 4; }

size_t leg_mfma_lds_bytes(int M, int mb) {
  const int key = M * 100 + mb;
  switch (key) {
    case 4204: return sizeof(float) * LegShape<42, 4>::LDS_FLOATS;
    case 4208: return sizeof(float) * LegShape<42, 8>::LDS_FLOATS;
    case 6302: return sizeof(float) * LegShape<63, 2>::LDS_FLOATS;
    case 6304: return sizeof(float) * LegShape<63, 4>::LDS_FLOATS;
    case 6308: return sizeof(float) * LegShape<63, 8>::LDS_FLOATS;
    case 10602: return sizeof(float) * LegShape<106, 2>::LDS_FLOATS;
    case 10604: return sizeof(float) * LegShape<106, 4>::LDS_FLOATS;
    case 10608: return sizeof(float) * LegShape<106, 8>::LDS_FLOATS;
    default: return 0;
  }
}

hipError_t launch_leg_mfma(const LegArgs<float>& a, int mb, hipStream_t s, bool shared) {
  switch (a.M * 100 + mb) {
    case 4204: return launch_leg_mfma_t<42, 4>(a, s, shared);
    case 4208: return launch_leg_mfma_t<42, 8>(a, s, shared);
    case 6302: return launch_leg_mfma_t<63, 2>(a, s, shared);
    case 6304: return launch_leg_mfma_t<63, 4>(a, s, shared);
    case 6308: return launch_leg_mfma_t<63, 8>(a, s, shared);
    case 10602: return launch_leg_mfma_t<106, 2>(a, s, shared);
    case 10604: return launch_leg_mfma_t<106, 4>(a, s, shared);
    case 10608: return launch_leg_mfma_t<106, 8>(a, s, shared);
    default: return hipErrorInvalidValue;
  }
}

template <int MA, int MB, bool FA, bool FB>
static hipError_t launch_grid_pair_ft(const GridArgs<float>& a0, const float* ti0, const float* tf0,
                                      const GridArgs<float>& a1, const float* ti1, const float* tf1, hipStream_t s) {
  constexpr int LA = FA ? MfmaFoldShape<MA>::LDS_FLOATS : MfmaShape<MA>::LDS_FLOATS;
  constexpr int LB = FB ? MfmaFoldShape<MB>::LDS_FLOATS : MfmaShape<MB>::LDS_FLOATS;
  const size_t smem = sizeof(float) * (size_t)cmax(LA, LB);
  const int n0 = (int)((long long)a0.B * a0.J / MfmaShape<MA>::ROWS);
  const int n1 = (int)((long long)a1.B * a1.J / MfmaShape<MB>::ROWS);
  hipLaunchKernelGGL((k_grid_mfma_pair<MA, MB, FA, FB>), dim3(n0 + n1), dim3(256), smem, s, a0, ti0, tf0, a1, ti1,
                     tf1, n0);
  return hipGetLastError();
}

template <int MA, int MB>
static hipError_t launch_grid_pair_t(const GridArgs<float>& a0, const float* ti0, const float* tf0, bool f0,
                                     const GridArgs<float>& a1, const float* ti1, const float* tf1, bool f1,
                                     hipStream_t s) {
  if (f0 && f1 && grid_fold_fast(a0) && grid_fold_fast(a1)) {
    const size_t smem = sizeof(float) * (size_t)cmax(MfmaFoldShape<MA>::LDS_FAST, MfmaFoldShape<MB>::LDS_FAST);
    const int n0 = (int)((long long)a0.B * a0.J / MfmaFoldShape<MA>::ROWS);
    const int n1 = (int)((long long)a1.B * a1.J / MfmaFoldShape<MB>::ROWS);
    constexpr int W = MB == 106 ? 1 : 2;
    hipLaunchKernelGGL((k_grid_mfma_pair_fast<MA, MB, W>), dim3(n0 + n1), dim3(256), smem, s, a0, ti0, tf0, a1, ti1,
                       tf1, n1);
    return hipGetLastError();
  }
  if (f0 && f1) return launch_grid_pair_ft<MA, MB, true, true>(a0, ti0, tf0, a1, ti1, tf1, s);
  if (f0) return launch_grid_pair_ft<MA, MB, true, false>(a0, ti0, tf0, a1, ti1, tf1, s);
  if (f1) return launch_grid_pair_ft<MA, MB, false, true>(a0, ti0, tf0, a1, ti1, tf1, s);
  return launch_grid_pair_ft<MA, MB, false, false>(a0, ti0, tf0, a1, ti1, tf1, s);
}

template <int MA, int MBA, int MB, int MBB>
static hipError_t launch_leg_pair_t(const LegArgs<float>& a0, const LegArgs<float>& a1, hipStream_t s) {
  const bool fold = a0.leg_fold && a1.leg_fold;
  const size_t smem = sizeof(float) * (size_t)(fold ? cmax(LegShape<MA, MBA>::LDS_FLOATS_F, LegShape<MB, MBB>::LDS_FLOATS_F)
                                                    : cmax(LegShape<MA, MBA>::LDS_FLOATS, LegShape<MB, MBB>::LDS_FLOATS));
  const int n0 = (MA + 1) * ((a0.B + MBA - 1) / MBA);
  const int n1 = (MB + 1) * ((a1.B + MBB - 1) / MBB);
  // The general update also covers the prev lineage of a pair's final window steps.  Both plans
  // take the folded contractions when both have them on.
  const bool gen = leg_generic(a0) || leg_generic(a1) || a0.lineage || a1.lineage;
  if (fold) {
    if (gen)
      hipLaunchKernelGGL((k_leg_mfma_pair<MA, MBA, MB, MBB, true, true>), dim3(n0 + n1), dim3(256), smem, s, a0, a1, n0);
    else
      hipLaunchKernelGGL((k_leg_mfma_pair<MA, MBA, MB, MBB, false, true>), dim3(n0 + n1), dim3(256), smem, s, a0, a1,
                         n0);
  } else {
    if (gen)
      hipLaunchKernelGGL((k_leg_mfma_pair<MA, MBA, MB, MBB, true, false>), dim3(n0 + n1), dim3(256), smem, s, a0, a1,
                         n0);
    else
      hipLaunchKernelGGL((k_leg_mfma_pair<MA, MBA, MB, MBB, false, false>), dim3(n0 + n1), dim3(256), smem, s, a0, a1,
                         n0);
  }
  return hipGetLastError();
}

bool mfma_pair_supported(int M0, int M1) {
  return (M0 == 42 && M1 == 63) || (M0 == 42 && M1 == 106) || (M0 == 63 && M1 == 106);
}

hipError_t launch_grid_mfma_pair(const GridArgs<float>& a0, const float* ti0, const float* tf0, bool fold0,
                                 const GridArgs<float>& a1, const float* ti1, const float* tf1, bool fold1,
                                 hipStream_t s) {
  if (a0.M == 42 && a1.M == 63) return launch_grid_pair_t<42, 63>(a0, ti0, tf0, fold0, a1, ti1, tf1, fold1, s);
  if (a0.M == 42 && a1.M == 106) return launch_grid_pair_t<42, 106>(a0, ti0, tf0, fold0, a1, ti1, tf1, fold1, s);
  if (a0.M == 63 && a1.M == 106) return launch_grid_pair_t<63, 106>(a0, ti0, tf0, fold0, a1, ti1, tf1, fold1, s);
  return hipErrorInvalidValue;
}

hipError_t launch_leg_mfma_pair(const LegArgs<float>& a0, const LegArgs<float>& a1, hipStream_t s) {
  if (a0.M == 42 && a1.M == 63 && a0.leg_fold && a1.leg_fold && !leg_generic(a0) && !leg_generic(a1)) {
    // T42 in groups of 4 members, T63 of 2: both fit the T42 body's 29 KB of LDS (5 workgroups
    // per CU) and its 96-register allocation.
    constexpr int MBA = 4, MBB = 2;
    const size_t smem = sizeof(float) * (size_t)cmax(LegShape<42, MBA>::LDS_FLOATS_F, LegShape<63, MBB>::LDS_FLOATS_F);
    const int g0 = (a0.B + MBA - 1) / MBA, g1 = (a1.B + MBB - 1) / MBB;
    const dim3 grid(43 * g0 + 64 * g1);
    if (a0.lineage && a1.lineage)
      hipLaunchKernelGGL((k_leg_mfma_pair_hf<42, MBA, 63, MBB, 5, true, true>), grid, dim3(256), smem, s, a0, a1, g0, g1);
    else if (a0.lineage)
      hipLaunchKernelGGL((k_leg_mfma_pair_hf<42, MBA, 63, MBB, 5, true, false>), grid, dim3(256), smem, s, a0, a1, g0,
                         g1);
    else if (a1.lineage)
      hipLaunchKernelGGL((k_leg_mfma_pair_hf<42, MBA, 63, MBB, 5, false, true>), grid, dim3(256), smem, s, a0, a1, g0,
                         g1);
    else
      hipLaunchKernelGGL((k_leg_mfma_pair_hf<42, MBA, 63, MBB, 5, false, false>), grid, dim3(256), smem, s, a0, a1, g0,
                         g1);
    return hipGetLastError();
  }
  if (a0.M == 42 && a1.M == 63) return launch_leg_pair_t<42, 8, 63, 4>(a0, a1, s);
  if (a0.M == 42 && a1.M == 106) return launch_leg_pair_t<42, 8, 106, 4>(a0, a1, s);
  if (a0.M == 63 && a1.M == 106) return launch_leg_pair_t<63, 4, 106, 4>(a0, a1, s);
  return hipErrorInvalidValue;
}

}  // namespace swx
