// Split-BF16x3 grid kernel: the MFMA grid kernel's two DFT GEMMs on the bf16
// matrix cores (v_mfma_f32_16x16x32_bf16), each operand x split as
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and a x b accumulated in
// FP32 as hi*hi' + hi*lo' + lo*hi' (the lo*lo' term, ~2^-16 relative, is
// dropped).  Same work decomposition and grid physics as k_grid_mfma
// (swx_mfma.hip): 16 rows (member, latitude) per workgroup, longitudes in
// chunks of 64, wave w owning one 16-wide longitude tile in GEMM 1 and fields
// 2w, 2w+1 in GEMM 2.  Reference functions: spectral_transform.py:290-348
// (DFTs), forcing.py, initial_conditions.py:283-292 (products).
//
// Operand layouts: a lane l supplies 8 consecutive k of row (A) or column (B)
// l % 16, k = 32 ks + 8 (l / 16) + e; the result is D[4 (l / 16) + q][l % 16].
#include "swx_grid_common.h"

namespace swx {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

constexpr int c32(int x) { return (x + 31) / 32 * 32; }
constexpr int c16(int x) { return (x + 15) / 16 * 16; }

template <int MT>
struct BfShape {
  static constexpr int Mp = MT + 1;
  static constexpr int I = MT == 42 ? 128 : (MT == 63 ? 192 : 320);
  static constexpr int K1 = 2 * Mp;       // inverse-DFT inner dimension (re, im of m = 0..M)
  static constexpr int K1b = c32(K1);     // padded to the bf16 k-step
  static constexpr int KS1 = K1b / 32;
  static constexpr int K2p = c16(2 * Mp); // forward-DFT output rows, 16-row tiles
  static constexpr int NT2 = K2p / 16;
  static constexpr int ICH = 64;          // longitudes per chunk
  static constexpr int KS2 = ICH / 32;
  static constexpr int XS = K1b + 8;      // LDS row stride (bf16) of X: 16 * odd bytes, b128 reads spread
  static constexpr int YS = ICH + 8;      // LDS row stride (bf16) of y
  // LDS (bytes): X hi, lo [5][16][XS]; y hi, lo [8][16][YS] (its region later holds the
  // FP32 accumulator exchange [2][8][16][16]).
  static constexpr int X_BYTES = 2 * 5 * 16 * XS * 2;
  static constexpr int Y_BYTES = 2 * 8 * 16 * YS * 2;
  static constexpr int LDS_BYTES = X_BYTES + Y_BYTES;
  static_assert(Y_BYTES >= 2 * 8 * 16 * 16 * 4, "exchange buffer fits in the y region");
  static_assert(I % ICH == 0, "chunking");
};

__device__ __forceinline__ void split2(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

__device__ __forceinline__ v4f mfma3(v8bf ah, v8bf al, v8bf bh, v8bf bl, v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
}

}  // namespace

template <int MT>
__global__ void __launch_bounds__(256) k_grid_bf16x3(GridArgs<float> a, const v8bf* __restrict__ TinvB,
                                                     const v8bf* __restrict__ TfwdB) {
  using S = BfShape<MT>;
  constexpr int Mp = S::Mp, I = S::I, K1b = S::K1b, KS1 = S::KS1, NT2 = S::NT2, ICH = S::ICH, KS2 = S::KS2;
  constexpr int XS = S::XS, YS = S::YS;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __bf16* sXh = reinterpret_cast<__bf16*>(lds_raw);            // [5][16][XS]
  __bf16* sXl = sXh + 5 * 16 * XS;
  __bf16* sYh = reinterpret_cast<__bf16*>(lds_raw + S::X_BYTES);  // [8][16][YS]
  __bf16* sYl = sYh + 8 * 16 * YS;
  float* sC = reinterpret_cast<float*>(lds_raw + S::X_BYTES);    // assembly tiles after the last chunk
  static_assert(S::Y_BYTES >= (int)sizeof(float) * kAsmFloats, "assembly tiles fit the product region");

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ln = lane & 15, lg = lane >> 4;
  const int J = a.J, N = a.N;
  const int row0 = blockIdx.x * 16;
  const int b = row0 / J, j0 = row0 % J;
  const double* cf = a.coef + (long long)b * a.ncoef;
  const CoefLayout cl = CoefLayout::make(N, J);
  const GeoLayout gl = GeoLayout::make(J, I);
  const double* geo = a.geo;
  const long long plane = (long long)J * I;
  double* phys = a.phys + (long long)b * 8 * plane;
  const double alpha = cf[C_ALPHA];

  // Phase 0: Fourier rows -> X split into hi / lo, [f][row][k] (k = 2m + re/im, zero padded).
  {
    constexpr int NE = 5 * (K1b / 2) * 8;
    constexpr int NQ = (NE + 255) / 256;
    v4f gv[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      const int jq = e & 7, fm = e >> 3;
      const int f = fm / (K1b / 2) < 5 ? fm / (K1b / 2) : 4, m = fm % (K1b / 2);
      const v4f v = *reinterpret_cast<const v4f*>(a.G + fidx(b, f, m < Mp ? m : Mp - 1, j0 + 2 * jq, Mp, J));
      gv[u] = (e < NE && m < Mp) ? v : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const int e = tid + 256 * u;
      if (e < NE) {
        const int jq = e & 7, fm = e >> 3;
        const int f = fm / (K1b / 2), m = fm % (K1b / 2);
#pragma unroll
        for (int r = 0; r < 2; ++r) {  // latitudes 2 jq, 2 jq + 1: (re, im) = gv[2r], gv[2r+1]
          const int off = (f * 16 + 2 * jq + r) * XS + 2 * m;
          __bf16 h0, l0, h1, l1;
          split2(gv[u][2 * r], h0, l0);
          split2(gv[u][2 * r + 1], h1, l1);
          sXh[off] = h0;
          sXh[off + 1] = h1;
          sXl[off] = l0;
          sXl[off + 1] = l1;
        }
      }
    }
  }
  __syncthreads();

  v4f acc[2 * NT2];
  float2 uv[(NT2 + 1) / 2][2];  // not prefetched here: the registers keep two waves per SIMD
#pragma unroll
  for (int p = 0; p < 2 * NT2; ++p) acc[p] = v4f{0.f, 0.f, 0.f, 0.f};

  const int n = ln;
  const int j = j0 + n;
  const PointConsts pc = point_consts(cf, geo, gl, j);
  for (int ic = 0; ic < I; ic += ICH) {
    // ---- Phase 1: inverse DFT of this wave's 16-wide longitude tile, all five fields.
    const int it = ic / 16 + wave;
    v8bf ah[KS1], al[KS1];
#pragma unroll
    for (int ks = 0; ks < KS1; ++ks) {
      const v8bf* src = TinvB + (((long long)it * KS1 + ks) * 64 + lane) * 2;
      ah[ks] = src[0];
      al[ks] = src[1];
    }
    // First k-step of this chunk's forward-DFT fragments, in flight during GEMM 1 and the physics.
    v8bf fh[NT2], fl[NT2];
#pragma unroll
    for (int t = 0; t < NT2; ++t) {
      const v8bf* src = TfwdB + (((long long)t * (I / 32) + ic / 32) * 64 + lane) * 2;
      fh[t] = src[0];
      fl[t] = src[1];
    }
    v4f c1[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      c1[f] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KS1; ++ks) {
        const int o = (f * 16 + ln) * XS + 32 * ks + 8 * lg;
        const v8bf bh = *reinterpret_cast<const v8bf*>(sXh + o);
        const v8bf bl = *reinterpret_cast<const v8bf*>(sXl + o);
        c1[f] = mfma3(ah[ks], al[ks], bh, bl, c1[f]);
      }
    }
    // ---- Grid physics on this lane's 4 points (row n, i = ic + 16 wave + 4 lg + q); the
    // products go to LDS split into hi / lo, [f][row][i].
    {
      float pr[8][4];
      float fe[4], fd[4], fp[4], fu[4], fv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        fe[q] = c1[0][q];
        fd[q] = c1[1][q];
        fp[q] = c1[2][q];
        fu[q] = c1[3][q];
        fv[q] = c1[4][q];
      }
      grid_bookkeeping4(a, cf, phys, plane, alpha, b, j, ic + 16 * wave + 4 * lg, fe, fd, fp, fu, fv);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float prod[8];
        grid_point_step(a, geo, gl, pc, ic + 16 * wave + 4 * lg + q, fe[q], fp[q], fu[q], fv[q], prod);
#pragma unroll
        for (int f = 0; f < 8; ++f) pr[f][q] = prod[f];
      }
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        const int o = (f * 16 + n) * YS + 16 * wave + 4 * lg;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          __bf16 h, l;
          split2(pr[f][q], h, l);
          sYh[o + q] = h;
          sYl[o + q] = l;
        }
      }
    }
    __syncthreads();
    // ---- Phase 2: forward DFT partial sums over this chunk, one 32-longitude k-step at a
    // time; the next k-step's fragments are fetched while the current one's MFMAs run.
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
      v8bf ch[NT2], cl2[NT2];
#pragma unroll
      for (int t = 0; t < NT2; ++t) {
        ch[t] = fh[t];
        cl2[t] = fl[t];
      }
      if (ks + 1 < KS2) {
#pragma unroll
        for (int t = 0; t < NT2; ++t) {
          const v8bf* src = TfwdB + (((long long)t * (I / 32) + ic / 32 + ks + 1) * 64 + lane) * 2;
          fh[t] = src[0];
          fl[t] = src[1];
        }
      }
      const int o0 = ((2 * wave) * 16 + ln) * YS + 32 * ks + 8 * lg;
      const int o1 = ((2 * wave + 1) * 16 + ln) * YS + 32 * ks + 8 * lg;
      const v8bf b0h = *reinterpret_cast<const v8bf*>(sYh + o0), b0l = *reinterpret_cast<const v8bf*>(sYl + o0);
      const v8bf b1h = *reinterpret_cast<const v8bf*>(sYh + o1), b1l = *reinterpret_cast<const v8bf*>(sYl + o1);
#pragma unroll
      for (int t = 0; t < NT2; ++t) {
        acc[2 * t] = mfma3(ch[t], cl2[t], b0h, b0l, acc[2 * t]);
        acc[2 * t + 1] = mfma3(ch[t], cl2[t], b1h, b1l, acc[2 * t + 1]);
      }
    }
    __syncthreads();
  }

  // ---- Phase 3: assemble the Legendre inputs (swx_grid_common.h).
  grid_assemble<Mp, NT2, 0, false>(a, acc, sC, b, j0, cf, cl, geo, gl, uv);
}

template <int MT>
static hipError_t launch_bf16x3_t(const GridArgs<float>& a, const void* TinvB, const void* TfwdB, hipStream_t s) {
  using S = BfShape<MT>;
  dim3 grid((unsigned)((long long)a.B * a.J / 16));
  hipLaunchKernelGGL(k_grid_bf16x3<MT>, grid, dim3(256), (size_t)S::LDS_BYTES, s, a,
                     static_cast<const v8bf*>(TinvB), static_cast<const v8bf*>(TfwdB));
  return hipGetLastError();
}

size_t grid_bf16x3_lds_bytes(int M) {
  switch (M) {
    case 42: return BfShape<42>::LDS_BYTES;
    case 63: return BfShape<63>::LDS_BYTES;
    case 106: return BfShape<106>::LDS_BYTES;
    default: return 0;
  }
}

int grid_bf16x3_ks1(int M) {
  switch (M) {
    case 42: return BfShape<42>::KS1;
    case 63: return BfShape<63>::KS1;
    case 106: return BfShape<106>::KS1;
    default: return 0;
  }
}

hipError_t launch_grid_bf16x3(const GridArgs<float>& a, const void* TinvB, const void* TfwdB, hipStream_t s) {
  switch (a.M) {
    case 42: return launch_bf16x3_t<42>(a, TinvB, TfwdB, s);
    case 63: return launch_bf16x3_t<63>(a, TinvB, TfwdB, s);
    case 106: return launch_bf16x3_t<106>(a, TinvB, TfwdB, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace swx
