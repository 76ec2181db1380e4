// Pieces shared by the matrix-core grid kernels (FP32 MFMA and split-BF16x3):
// the grid-space work of one point and the assembly of the Legendre inputs from
// the forward-DFT accumulators.
#pragma once

#include <cfloat>

#include "swx_kernels.h"
#include "swx_physics.h"

namespace swx {

typedef float v4f __attribute__((ext_vector_type(4)));

// Workgroup barrier for an LDS hand-off: waits for the wave's own LDS operations only, so global
// loads issued before it stay in flight across it (__syncthreads() also drains those).  For
// kernels whose threads exchange data through LDS alone.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Per-member and per-row constants of the FP32 grid physics.
struct PointConsts {
  float dph, itr, itd, Phibar, s1m, inv_denE;
};

__device__ __forceinline__ PointConsts point_consts(const double* cf, const double* geo, const GeoLayout& gl, int j) {
  PointConsts c;
  c.dph = (float)cf[C_DPHIEQ];
  c.itr = (float)cf[C_INVTAURAD];
  c.itd = (float)cf[C_INVTD];  // 0 for the taudrag == -1 sentinel
  c.Phibar = (float)cf[C_PHIBAR];
  c.s1m = (float)geo[gl.s1m + j];
  c.inv_denE = (float)(1.0 / geo[gl.denomE + j]);
  return c;
}

// The 8 products of the forward transform at one point (forcing.py:14-155,
// initial_conditions.py:283-292) in FP32, the precision of the transforms around them:
//   Q = (Phieq - (Phi + Phibar)) / taurad, formed as (DPhieq cos(lambda) sqrt(1 - mu^2) - Phi) / taurad
//       (Phieq - Phibar is that term on the strict dayside and 0 elsewhere, dcos = day * cos(lambda)),
//   R = -max(Q, 0) / (Phi + Phibar) (a zero denominator replaced by the smallest normal),
//   F = U R - U / taudrag, G = V R - V / taudrag; A..E = U eta, V eta, U Phi, V Phi,
//   (U^2 + V^2) / (2 (1 - mu^2)); PhiF = Q.  Unforced tests: PhiF = F = G = 0.
__device__ __forceinline__ void grid_products_f32(const PointConsts& pc, int test, float dcos, float eta, float Phi,
                                                  float U, float V, float out[8]) {
  // Unfused arithmetic: every kernel variant that inlines this rounds identically.
#pragma clang fp contract(off)
  float PhiF = 0.f, F = 0.f, G = 0.f;
  if (test < 0) {
    const float Q = pc.itr * ((pc.dph * dcos) * pc.s1m - Phi);
    const float Qc = Q < 0.f ? 0.f : Q;
    const float tot = Phi + pc.Phibar;
    const float r = -Qc * __builtin_amdgcn_rcpf(tot != 0.f ? tot : FLT_MIN);
    PhiF = Q;
    F = U * r - U * pc.itd;
    G = V * r - V * pc.itd;
  }
  out[0] = U * eta;
  out[1] = V * eta;
  out[2] = U * Phi;
  out[3] = V * Phi;
  out[4] = (U * U + V * V) * pc.inv_denE;
  out[5] = PhiF;
  out[6] = F;
  out[7] = G;
}

// The 8 FP32 products of one grid point (row j, longitude i); the FP64 bookkeeping of its
// fields (residue, pinned winds, physical carry) is grid_bookkeeping4's, done before.
__device__ __forceinline__ void grid_point_step(const GridArgs<float>& a, const double* geo, const GeoLayout& gl,
                                                const PointConsts& pc, int i, float eta, float Phi, float U, float V,
                                                float prod[8]) {
  grid_products_f32(pc, a.test, (float)geo[gl.dcos + i], eta, Phi, U, V, prod);
}

typedef double v4d __attribute__((ext_vector_type(4)));

// A frozen member's rows j0 .. j0 + nrows - 1 (diagnostics on, member dead; nrows <= 16, 256
// threads): the history receives its current level (eta, delta, Phi, U, V of the carry) and the
// Phi extrema of its current Phi, the plain grid kernel's frozen path (model.py:847-864).
__device__ inline void grid_frozen_rows(const GridArgs<float>& a, const double* phys, long long plane, int b, int j0,
                                        int nrows) {
  const int tid = threadIdx.x, I = a.I;
  const int r = tid >> 4, part = tid & 15;
  double mn = DBL_MAX, mx = -DBL_MAX;
  if (r < nrows) {
    const long long row = (long long)(j0 + r) * I;
    const int src[5] = {1, 3, 5, 6, 7};
    for (int i = part; i < I; i += 16) {
      const double phi = phys[5 * plane + row + i];
      mn = mn < phi ? mn : phi;
      mx = mx > phi ? mx : phi;
      if (a.hist_fields)
        for (int f = 0; f < 5; ++f)
          a.hist_fields[(((long long)b * 5 + f) * a.hist_T + a.hist_k) * plane + row + i] = phys[src[f] * plane + row + i];
    }
  }
  for (int o = 1; o < 16; o <<= 1) {  // the 16 threads of a row are 16 consecutive lanes
    const double omn = __shfl_xor(mn, o), omx = __shfl_xor(mx, o);
    mn = mn < omn ? mn : omn;
    mx = mx > omx ? mx : omx;
  }
  if (part == 0 && r < nrows && a.hist_phiext) {
    double* h = a.hist_phiext + (((long long)b * a.hist_T + a.hist_k) * a.J + j0 + r) * 2;
    h[0] = mn;
    h[1] = mx;
  }
}

// Row totals of the diagnostics partials of grid_bookkeeping4 (lane l holds row l % 16 of rows
// j0 ..; 4 waves, 256 threads; `red`: 256 doubles of LDS, free): the RMS integrand and min |wind|
// to part_uv (read by the next step's blow-up gate), the Phi extrema to the history.
__device__ inline void grid_diag_rows(const GridArgs<float>& a, double (&diag)[4], int b, int j0, double* red) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15;
  for (int o = 16; o < 64; o <<= 1) {
    const double s0 = __shfl_xor(diag[0], o), s1 = __shfl_xor(diag[1], o);
    const double s2 = __shfl_xor(diag[2], o), s3 = __shfl_xor(diag[3], o);
    diag[0] += s0;
    diag[1] = diag[1] < s1 ? diag[1] : s1;
    diag[2] = diag[2] < s2 ? diag[2] : s2;
    diag[3] = diag[3] > s3 ? diag[3] : s3;
  }
  if (lane < 16)
    for (int k = 0; k < 4; ++k) red[(wave * 16 + ln) * 4 + k] = diag[k];
  __syncthreads();
  if (tid < 16) {
    double v[4];
    for (int k = 0; k < 4; ++k) v[k] = red[tid * 4 + k];
    for (int w = 1; w < 4; ++w) {
      const double* q = red + (w * 16 + tid) * 4;
      v[0] += q[0];
      v[1] = v[1] < q[1] ? v[1] : q[1];
      v[2] = v[2] < q[2] ? v[2] : q[2];
      v[3] = v[3] > q[3] ? v[3] : q[3];
    }
    const int j = j0 + tid;
    if (a.diagnostics) {
      a.part_uv[((long long)b * a.J + j) * 2 + 0] = v[0];
      a.part_uv[((long long)b * a.J + j) * 2 + 1] = v[1];
    }
    if (a.hist_phiext) {
      double* h = a.hist_phiext + (((long long)b * a.hist_T + a.hist_k) * a.J + j) * 2;
      h[0] = v[2];
      h[1] = v[3];
    }
  }
}

// The FP64 bookkeeping of the 4 consecutive longitudes i0 .. i0 + 3 (i0 % 4 == 0) of row j of
// member b, before their products: the initial fields' residue (RAW), pinned winds (test 1),
// the time filter of the physical prev level and the physical carry (when a.phys_on;
// swx_physics.h:physical_carry, the same arithmetic on 4 points).  The FP64 planes move as
// 32-byte vectors, so the 4 lanes of a row cover whole 128-byte lines (a per-point access
// pattern writes each line in 8-byte pieces).  The fields come in and go out as FP32
// transform values.  Does nothing when none of the three applies.
// DIAG: steps with history (a.hist_fields) also write the 5 fields of the step to it, and `diag`
// (non-null with diagnostics or per-step Phi extrema) accumulates the row's partials as the plain grid kernel
// forms them (time_stepping.py:537-576; model.py:838-845, :1112-1117): diag[0] += wrms ((U/cos)^2 +
// (V/cos)^2) / aplanet, diag[1] = min |wind|, diag[2] / diag[3] = min / max of the filtered prev
// Phi (the plain kernel's expressions and comparisons).
template <bool DIAG = false>
__device__ __forceinline__ void grid_bookkeeping4(const GridArgs<float>& a, const double* cf, double* phys,
                                                  long long plane, double alpha, int b, int j, int i0,
                                                  float (&eta)[4], float (&delta)[4], float (&Phi)[4],
                                                  float (&U)[4], float (&V)[4], double* diag = nullptr,
                                                  double wrms = 0.0, double aplanet = 1.0, double cosphi = 1.0) {
  if (!(a.raw_filter || a.test == 1 || a.phys_on || (DIAG && (a.hist_fields || diag)))) return;
  const long long pix = (long long)j * a.I + i0;
  v4d e, d, p, u, v;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    e[q] = eta[q];
    d[q] = delta[q];
    p[q] = Phi[q];
    u[q] = U[q];
    v[q] = V[q];
  }
  if (a.raw_filter) {
    const double c = a.resid_c[2 * b + 1];
    if (fabs(c) > 1e-20) {
      const double* r = a.R0p + (long long)b * 3 * plane + pix;
      e += c * *reinterpret_cast<const v4d*>(r);
      d += c * *reinterpret_cast<const v4d*>(r + plane);
      p += c * *reinterpret_cast<const v4d*>(r + 2 * plane);
    }
  }
  if (a.test == 1) {
    u = *reinterpret_cast<const v4d*>(a.uvic + ((long long)b * 2 + 0) * plane + pix);
    v = *reinterpret_cast<const v4d*>(a.uvic + ((long long)b * 2 + 1) * plane + pix);
  }
  v4d pmid = p;  // the filtered prev Phi of the diagnostics (the current level without the carry)
  if (a.phys_on) {
    double* ph = phys + pix;
    const v4d ep = *reinterpret_cast<const v4d*>(ph), ec = *reinterpret_cast<const v4d*>(ph + plane);
    const v4d dp = *reinterpret_cast<const v4d*>(ph + 2 * plane), dc = *reinterpret_cast<const v4d*>(ph + 3 * plane);
    const v4d pp = *reinterpret_cast<const v4d*>(ph + 4 * plane), pc = *reinterpret_cast<const v4d*>(ph + 5 * plane);
    v4d em = ec, dm = dc, pm = pc;
    if (a.do_ra) {
      double g = alpha;
      if (a.raw_filter) {
        const double w = cf[C_WILLIAMS];
        g = alpha * w / (1.0 - alpha * (1.0 - w));
      }
      em = ec + g * (ep - 2.0 * ec + e);
      dm = dc + g * (dp - 2.0 * dc + d);
      pm = pc + g * (pp - 2.0 * pc + p);
    }
    *reinterpret_cast<v4d*>(ph) = em;
    *reinterpret_cast<v4d*>(ph + plane) = e;
    *reinterpret_cast<v4d*>(ph + 2 * plane) = dm;
    *reinterpret_cast<v4d*>(ph + 3 * plane) = d;
    *reinterpret_cast<v4d*>(ph + 4 * plane) = pm;
    *reinterpret_cast<v4d*>(ph + 5 * plane) = p;
    *reinterpret_cast<v4d*>(ph + 6 * plane) = u;
    *reinterpret_cast<v4d*>(ph + 7 * plane) = v;
    pmid = pm;
  }
  if (DIAG && a.hist_fields) {
    double* h = a.hist_fields + (long long)b * 5 * a.hist_T * plane + (long long)a.hist_k * plane + pix;
    const long long fs = (long long)a.hist_T * plane;
    *reinterpret_cast<v4d*>(h) = e;
    *reinterpret_cast<v4d*>(h + fs) = d;
    *reinterpret_cast<v4d*>(h + 2 * fs) = p;
    *reinterpret_cast<v4d*>(h + 3 * fs) = u;
    *reinterpret_cast<v4d*>(h + 4 * fs) = v;
  }
  if (DIAG && diag) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double uc = u[q] / cosphi, vc = v[q] / cosphi;
      const double sp = sqrt(u[q] * u[q] + v[q] * v[q]);
      diag[0] += wrms * (uc * uc + vc * vc) / aplanet;
      diag[1] = diag[1] < sp ? diag[1] : sp;
      diag[2] = diag[2] < pmid[q] ? diag[2] : pmid[q];
      diag[3] = diag[3] > pmid[q] ? diag[3] : pmid[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    eta[q] = (float)e[q];
    delta[q] = (float)d[q];
    Phi[q] = (float)p[q];
    U[q] = (float)u[q];
    V[q] = (float)v[q];
  }
}

// Assembly of the 9 Legendre inputs of 16 rows (member b, latitudes j0..j0+15)
// from the forward-DFT accumulators: wave w holds acc[2t + h] = fields 2w + h,
// wavenumber rows 16t + 4 (lane/16) + q, latitude column lane % 16.  Two
// 16-row k' tiles go through LDS (sC, [2][8][16][17] floats: an odd row stride, so the
// tile writes and the per-thread row reads hit distinct banks) per round, then one
// (m, row) pair per thread assembles (swx_physics.h:assemble_slots, with the row's
// constants computed once) and writes T.  U_m, V_m of the step are read from the
// Fourier rows G.
//
// Row order of the accumulator tiles: TE == 0, row r of the 16 TE-free tiles is
// (m, re/im) = (r / 2, r % 2); TE > 0 (half-period folded forward DFT), tiles t < TE hold the
// even wavenumbers (m = 2 (r / 2)) and tiles t >= TE the odd ones (m = 2 (r' / 2) + 1, r' the
// row counted from tile TE).
constexpr int kAsmStride = 17;
constexpr int kAsmFloats = 2 * 8 * 16 * kAsmStride;

// Wavenumber handled by this thread in assembly round t0 / 2 (-1: none).
template <int Mp, int NT2, int TE>
__device__ __forceinline__ int assemble_m(int t0) {
  const int tid = threadIdx.x, dt2 = tid >> 7, ml = (tid >> 4) & 7;
  const int tt = t0 + dt2;
  const int m = TE == 0 ? 8 * tt + ml : (tt < TE ? 2 * (8 * tt + ml) : 2 * (8 * (tt - TE) + ml) + 1);
  return (tt < NT2 && m < Mp) ? m : -1;
}

// U_m, V_m of every assembly round of this thread, loaded ahead (the caller issues this
// before its forward-DFT GEMM so that the loads complete under the matrix-core work).  Up
// to 4 rounds (T42, T63); with more the registers are needed elsewhere and each round
// loads its own.
template <int NT2>
constexpr bool kAsmPrefetch = (NT2 + 1) / 2 <= 4;

template <int Mp, int NT2, int TE = 0, bool PRE = true>
__device__ __forceinline__ void assemble_prefetch_uv(const GridArgs<float>& a, int b, int j0,
                                                     float2 (&uv)[(NT2 + 1) / 2][2]) {
  if constexpr (!(PRE && kAsmPrefetch<NT2>)) return;
  if (!a.si && !a.expflag) return;  // the modified-Euler slots do not read U_m, V_m
  const int jj = j0 + (threadIdx.x & 15);
#pragma unroll
  for (int r = 0; r < (NT2 + 1) / 2; ++r) {
    const int m = assemble_m<Mp, NT2, TE>(2 * r);
    const int mc = m < 0 ? 0 : m;  // in-bounds address; unused when m < 0
    uv[r][0] = *reinterpret_cast<const float2*>(a.G + fidx(b, 3, mc, jj, Mp, a.J));
    uv[r][1] = *reinterpret_cast<const float2*>(a.G + fidx(b, 4, mc, jj, Mp, a.J));
  }
}

// FAST: the caller guarantees the modified-Euler slots (no semi-implicit or explicit scheme).
template <int Mp, int NT2, int TE = 0, bool PRE = true, bool FAST = false>
__device__ __forceinline__ void grid_assemble(const GridArgs<float>& a, const v4f* acc, float* sC, int b, int j0,
                                              const double* cf, const CoefLayout& cl, const double* geo,
                                              const GeoLayout& gl, const float2 (&uv)[(NT2 + 1) / 2][2]) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int J = a.J;
  // This thread's latitude is the same in every round.
  const int dt2 = tid >> 7, ml = (tid >> 4) & 7, nn = tid & 15;
  const int jj = j0 + nn;
  const double tcj = cf[cl.n_tc + jj], wj = geo[gl.w + jj];
  const SlotConsts sk = slot_consts(cf, tcj, wj, a.expflag, a.forcflag);
  const SlotConstsF skf = slot_consts_f32(sk);
#pragma unroll
  for (int t0 = 0; t0 < NT2; t0 += 2) {
#pragma unroll
    for (int d2 = 0; d2 < 2; ++d2) {
      const int t = t0 + d2;
      if (t < NT2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int f = 2 * wave + h;
#pragma unroll
          for (int q = 0; q < 4; ++q) sC[((d2 * 8 + f) * 16 + 4 * lg + q) * kAsmStride + ln] = acc[2 * t + h][q];
        }
      }
    }
    __syncthreads();
    {
      const int m = assemble_m<Mp, NT2, TE>(t0);
      if (m >= 0) {
        float fs[8][2];
#pragma unroll
        for (int f = 0; f < 8; ++f) {
          fs[f][0] = sC[((dt2 * 8 + f) * 16 + 2 * ml) * kAsmStride + nn];
          fs[f][1] = sC[((dt2 * 8 + f) * 16 + 2 * ml + 1) * kAsmStride + nn];
        }
        if (a.export_raw) {
          float* rb = a.raw + (long long)b * kNRaw * J * Mp * 2;
#pragma unroll
          for (int f = 0; f < kNRaw; ++f)
            *reinterpret_cast<float2*>(rb + ((long long)f * J + jj) * Mp * 2 + 2 * m) = make_float2(fs[f][0], fs[f][1]);
        }
        float* Tm = a.T + (((long long)b * Mp + m) * kNT) * J * 2;
        if (FAST || (!a.si && !a.expflag)) {
          float out[kNT][2];
          assemble_slots_f32(skf, m, fs, out);
#pragma unroll
          for (int k = 0; k < kNT; ++k)
            *reinterpret_cast<float2*>(Tm + (k * J + jj) * 2) = make_float2(out[k][0], out[k][1]);
        } else {
          double fm[8][2];
#pragma unroll
          for (int f = 0; f < 8; ++f) {
            fm[f][0] = fs[f][0];
            fm[f][1] = fs[f][1];
          }
          float2 gu, gv;
          if constexpr (PRE && kAsmPrefetch<NT2>) {
            gu = uv[t0 / 2][0];
            gv = uv[t0 / 2][1];
          } else {
            gu = *reinterpret_cast<const float2*>(a.G + fidx(b, 3, m, jj, Mp, J));
            gv = *reinterpret_cast<const float2*>(a.G + fidx(b, 4, m, jj, Mp, J));
          }
          double Um[2] = {gu.x, gu.y}, Vm[2] = {gv.x, gv.y};
          double out[kNT][2];
          if (a.si) {
            double lagR[2][2];
            si_lag_exchange(a, cf, b, jj, m, fm, Um, Vm, false, lagR);
            assemble_slots_si(cf, tcj, wj, m, a.forcflag, a.relax, fm, lagR, out);
          } else {
            assemble_slots(sk, m, fm, Um, Vm, out);
          }
#pragma unroll
          for (int k = 0; k < kNT; ++k)
            *reinterpret_cast<float2*>(Tm + (k * J + jj) * 2) = make_float2((float)out[k][0], (float)out[k][1]);
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace swx
