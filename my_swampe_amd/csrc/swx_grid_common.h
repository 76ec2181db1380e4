// Pieces shared by the matrix-core grid kernels (FP32 MFMA and split-BF16x3):
// the grid-space work of one point and the assembly of the Legendre inputs from
// the forward-DFT accumulators.
#pragma once

#include "swx_kernels.h"
#include "swx_physics.h"

namespace swx {

typedef float v4f __attribute__((ext_vector_type(4)));

// One grid point (row j, longitude i) of member b: the initial fields' residue
// (RAW), pinned winds (test 1), the time filter of the physical prev level and
// the physical carry (when a.phys_on; swx_physics.h:physical_carry), then the 8
// products of the forward transform (forcing.py, initial_conditions.py:283-292)
// in reciprocal form (inv_denE = 1 / (2 (1 - mu_j^2))).
__device__ __forceinline__ void grid_point_step(const GridArgs<float>& a, const double* cf, const double* geo,
                                                const GeoLayout& gl, double* phys, long long plane, double alpha,
                                                int b, int j, int i, double eta, double delta, double Phi, double U,
                                                double V, double inv_denE, double prod[8]) {
  const long long pix = (long long)j * a.I + i;
  if (a.raw_filter) add_residue(a.resid_c, a.R0p, b, plane, pix, eta, delta, Phi);
  if (a.test == 1) {
    U = a.uvic[((long long)b * 2 + 0) * plane + pix];
    V = a.uvic[((long long)b * 2 + 1) * plane + pix];
  }
  if (a.phys_on) {
    double Phi_mid;
    physical_carry(cf, a.do_ra, a.raw_filter, alpha, phys, plane, pix, eta, delta, Phi, U, V, Phi_mid);
  }
  grid_products_recip(cf, geo, gl, j, i, a.test, eta, Phi, U, V, inv_denE, prod);
}

// Assembly of the 9 Legendre inputs of 16 rows (member b, latitudes j0..j0+15)
// from the forward-DFT accumulators: wave w holds acc[2t + h] = fields 2w + h,
// wavenumber rows 16t + 4 (lane/16) + q, latitude column lane % 16.  Two
// 16-row k' tiles go through LDS (sC, 2*8*16*16 floats) per round, then one
// (m, row) pair per thread assembles (swx_physics.h:assemble_slots) and writes T.
// U_m, V_m of the step are read from the Fourier rows G.
//
// Row order of the accumulator tiles: TE == 0, row r of the 16 TE-free tiles is
// (m, re/im) = (r / 2, r % 2); TE > 0 (half-period folded forward DFT), tiles t < TE hold the
// even wavenumbers (m = 2 (r / 2)) and tiles t >= TE the odd ones (m = 2 (r' / 2) + 1, r' the
// row counted from tile TE).
template <int Mp, int NT2, int TE = 0>
__device__ __forceinline__ void grid_assemble(const GridArgs<float>& a, const v4f* acc, float* sC, int b, int j0,
                                              const double* cf, const CoefLayout& cl, const double* geo,
                                              const GeoLayout& gl) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, ln = lane & 15, lg = lane >> 4;
  const int J = a.J;
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
      const int tt = t0 + dt2;
      const int m = TE == 0 ? 8 * tt + ml : (tt < TE ? 2 * (8 * tt + ml) : 2 * (8 * (tt - TE) + ml) + 1);
      if (t0 + dt2 < NT2 && m < Mp) {
        const int jj = j0 + nn;
        double fm[8][2];
#pragma unroll
        for (int f = 0; f < 8; ++f) {
          fm[f][0] = sC[((dt2 * 8 + f) * 16 + 2 * ml) * 16 + nn];
          fm[f][1] = sC[((dt2 * 8 + f) * 16 + 2 * ml + 1) * 16 + nn];
        }
        const float* gu = a.G + fidx(b, 3, m, jj, Mp, J);
        const float* gv = a.G + fidx(b, 4, m, jj, Mp, J);
        double Um[2] = {gu[0], gu[1]}, Vm[2] = {gv[0], gv[1]};
        if (a.export_raw) {
          float* rb = a.raw + (long long)b * kNRaw * J * Mp * 2;
#pragma unroll
          for (int f = 0; f < kNRaw; ++f) {
            rb[((long long)f * J + jj) * Mp * 2 + 2 * m] = (float)fm[f][0];
            rb[((long long)f * J + jj) * Mp * 2 + 2 * m + 1] = (float)fm[f][1];
          }
        }
        double out[kNT][2];
        if (a.si) {
          double lagR[2][2];
          si_lag_exchange(a, cf, b, jj, m, fm, Um, Vm, false, lagR);
          assemble_slots_si(cf, cf[cl.n_tc + jj], geo[gl.w + jj], m, a.forcflag, a.relax, fm, lagR, out);
        } else {
          assemble_slots(cf, cf[cl.n_tc + jj], geo[gl.w + jj], m, a.expflag, a.forcflag, fm, Um, Vm, out);
        }
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

}  // namespace swx
