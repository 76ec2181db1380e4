// Kernel argument blocks and launchers shared by the plan code.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace swx {

// Per-member coefficient table (FP64), layout [B][ncoef]:
//   scalars [0, kNScal), then sigma[N+1], sigmaPhi[N+1], tcmn[N+1], tc[J], areac[J],
//   sigma_exp[N+1], sigmaPhi_exp[N+1].
enum : int {
  C_DT = 0,
  C_A,
  C_OMEGA,
  C_PHIBAR,
  C_TAURAD,
  C_TAUDRAG,
  C_DPHIEQ,
  C_ALPHA,
  C_BLOWUP,
  C_INVTD,      // 0 if taudrag == -1 else 1/taudrag
  C_AREAPLANET, // 4 pi a^2
  C_FMN01,      // omega / sqrt(0.375)
  C_INVTAURAD,  // 1 / taurad
  C_SIALPHA,    // semi-implicit off-centering
  C_WILLIAMS,   // RAW filter weight
  kNScal = 16
};

struct CoefLayout {
  int n_sigma, n_sigmaphi, n_tcmn, n_tc, n_areac, n_sigexp, n_sigexpphi, ncoef;
  __host__ __device__ static CoefLayout make(int N, int J) {
    CoefLayout c;
    c.n_sigma = kNScal;
    c.n_sigmaphi = c.n_sigma + (N + 1);
    c.n_tcmn = c.n_sigmaphi + (N + 1);
    c.n_tc = c.n_tcmn + (N + 1);
    c.n_areac = c.n_tc + J;
    c.n_sigexp = c.n_areac + J;          // exponential hyperdiffusion (semi-implicit scheme)
    c.n_sigexpphi = c.n_sigexp + (N + 1);
    c.ncoef = c.n_sigexpphi + (N + 1);
    return c;
  }
};

// Scheme flags of one step of the spectral update (host fills them per launch).
struct StepFlags {
  int expflag, diffflag;
  int si;      // semi-implicit leapfrog
  int raw;     // Robert-Asselin-Williams filter
  int relax;   // semi-implicit: implicit relaxation/drag (forcflag and test None)
  int do_ra;   // time filter active this step (modalflag and t > 2)
  // Modified Euler: carry the time-filtered prev lineage of the physical fields in the
  // spectral prev level, S_prev' = S_curr + alpha (S_prev - 2 S_curr + S_new) when filtered,
  // else S_curr (the same recurrence the physical prev level follows, in spectral form).
  int lineage;
};

// Shared geometry table (FP64): mus[J], w[J], s1m[J], cosphi[J], denomE[J],
// cosl[I], day[I] (1.0 on the strict dayside, else 0.0), dcos[I] = day * cosl.
struct GeoLayout {
  int mus, w, s1m, cosphi, denomE, cosl, day, dcos, n;
  __host__ __device__ static GeoLayout make(int J, int I) {
    GeoLayout g;
    g.mus = 0;
    g.w = J;
    g.s1m = 2 * J;
    g.cosphi = 3 * J;
    g.denomE = 4 * J;
    g.cosl = 5 * J;
    g.day = 5 * J + I;
    g.dcos = 5 * J + 2 * I;
    g.n = 5 * J + 3 * I;
    return g;
  }
};

// Number of Legendre-input slots written by the grid kernel per (member, m, j):
// P-basis slots 0..4, H-basis slots 5..8.
constexpr int kNP = 5;
constexpr int kNH = 4;
constexpr int kNT = kNP + kNH;
// Fourier fields produced by the Legendre kernel per (member, j, m):
// eta, delta, Phi, U, V.
constexpr int kNG = 5;
// Raw forward-DFT products per row: A, B, C, D, E, PhiF, F, G.
constexpr int kNRaw = 8;

// Degree rows padded to whole 16-row MFMA tiles.
__host__ __device__ constexpr int leg_knp(int N) { return (N + 1 + 15) / 16 * 16; }

// Equatorially folded Legendre contractions (FP32 matrix-core kernel).  The Gauss nodes pair
// as mu_{J-1-j} = -mu_j and the basis has parity P(-mu) = (-1)^{n+m} P(mu),
// H(-mu) = -(-1)^{n+m} H(mu), so the forward sums run over j < J/2 on s = T(j) + T(J-1-j) or
// a = T(j) - T(J-1-j), and the inverse forms the even-(n-m) and odd-(n-m) partial sums E, O on
// j < J/2 with X(j) = E + O, X(J-1-j) = E - O (P; H: O - E).  Per m, with K = N + 1 - m degrees:
// Ke = (K + 1) / 2 even and Ko = K / 2 odd ones.
//   forward: row tiles of 16 degrees, the ceil(Ke/16) even tiles first, then the odd ones;
//   inverse: the rows of the inputs are the Ke even degrees then the Ko odd ones; E takes
//   k-steps [0, leg_fold_ne(K)), O the k-steps [leg_fold_kb(K), leg_fold_kb(K) + leg_fold_no(K))
//   (each table zero on the other parity's rows).
__host__ __device__ constexpr int leg_fold_nte(int K) { return ((K + 1) / 2 + 15) / 16; }
__host__ __device__ constexpr int leg_fold_nto(int K) { return (K / 2 + 15) / 16; }
__host__ __device__ constexpr int leg_fold_ne(int K) { return ((K + 1) / 2 + 3) / 4; }
__host__ __device__ constexpr int leg_fold_kb(int K) { return ((K + 1) / 2) / 4; }
__host__ __device__ constexpr int leg_fold_no(int K) { return (K + 3) / 4 - leg_fold_kb(K); }
// Table extents over all m (K = 1 .. N + 1): forward row tiles per m, inverse E and O k-step
// slots per lane (each padded to a multiple of 4: float4 loads).
__host__ __device__ constexpr int leg_fold_ntf(int N) {
  int r = 0;
  for (int K = 1; K <= N + 1; ++K) r = leg_fold_nte(K) + leg_fold_nto(K) > r ? leg_fold_nte(K) + leg_fold_nto(K) : r;
  return r;
}
__host__ __device__ constexpr int leg_fold_kse(int N) {
  int r = 0;
  for (int K = 1; K <= N + 1; ++K) r = leg_fold_ne(K) > r ? leg_fold_ne(K) : r;
  return (r + 3) / 4 * 4;
}
__host__ __device__ constexpr int leg_fold_kso(int N) {
  int r = 0;
  for (int K = 1; K <= N + 1; ++K) r = leg_fold_no(K) > r ? leg_fold_no(K) : r;
  return (r + 3) / 4 * 4;
}

// Work-item order of a Legendre launch with grid (M+1, groups).  heavy_first = 1: workgroup
// L = x + y (M+1) takes zonal wavenumber m = L / groups and group L % groups; dispatch runs in
// L order, so the heavy items (small m: K = N + 1 - m degrees) start first and the light ones
// fill the tail, and the workgroups resident at any time share a few m's basis tables (L2
// locality).  heavy_first = 0: m = x, group = y (every group's m's in turn), which measures
// faster when the launch shares the GPU with concurrently stepped plans.
__device__ __forceinline__ void leg_item(int heavy_first, int& m, int& grp) {
  if (heavy_first) {
    const int L = blockIdx.x + blockIdx.y * gridDim.x;
    m = L / gridDim.y;
    grp = L % gridDim.y;
  } else {
    m = blockIdx.x;
    grp = blockIdx.y;
  }
}

// Fourier rows G / Gprev / Gbar are stored m-major: [member][field][m][j]
// complex, so a Legendre block (one m) reads and writes contiguous j.  Offset
// in reals of (b, f, m, j):
__host__ __device__ inline long long fidx(int b, int f, int m, int j, int Mp, int J) {
  return ((((long long)b * kNG + f) * Mp + m) * J + j) * 2;
}

// Offset (in elements of one J-row block) of the triangular basis block of m:
// the block holds J * (N + 1 - m) values.
__host__ __device__ inline long long tri_off(int m, int N, int J) {
  return (long long)J * ((long long)m * (N + 1) - (long long)m * (m - 1) / 2);
}

enum LegMode : int { LEG_INIT = 0, LEG_STEP = 1, LEG_PRIME = 2 };
enum GridMode : int { GRID_DFT = 0, GRID_INIT = 1, GRID_STEP = 2 };

template <typename R>
struct LegArgs {
  int M, N, J, B;
  int mode, expflag, diffflag, diagnostics, diag_winds;
  const R* Pf;  // forward layout  [m-block][j][n-m]
  const R* Hf;
  const R* Pi;  // inverse layout  [m-block][n-m][j]
  const R* Hi;
  // MFMA-fragment tables of the same blocks (FP32 plans; KNP = N+1 rounded up to 16):
  //   fwd[m][nt][lane][ks] = basis(j = 4 ks + lane/16, n = m + 16 nt + lane%16)   ks < J/4
  //   inv[m][jt][lane][ks] = basis(j = 16 jt + lane%16, n = m + 4 ks + lane/16)   ks < KNP/4
  // zero where n > N; each lane's ks run is contiguous (float4 loads).
  const float* fwdP;
  const float* fwdH;
  const float* invP;
  const float* invH;
  // Equatorially folded tables (leg_fold_* extents; J2 = J / 2):
  //   ffwd[m][r][lane][ks] = basis(j = 4 ks + lane/16, n = m + np), ks < J2 / 4, with row
  //     tile r < nte: np = 2 (16 r + lane%16), else np = 2 (16 (r - nte) + lane%16) + 1;
  //   finv[m][jt][lane][slot] = basis(j = 16 jt + lane%16, n = m + np(row)), jt < J2 / 16, E
  //     slots e < kse: row 4 e + lane/16 (np = 2 row while row < Ke), O slots kse + o: row
  //     4 (kb + o) + lane/16 (np = 2 (row - Ke) + 1 while Ke <= row < K); zero elsewhere.
  const float* ffwdP;
  const float* ffwdH;
  const float* finvP;
  const float* finvH;
  int leg_fold;  // 1: the matrix-core kernel uses the folded tables
  int bf16x3;    // 1: its contractions as split-BF16x3 products (folded kernel only)
  const double* coef;
  int ncoef;
  // [M+1][N+1]: Pw(invrs_leg(X)) = f2 * X for the basis in use (1 for an
  // orthonormal basis; < 1 or 0 where the reference normalization underflows).
  const double* f2;
  double* S;        // [B][2][3][M+1][N+1] complex, level 0 = prev, 1 = curr
  const R* T;       // [B][M+1][kNT][J] complex
  R* G;             // [B][kNG][J][M+1] complex
  const double* part_uv;  // [B][J][2]
  unsigned char* dead;    // [B]
  double* hist_scal;      // [B][hist_T][4] or null
  unsigned char* hist_dead;
  int hist_T, hist_k;
  // Semi-implicit / RAW (LEG_STEP): scheme flags of this step and the per-member
  // coefficients (c_prev, c_curr) of the initial fields' non-band-limited residue
  // in the filtered lineages, [B][2] (advanced by the m = 0 workgroups).
  int si, raw_filter, relax, do_ra;
  double* resid_c;
  // LEG_PRIME: 1 = transform the prev level instead of the current one.
  int prime_prev;
  // LEG_STEP, modified Euler: 1 = advance the spectral prev lineage (StepFlags::lineage).
  int lineage;
  __host__ __device__ StepFlags flags() const {
    return StepFlags{expflag, diffflag, si, raw_filter, relax, do_ra, lineage};
  }
};

template <typename R>
struct GridArgs {
  int M, N, J, I, B;
  int mode, expflag, forcflag, test, diagnostics, do_ra, phys_on, export_raw, diag_winds, have_uv, init_from_g;
  const R* tw;         // [2][I]: cos(2 pi k / I), sin(2 pi k / I)
  const double* geo;
  const double* coef;
  int ncoef;
  const R* G;          // [B][kNG][J][M+1] (GRID_STEP, GRID_INIT with diag_winds)
  R* Gout;             // GRID_DFT / GRID_INIT write targets
  R* T;                // [B][M+1][kNT][J]
  const double* in_eta;  // GRID_DFT / GRID_INIT inputs [B][J][I]
  const double* in_delta;
  const double* in_Phi;
  const double* in_U;
  const double* in_V;
  double* phys;        // [B][8][J][I]: eta p,c; delta p,c; Phi p,c; U; V
  double* uvic;        // [B][2][J][I] pinned winds (test == 1)
  double* part_uv;     // [B][J][2]
  const unsigned char* dead;
  double* hist_fields; // [B][5][hist_T][J][I]
  double* hist_phiext; // [B][hist_T][J][2]
  int hist_T, hist_k;
  R* raw;              // [B][kNRaw][J][M+1]
  R* lag;              // [B][2][J][M+1]
  R* g0uv;             // [B][2][J][M+1]
  // Semi-implicit: the Legendre inputs use the lagged momentum remainder `lag`
  // (read, then overwritten with the current one; the value read is copied to
  // lagx on export_raw steps).  RAW: the physical fields gain c_curr * R0p (the
  // initial fields' residue, [B][3][J][I]) and the physical carry follows the
  // Williams filter.
  int si, raw_filter, relax;
  R* lagx;
  const double* resid_c;  // [B][2]
  const double* R0p;
};

template <typename R>
hipError_t launch_leg(const LegArgs<R>& a, hipStream_t s);
// The blow-up gate of a diagnostics step on its own (swx_kernels.hip:k_blowup_gate), for the
// matrix-core Legendre kernel, which then skips the members it marks dead.
template <typename R>
hipError_t launch_blowup_gate(const LegArgs<R>& a, hipStream_t s);
template <typename R>
hipError_t launch_grid(const GridArgs<R>& a, hipStream_t s);
// MFMA grid kernel (FP32, GRID_STEP without history/diagnostics).  DFT tables
// in MFMA-fragment order (lane = 16 * (lane / 16) + lane % 16, ln = lane % 16,
// lg = lane / 16), zero outside the matrices:
//   inv[it][lane][ks] = Tinv(i = 16 it + ln, k = 4 ks + lg),       ks < kq1
//   fwd[t][g][lane][q] = Tfwd(k' = 16 t + ln, i = 16 g + 4 q + lg), q < 4
// with Tinv(i, 2m) = w_m cos(2 pi m i / I), Tinv(i, 2m+1) = -w_m sin(...),
// Tfwd(2m, i) = cos(2 pi m i / I) / I, Tfwd(2m+1, i) = -sin(...) / I.
hipError_t launch_grid_mfma(const GridArgs<float>& a, const float* TinvF, const float* TfwdF, hipStream_t s);
int grid_mfma_kq1(int M);
size_t grid_mfma_lds_bytes(int M);
// Half-period folded MFMA grid kernel (same eligibility; half the matrix-core work).  k rows
// (m, re/im) in folded order: slot s = k / 2 is m = 2 s for s < KE4 / 2, else
// m = 2 (s - KE4 / 2) + 1; output tile t < TE holds even m (row r: m = 2 (r / 2)), t >= TE odd
// m (row r counted from tile TE: m = 2 (r / 2) + 1); invalid where m > M.  Tables:
//   inv[it][lane][ks] = Tinv(i = 16 it + ln, k(4 ks + lg)),  it < H / 16, ks < KQ1
//   fwd[t][g][lane][q] = Tfwd(k(16 t + ln), i = 16 g + 4 q + lg),  g < H / 16, q < 4
struct FoldDims {
  int H, K1, KE4, KQ1, TE, NT2;
};
bool grid_mfma_fold_dims(int M, FoldDims* d);
size_t grid_mfma_fold_lds_bytes(int M);
// shared = true: the launch shares the device with another plan's kernels (swx_pipeline), and
// the steady fast variant runs at its lower occupancy, which leaves room for them.
hipError_t launch_grid_mfma_fold(const GridArgs<float>& a, const float* TinvF, const float* TfwdF, hipStream_t s,
                                 bool shared = false);
// Split-BF16x3 grid kernel (same eligibility as the MFMA grid kernel); tables
// of bf16 hi / lo pairs in its fragment order (swx_bf16x3.hip).
hipError_t launch_grid_bf16x3(const GridArgs<float>& a, const void* TinvB, const void* TfwdB, hipStream_t s);
size_t grid_bf16x3_lds_bytes(int M);
int grid_bf16x3_ks1(int M);
// Two FP32 plans of different truncations (M0 < M1: 42/63, 42/106, 63/106) in one launch.
bool mfma_pair_supported(int M0, int M1);
// fold0 / fold1: that plan's half-period folded variant (its tables passed as ti / tf).
hipError_t launch_grid_mfma_pair(const GridArgs<float>& a0, const float* ti0, const float* tf0, bool fold0,
                                 const GridArgs<float>& a1, const float* ti1, const float* tf1, bool fold1,
                                 hipStream_t s);
hipError_t launch_leg_mfma_pair(const LegArgs<float>& a0, const LegArgs<float>& a1, hipStream_t s);
// MFMA Legendre kernel (FP32, LEG_STEP without diagnostics).
// mb = members per workgroup, one of leg_mfma_candidates(M) (the plan picks the fastest).
// shared = true: the launch shares the GPU with concurrently stepped plans (work items in
// member-group order instead of heaviest first).
hipError_t launch_leg_mfma(const LegArgs<float>& a, int mb, hipStream_t s, bool shared = false);
size_t leg_mfma_lds_bytes(int M, int mb);
int leg_mfma_candidates(int M, int* mbs);  // up to 3 values
int leg_mfma_default_mb(int M);
int grid_mfma_k2p(int M);

// One single-wave launch that spins for `ticks` of the device wall clock (bounded by it).
hipError_t launch_wait_ticks(unsigned long long ticks, hipStream_t s);
// Measured ceilings (swx_peaks.hip): out[0..2] matrix-core TFLOP/s (FP32 16x16x4, FP64 16x16x4,
// BF16 16x16x16), out[3..5] HBM TB/s (copy: bytes read + written; read-only; write-only).
hipError_t measure_ceilings(double* out, size_t bytes, hipStream_t s);
// dst <- src, `bytes` a multiple of 16 and both 16-byte aligned (hipErrorInvalidValue otherwise).
hipError_t launch_copy_words(void* dst, const void* src, size_t bytes, hipStream_t s);
template <typename R>
hipError_t launch_copy_real(double* dst, long long dstride, const R* src, long long sstride, long long n, int B,
                            hipStream_t s);
// Field f of an m-major Fourier-row buffer -> complex128 [B][J][M+1]; with `res`, plus
// rc[2 b + lvl] times the same field of `res` (the initial fields' residue).
template <typename R>
hipError_t launch_copy_fourier(double* dst, const R* src, int f, int B, int J, int Mp, hipStream_t s,
                               const R* res = nullptr, const double* rc = nullptr, int lvl = 0);
// Physical carry of a modified-Euler plan from its Fourier rows: prev levels (planes 0, 2, 4 of
// phys [B][8][J][I]) = inverse DFT of fields 0..2 of Q (the rows of the spectral prev lineage),
// current levels and winds (planes 1, 3, 5, 6, 7) = inverse DFT of fields 0..4 of G.
template <typename R>
hipError_t launch_phys_synth(const R* G, const R* Q, double* phys, const R* tw, int B, int J, int I, int Mp,
                             hipStream_t s);
// Initial fields' residue of semi-implicit / RAW plans (swx_kernels.hip:k_residue_init).
template <typename R>
hipError_t launch_residue_init(const R* G, R* R0f, double* R0p, const double* in_eta, const double* in_delta,
                               const double* in_Phi, const R* tw, double* resid_c, int B, int J, int I, int Mp,
                               hipStream_t s);

}  // namespace swx
