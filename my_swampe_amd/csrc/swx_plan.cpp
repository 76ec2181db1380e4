// Plan management and the extern "C" entry points of libswampe_hip.so.
//
// A plan holds one truncation group (M, J, I) with B ensemble members: the
// FP64 geometry (host), the basis/twiddle tables on the device in the
// transform precision, per-member coefficient tables, and the device carry.
// Stepping is two kernel launches per step (k_leg, k_grid); forward-only
// stepping (swx_run_graph) replays captured HIP graphs of the step sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/swampe_capi.h"
#include "geometry.h"
#include "swx_adjoint.h"
#include "swx_kernels.h"
#include "swx_observe.h"
#include "swx_transforms.h"

using namespace swx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define SWX_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(SWX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));                 \
  } while (0)

struct Geometry {
  int M = 0, N = 0, J = 0, I = 0;
  std::vector<double> mus, w, lam, P, H;
};

// Geometry of truncation M; quadrature from (qmus, qw) when given (checked by the caller).
bool build_geometry(int M, bool exact_basis, Geometry& g, const double* qmus = nullptr, const double* qw = nullptr) {
  int J, I;
  if (!resolution(M, &J, &I, nullptr)) return false;
  g.M = M;
  g.N = M;
  g.J = J;
  g.I = I;
  g.mus.resize(J);
  g.w.resize(J);
  g.lam.resize(I);
  if (qmus && qw) {
    g.mus.assign(qmus, qmus + J);
    g.w.assign(qw, qw + J);
  } else {
    gauss_legendre(J, g.mus.data(), g.w.data());
  }
  lambdas(I, g.lam.data());
  g.P.resize((size_t)J * (M + 1) * (M + 1));
  g.H.resize(g.P.size());
  legendre_basis(M, M, J, g.mus.data(), g.P.data(), g.H.data(), !exact_basis);
  return true;
}

}  // namespace

// Captured step sequences of a plan or an ensemble, keyed by their shape: steps, the first
// grid-carry and the first lineage step of each plan (equal to steps where none), whether the last
// step saves the previous Fourier rows, and the stepping mode (standalone / in a pipeline).  Every
// captured step has t > 2, so the same key always records the same launches.  Up to kSlots are
// kept; the least recently replayed one is replaced.
struct GraphCache {
  static constexpr int kSlots = 8, kKey = 7;
  struct Slot {
    hipGraphExec_t exec = nullptr;
    int key[kKey] = {};
    hipStream_t stream = nullptr;
    unsigned long long used = 0;
    hipEvent_t done = nullptr;  // recorded after each launch of exec
  };
  Slot slot[kSlots];
  unsigned long long clock = 0;
  // A graph is destroyed only once its last launch has completed (its kernel arguments are the
  // executable graph's own memory).
  static void release(Slot& sl) {
    if (sl.exec) {
      if (sl.done) (void)hipEventSynchronize(sl.done);
      (void)hipGraphExecDestroy(sl.exec);
    }
    if (sl.done) (void)hipEventDestroy(sl.done);
    sl = Slot{};
  }
  void clear() {
    for (Slot& sl : slot) release(sl);
  }
  int live() const {
    int n = 0;
    for (const Slot& sl : slot) n += sl.exec != nullptr;
    return n;
  }
};

struct swx_plan {
  swx_plan_desc d{};
  Geometry geo;
  int device = 0;
  size_t rsz = 8;  // bytes per transform real
  CoefLayout cl{};
  GeoLayout gl{};
  std::vector<double> params;  // [B][SWX_NPARAM]
  // device buffers
  void* Pf = nullptr;
  void* Hf = nullptr;
  void* Pi = nullptr;
  void* Hi = nullptr;
  float* fragP[4] = {nullptr, nullptr, nullptr, nullptr};  // MFMA Legendre tables: fwd P, fwd H, inv P, inv H
  float* fragF[4] = {nullptr, nullptr, nullptr, nullptr};  // the same, equatorially folded (LegArgs ffwd*, finv*)
  int leg_fold = 2;  // SWX_OPT_LEG_FOLD: 1 = full contractions, 2 = folded
  void* tw = nullptr;
  float* TinvF = nullptr;  // MFMA grid kernel DFT tables, fragment order (FP32 plans)
  float* TfwdF = nullptr;
  float* TinvH = nullptr;  // tables of the half-period folded MFMA grid kernel
  float* TfwdH = nullptr;
  bool has_fold = false;   // folded kernel available (its LDS fits)
  int grid_fold = 1;       // SWX_OPT_GRID_FOLD in use: 1 = full DFTs, 2 = folded
  bool use_mfma = false;       // MFMA grid kernel available for this plan
  bool use_bf16x3 = false;     // split-BF16x3 grid kernel (precision SWX_PREC_BF16X3)
  bool use_adj_mfma = false;   // MFMA grid adjoint kernel (FP32 transforms, where its LDS fits)
  bool use_leg_adj_mfma = false;  // MFMA Legendre adjoint kernels (FP32 transforms)
  bool use_adj_fold = false;      // folded MFMA grid adjoint available (its LDS fits; two passes at T106)
  void* TinvB = nullptr;       // its DFT tables: bf16 hi / lo pairs in fragment order
  void* TfwdB = nullptr;
  bool opt_plain_grid = false;  // SWX_OPT_GRID_KERNEL = 1
  bool opt_full_phys = false;   // SWX_OPT_FULL_PHYS_CARRY = 1
  bool in_pipeline = false;     // stepped beside other plans by a swx_pipeline
  bool pipe_same_m = false;     // ... whose plans all share this plan's truncation
  bool use_leg_mfma = false;    // MFMA Legendre kernel available
  int leg_mb = 0;               // its members per workgroup (timed at plan creation)
  float leg_mb_ms[3] = {0, 0, 0};  // the candidates' measured launch times
  bool opt_plain_leg = false;   // SWX_OPT_LEG_KERNEL = 1
  int opt_phys_steps = 0;       // SWX_OPT_PHYS_STEPS (0 = automatic, -1 = never)
  int opt_lineage = 0;          // SWX_OPT_PHYS_LINEAGE (0 = automatic, -1 = off, k > 0 = window)
  // The grid carry `phys` is behind the state: the last call advanced the physical prev level as
  // the spectral lineage (S prev level), and `phys` is rebuilt from the rows on demand.
  bool phys_stale = false;
  void* Gsyn = nullptr;         // rows of the lineage for that rebuild (allocated on first use)
  // Steps run since creation: graph-replayed, eager, with the lineage, with the grid carry written.
  long long nstep[4] = {0, 0, 0, 0};
  double* geo_d = nullptr;
  double* f2 = nullptr;
  double* coef = nullptr;
  double* S = nullptr;
  void* T = nullptr;
  void* G = nullptr;
  void* Gprev = nullptr;
  void* raw = nullptr;
  void* lag = nullptr;
  void* g0uv = nullptr;
  double* phys = nullptr;
  double* uvic = nullptr;
  // semi-implicit / RAW plans: the initial fields' residue (Fourier rows in the G layout,
  // physical [B][3][J][I] for RAW), its per-member coefficients [B][2] and the exported
  // lagged momentum remainder (semi-implicit)
  void* R0f = nullptr;
  double* R0p = nullptr;
  double* resid_c = nullptr;
  void* lagx = nullptr;
  double* part_uv = nullptr;
  unsigned char* dead = nullptr;
  bool initialized = false;
  // adjoint workspace (allocated on first use)
  double* Lam = nullptr;
  void* Tbar = nullptr;
  void* Gbar = nullptr;
  double* pgrad = nullptr;
  double* dgrad = nullptr;        // diffusion partials [B][2] (dL/dK6, dL/dK6Phi)
  bool opt_adj_diffusion = false; // SWX_OPT_ADJ_DIFFUSION
  void* adj_stc = nullptr;       // S_t cache of the folded MFMA grid adjoint (T106)
  // captured forward step sequences (swx_run_graph)
  GraphCache graphs;
  unsigned opt_epoch = 0;        // swx_plan_set_option calls so far (ensembles re-capture on a change)
  size_t dev_bytes = 0;  // device memory held by the plan
};

namespace {

// Device allocation owned by `owner` (counted in owner->dev_bytes).
template <typename T>
int dalloc(swx_plan* owner, T** p, size_t bytes) {
  void* q = nullptr;
  const size_t nb = bytes < 16 ? 16 : bytes;
  hipError_t e = hipMalloc(&q, nb);
  if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  *p = static_cast<T*>(q);
  owner->dev_bytes += nb;
  return SWX_OK;
}

// bf16 (round to nearest even) bit pattern of a float, and its value.
uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

float bf16_value(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// x = hi + lo + O(2^-16 |x|): hi = bf16(x), lo = bf16(x - hi).
void bf16_split(double x, uint16_t& hi, uint16_t& lo) {
  hi = bf16_bits((float)x);
  lo = bf16_bits((float)(x - (double)bf16_value(hi)));
}

// Upload a host double array in the transform precision.
int upload_real(void* dst, const std::vector<double>& src, size_t rsz) {
  if (rsz == 8) {
    hipError_t e = hipMemcpy(dst, src.data(), src.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  } else {
    std::vector<float> f(src.size());
    for (size_t k = 0; k < src.size(); ++k) f[k] = (float)src[k];
    hipError_t e = hipMemcpy(dst, f.data(), f.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  return SWX_OK;
}

// Per-member coefficient tables (model.py:498-518; time_stepping.py:491-522;
// filters.py:95-151; forcing.py:94; time_stepping.py:568-573).
void fill_coef(const swx_plan* p, std::vector<double>& out) {
  const int N = p->geo.N, J = p->geo.J, B = p->d.n_members;
  const CoefLayout& cl = p->cl;
  out.assign((size_t)B * cl.ncoef, 0.0);
  const double pi = 3.14159265358979323846;
  const std::vector<double>& mus = p->geo.mus;
  const std::vector<double>& lam = p->geo.lam;
  for (int b = 0; b < B; ++b) {
    const double* pr = &p->params[(size_t)b * SWX_NPARAM];
    double* c = &out[(size_t)b * cl.ncoef];
    const double dt = pr[SWX_P_DT], a = pr[SWX_P_A], K6 = pr[SWX_P_K6];
    const double K6Phi = pr[SWX_P_K6PHI] < 0 ? K6 : pr[SWX_P_K6PHI];
    c[C_DT] = dt;
    c[C_A] = a;
    c[C_OMEGA] = pr[SWX_P_OMEGA];
    c[C_PHIBAR] = pr[SWX_P_PHIBAR];
    c[C_TAURAD] = pr[SWX_P_TAURAD];
    c[C_TAUDRAG] = pr[SWX_P_TAUDRAG];
    c[C_DPHIEQ] = pr[SWX_P_DPHIEQ];
    c[C_ALPHA] = pr[SWX_P_ALPHA];
    c[C_BLOWUP] = pr[SWX_P_BLOWUP];
    c[C_INVTD] = pr[SWX_P_TAUDRAG] == -1.0 ? 0.0 : 1.0 / pr[SWX_P_TAUDRAG];
    c[C_AREAPLANET] = 4.0 * pi * a * a;
    c[C_FMN01] = pr[SWX_P_OMEGA] / std::sqrt(0.375);
    c[C_INVTAURAD] = 1.0 / pr[SWX_P_TAURAD];
    c[C_SIALPHA] = pr[SWX_P_SI_ALPHA];
    c[C_WILLIAMS] = pr[SWX_P_WILLIAMS_ALPHA];
    const double a3 = a * a * a, a6 = a3 * a3;
    for (int n = 0; n <= N; ++n) {
      const double nv = n, n1 = n + 1.0;
      const double ncoeff = ((nv * nv * nv) / a3) * ((n1 * n1 * n1) / a3);
      c[cl.n_sigma + n] = 1.0 / (1.0 + (2.0 * dt * K6) * (ncoeff - 8.0 / a6));
      c[cl.n_sigmaphi + n] = 1.0 / (1.0 + (2.0 * dt * K6Phi) * ncoeff);
      c[cl.n_tcmn + n] = n == 0 ? 0.0 : a / (nv * n1);
      // Exponential hyperdiffusion of the semi-implicit leapfrog (filters.py:154-213).
      c[cl.n_sigexp + n] = std::exp(-(2.0 * dt * K6) * std::max(ncoeff - 8.0 / a6, 0.0));
      c[cl.n_sigexpphi + n] = std::exp(-(2.0 * dt * K6Phi) * ncoeff);
    }
    const double dl = lam[2] - lam[1];
    const double dp = std::asin(mus[2]) - std::asin(mus[1]);
    for (int j = 0; j < J; ++j) {
      c[cl.n_tc + j] = (2.0 * dt) / (a * (1.0 - mus[j] * mus[j]));
      const double sp = std::sin(std::asin(mus[j]) + pi / 2.0);
      c[cl.n_areac + j] = (a * a) * (sp * sp) * dp * dl;
    }
  }
}

// Semi-implicit relaxation/drag taken implicitly: forced runs with test None
// (model.py:905-925).
int relax_implicit(const swx_plan* p) { return p->d.semi_implicit && p->d.forcflag && p->d.test == -1 ? 1 : 0; }

// Time filter active at step t: modalflag and t > 2 (model.py:991).
int filter_on(const swx_plan* p, int t) { return (p->d.modalflag && t > 2) ? 1 : 0; }

template <typename R>
LegArgs<R> leg_args(swx_plan* p, int mode) {
  LegArgs<R> a{};
  a.M = p->geo.M;
  a.N = p->geo.N;
  a.J = p->geo.J;
  a.B = p->d.n_members;
  a.mode = mode;
  a.expflag = p->d.expflag;
  a.diffflag = p->d.diffflag;
  a.diagnostics = p->d.diagnostics;
  a.Pf = (const R*)p->Pf;
  a.Hf = (const R*)p->Hf;
  a.Pi = (const R*)p->Pi;
  a.Hi = (const R*)p->Hi;
  a.fwdP = p->fragP[0];
  a.fwdH = p->fragP[1];
  a.invP = p->fragP[2];
  a.invH = p->fragP[3];
  a.ffwdP = p->fragF[0];
  a.ffwdH = p->fragF[1];
  a.finvP = p->fragF[2];
  a.finvH = p->fragF[3];
  a.leg_fold = p->leg_fold == 2 && p->fragF[0] ? 1 : 0;
  a.bf16x3 = p->d.precision == SWX_PREC_BF16X3 && a.leg_fold ? 1 : 0;
  a.coef = p->coef;
  a.ncoef = p->cl.ncoef;
  a.f2 = p->f2;
  a.S = p->S;
  a.T = (const R*)p->T;
  a.G = (R*)p->G;
  a.part_uv = p->part_uv;
  a.dead = p->dead;
  a.si = p->d.semi_implicit;
  a.raw_filter = p->d.raw_filter;
  a.relax = relax_implicit(p);
  a.resid_c = p->resid_c;
  return a;
}

template <typename R>
GridArgs<R> grid_args(swx_plan* p, int mode) {
  GridArgs<R> a{};
  a.M = p->geo.M;
  a.N = p->geo.N;
  a.J = p->geo.J;
  a.I = p->geo.I;
  a.B = p->d.n_members;
  a.mode = mode;
  a.expflag = p->d.expflag;
  a.forcflag = p->d.forcflag;
  a.test = p->d.test;
  a.diagnostics = p->d.diagnostics;
  a.phys_on = 1;
  a.tw = (const R*)p->tw;
  a.geo = p->geo_d;
  a.coef = p->coef;
  a.ncoef = p->cl.ncoef;
  a.G = (const R*)p->G;
  a.Gout = (R*)p->G;
  a.T = (R*)p->T;
  a.phys = p->phys;
  a.uvic = p->uvic;
  a.part_uv = p->part_uv;
  a.dead = p->dead;
  a.raw = (R*)p->raw;
  a.lag = (R*)p->lag;
  a.g0uv = (R*)p->g0uv;
  a.si = p->d.semi_implicit;
  a.raw_filter = p->d.raw_filter;
  a.relax = relax_implicit(p);
  a.lagx = (R*)p->lagx;
  a.resid_c = p->resid_c;
  a.R0p = p->R0p;
  return a;
}

// The MFMA grid kernels cover the forward stepping of FP32 plans; steps with per-step history or
// diagnostics need the folded FP32 kernel (its FP64 bookkeeping variant writes them); everything
// else uses the plain kernel.
template <typename R>
bool grid_uses_mfma(const swx_plan* p, const GridArgs<R>& g) {
  const bool per_step = g.diagnostics || g.hist_fields || g.hist_phiext;
  const bool fold = p->has_fold && p->grid_fold == 2 && !p->use_bf16x3;
  return sizeof(R) == 4 && p->use_mfma && !p->opt_plain_grid && g.mode == GRID_STEP && !(g.test == 1 && g.expflag) &&
         (!per_step || fold);
}

template <typename R>
bool leg_uses_mfma(const swx_plan* p, const LegArgs<R>& l) {
  return sizeof(R) == 4 && p->use_leg_mfma && !p->opt_plain_leg && l.mode == LEG_STEP;
}

template <typename R>
hipError_t launch_grid_any(swx_plan* p, const GridArgs<R>& g, hipStream_t s) {
  if constexpr (sizeof(R) == 4) {
    if (grid_uses_mfma(p, g)) {
      if (p->use_bf16x3) return launch_grid_bf16x3(g, p->TinvB, p->TfwdB, s);
      if (p->has_fold && p->grid_fold == 2) return launch_grid_mfma_fold(g, p->TinvH, p->TfwdH, s, p->in_pipeline);
      return launch_grid_mfma(g, p->TinvF, p->TfwdF, s);
    }
  }
  return launch_grid<R>(g, s);
}

template <typename R>
hipError_t launch_leg_any(swx_plan* p, const LegArgs<R>& l, hipStream_t s) {
  if constexpr (sizeof(R) == 4) {
    if (leg_uses_mfma(p, l)) {
      if (l.diagnostics) {
        hipError_t e = launch_blowup_gate<R>(l, s);
        if (e != hipSuccess) return e;
      }
      // Beside plans of other truncations the member-group work order (1-2 % faster there);
      // alone or beside plans of its own truncation heaviest first (1-2 % faster for two halves
      // of one T42 batch).
      return launch_leg_mfma(l, p->leg_mb, s, p->in_pipeline && !p->pipe_same_m);
    }
  }
  return launch_leg<R>(l, s);
}

template <typename R>
int do_init(swx_plan* p, const double* eta0, const double* delta0, const double* Phi0, const double* U0,
            const double* V0, hipStream_t s) {
  const int have_uv = (U0 != nullptr);
  const int B = p->d.n_members;
  SWX_HIP(hipMemsetAsync(p->dead, 0, (size_t)B, s));
  // 1) DFT of the initial fields -> raw Fourier (G) and Legendre inputs (T slots 0..2).
  GridArgs<R> g = grid_args<R>(p, GRID_DFT);
  g.in_eta = eta0;
  g.in_delta = delta0;
  g.in_Phi = Phi0;
  g.in_U = U0;
  g.in_V = V0;
  g.have_uv = have_uv;
  SWX_HIP(launch_grid<R>(g, s));
  // 2) Spectral state (and, when diagnosing, the wind Fourier coefficients).
  LegArgs<R> l = leg_args<R>(p, LEG_INIT);
  l.diag_winds = !have_uv;
  SWX_HIP(launch_leg<R>(l, s));
  // 3) Physical carry, forcing, nonlinear products, first Legendre inputs.
  GridArgs<R> gi = grid_args<R>(p, GRID_INIT);
  gi.in_eta = eta0;
  gi.in_delta = delta0;
  gi.in_Phi = Phi0;
  gi.in_U = U0;
  gi.in_V = V0;
  gi.have_uv = have_uv;
  gi.diag_winds = !have_uv;
  SWX_HIP(launch_grid<R>(gi, s));
  const size_t gbytes = (size_t)B * kNG * p->geo.J * (p->geo.M + 1) * 2 * p->rsz;
  SWX_HIP(hipMemcpyAsync(p->Gprev, p->G, gbytes, hipMemcpyDeviceToDevice, s));
  if (p->d.semi_implicit || p->d.raw_filter) {
    // 4) Residue of the initial fields outside the band-limited state: the filtered levels
    //    of these schemes carry it (swx_kernels.hip:k_residue_init).
    LegArgs<R> lp = leg_args<R>(p, LEG_PRIME);
    lp.diagnostics = 0;
    lp.G = (R*)p->R0f;
    SWX_HIP(launch_leg<R>(lp, s));
    SWX_HIP(launch_residue_init<R>((const R*)p->G, (R*)p->R0f, p->R0p, eta0, delta0, Phi0, (const R*)p->tw, p->resid_c,
                                   B, p->geo.J, p->geo.I, p->geo.M + 1, s));
  }
  p->initialized = true;
  p->phys_stale = false;
  return SWX_OK;
}

// Number of final steps whose physical (grid) carry can still reach the
// exported state: the Robert-Asselin recurrence prev' = a prev + (1-2a) curr
// + a new contracts earlier levels by |a| per step, so older contributions
// fall below 2^-64 relative after ceil(64 ln2 / ln(1/|a|)) steps.  Steps
// before the window do not need to write the physical carry.
int phys_window(const swx_plan* p) {
  if (p->opt_full_phys || p->d.diagnostics) return 1 << 30;
  if (p->opt_phys_steps > 0) return p->opt_phys_steps;
  if (p->opt_phys_steps < 0) return 0;
  // Contraction of the prev level per step: alpha (classic RA) or, for RAW,
  // alpha w / (1 - alpha (1 - w)) (swx_physics.h:physical_carry).
  double amax = 0.0;
  for (int b = 0; b < p->d.n_members; ++b) {
    const double* pr = &p->params[(size_t)b * SWX_NPARAM];
    double c = std::fabs(pr[SWX_P_ALPHA]);
    if (p->d.raw_filter) {
      const double al = pr[SWX_P_ALPHA], w = pr[SWX_P_WILLIAMS_ALPHA];
      c = std::fabs(al * w / (1.0 - al * (1.0 - w)));
    }
    amax = std::max(amax, c);
  }
  if (amax >= 0.5) return 1 << 30;
  if (amax == 0.0) return 16;
  const int k = (int)std::ceil(64.0 * std::log(2.0) / std::log(1.0 / amax)) + 2;
  return std::max(16, k);
}

// Modified-Euler plans whose physical prev level follows the classic Robert-Asselin filter
// alone can carry it as a spectral lineage in the (otherwise unused) spectral prev level: the
// Legendre kernel advances it in the final steps of a call (swx_physics.h:spectral_update,
// StepFlags::lineage) and the grid carry is rebuilt from the rows when it is read
// (rebuild_phys), so no step streams the FP64 grid planes.  Test 1 (pinned winds), diagnostics
// (the filtered Phi extrema are per-step grid values) and the other schemes keep the grid carry.
bool lineage_eligible(const swx_plan* p) {
  return !p->d.semi_implicit && !p->d.raw_filter && !p->d.expflag && p->d.test != 1 && !p->d.diagnostics &&
         !p->opt_full_phys && p->opt_phys_steps == 0 && p->opt_lineage >= 0;
}

// Final steps of a call in which the lineage is advanced.  The lineage a call starts from is
// not read: like the grid carry, its memory contracts by |alpha| per filtered step (an
// unfiltered step resets it to the current level), so the window of phys_window steps reaches
// the exported state exactly.  It starts after the call's first step (nsteps > window), whose
// current level is band-limited, so the initial fields' residue outside the band (which the
// grid carry holds after an initialization from grids) has left the lineage's reach too.
int lineage_steps(const swx_plan* p, int nsteps) {
  if (!lineage_eligible(p)) return 0;
  const int w = p->opt_lineage > 0 ? p->opt_lineage : phys_window(p);
  return nsteps > w ? w : 0;
}

// The grid carry of a plan whose last call left it behind (phys_stale): the prev levels from the
// inverse transforms of the spectral lineage, the current levels and winds from the current rows.
template <typename R>
int rebuild_phys_t(swx_plan* p, void* rows, hipStream_t s) {
  const int B = p->d.n_members, J = p->geo.J, I = p->geo.I, Mp = p->geo.M + 1;
  LegArgs<R> l = leg_args<R>(p, LEG_PRIME);
  l.diagnostics = 0;
  l.G = (R*)rows;
  l.prime_prev = 1;
  SWX_HIP(launch_leg<R>(l, s));
  SWX_HIP(launch_phys_synth<R>((const R*)p->G, (const R*)rows, p->phys, (const R*)p->tw, B, J, I, Mp, s));
  p->phys_stale = false;
  return SWX_OK;
}

// Bytes of the lineage rows a rebuild writes: [B][kNG][M+1][J] complex.
size_t lineage_rows_bytes(const swx_plan* p) {
  return (size_t)p->d.n_members * kNG * p->geo.J * (p->geo.M + 1) * 2 * p->rsz;
}

// rows: device buffer of lineage_rows_bytes(p) for the lineage rows, or null for the plan's own
// (allocated on the first rebuild).
int rebuild_phys(swx_plan* p, hipStream_t s, void* rows = nullptr) {
  if (!p->phys_stale) return SWX_OK;
  int rc;
  if (!rows) {
    if (!p->Gsyn && (rc = dalloc(p, &p->Gsyn, lineage_rows_bytes(p)))) return rc;
    rows = p->Gsyn;
  }
  return p->rsz == 8 ? rebuild_phys_t<double>(p, rows, s) : rebuild_phys_t<float>(p, rows, s);
}

template <typename R>
int do_steps(swx_plan* p, int t0, int nsteps, double* hist_fields, double* hist_scal, double* hist_phiext,
             unsigned char* hist_dead, int hist_T, int hist_k0, bool save_last, hipStream_t s, int phys_from = 0,
             int lineage_from = 1 << 30, bool capturing = false) {
  const int B = p->d.n_members;
  const size_t gbytes = (size_t)B * kNG * p->geo.J * (p->geo.M + 1) * 2 * p->rsz;
  for (int k = 0; k < nsteps; ++k) {
    const int t = t0 + k;
    const bool last = save_last && (k == nsteps - 1);
    if (last) SWX_HIP(launch_copy_words(p->Gprev, p->G, gbytes, s));
    LegArgs<R> l = leg_args<R>(p, LEG_STEP);
    l.hist_scal = hist_scal;
    l.hist_dead = hist_dead;
    l.hist_T = hist_T;
    l.hist_k = hist_k0 + k;
    l.do_ra = filter_on(p, t);
    l.lineage = k >= lineage_from ? 1 : 0;
    SWX_HIP(launch_leg_any<R>(p, l, s));
    GridArgs<R> g = grid_args<R>(p, GRID_STEP);
    g.do_ra = filter_on(p, t);
    g.phys_on = (k >= phys_from || hist_fields || hist_phiext) ? 1 : 0;
    g.export_raw = last ? 1 : 0;
    if (!capturing) {
      p->nstep[1] += 1;
      p->nstep[2] += l.lineage;
      p->nstep[3] += g.phys_on;
    }
    g.hist_fields = hist_fields;
    g.hist_phiext = hist_phiext;
    g.hist_T = hist_T;
    g.hist_k = hist_k0 + k;
    SWX_HIP(launch_grid_any<R>(p, g, s));
  }
  return SWX_OK;
}

// Replays the step sequence that `capture` records on stream s, cached under `key`; a sequence
// the cache does not hold is captured and instantiated first.
template <class F>
int graph_replay(GraphCache& gc, const int (&key)[GraphCache::kKey], hipStream_t s, F&& capture) {
  GraphCache::Slot* hit = nullptr;
  GraphCache::Slot* victim = nullptr;
  for (GraphCache::Slot& sl : gc.slot) {
    if (sl.exec && sl.stream == s && std::equal(key, key + GraphCache::kKey, sl.key)) {
      hit = &sl;
      break;
    }
    if (!victim || (!sl.exec && victim->exec) || (sl.exec && victim->exec && sl.used < victim->used)) victim = &sl;
  }
  if (!hit) {
    GraphCache::release(*victim);
    hipGraph_t graph = nullptr;
    SWX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = capture();
    hipError_t ec = hipStreamEndCapture(s, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ec != hipSuccess) return fail(SWX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ec != hipSuccess) return fail(SWX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ec));
    victim->exec = exec;
    std::copy(key, key + GraphCache::kKey, victim->key);
    victim->stream = s;
    hit = victim;
  }
  hit->used = ++gc.clock;
  SWX_HIP(hipGraphLaunch(hit->exec, s));
  if (!hit->done) SWX_HIP(hipEventCreateWithFlags(&hit->done, hipEventDisableTiming));
  SWX_HIP(hipEventRecord(hit->done, s));
  return SWX_OK;
}

int check_plan(swx_plan* p) {
  if (!p) return fail(SWX_ERR_ARG, "null plan");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return fail(SWX_ERR_HIP, "hipGetDevice failed");
  if (dev != p->device) return fail(SWX_ERR_STATE, "plan used on a different device than the one it was created on");
  return SWX_OK;
}

// Copy between the current level of the plan's spectral state ([B][2][3][Mp][Np]
// complex) and a compact [B][3][Mp][Np] complex array.
int copy_state(swx_plan* p, double* dst, const double* src, bool to_plan, bool both_levels, hipStream_t s) {
  const size_t per = (size_t)3 * (p->geo.M + 1) * (p->geo.N + 1) * 2 * sizeof(double);
  const int B = p->d.n_members;
  if (to_plan) {
    for (int lvl = both_levels ? 0 : 1; lvl < 2; ++lvl)
      SWX_HIP(hipMemcpy2DAsync((char*)p->S + lvl * per, 2 * per, src, per, per, B, hipMemcpyDeviceToDevice, s));
  } else {
    SWX_HIP(hipMemcpy2DAsync(dst, per, (const char*)p->S + per, 2 * per, per, B, hipMemcpyDeviceToDevice, s));
  }
  return SWX_OK;
}

// Fourier rows and Legendre inputs of the stored current level: invLeg then
// the grid kernel (GRID_INIT from the rows when `init`, else GRID_STEP
// without the physical carry).
template <typename R>
int prime(swx_plan* p, bool init, hipStream_t s) {
  LegArgs<R> l = leg_args<R>(p, LEG_PRIME);
  l.diagnostics = 0;
  SWX_HIP(launch_leg<R>(l, s));
  GridArgs<R> g = grid_args<R>(p, init ? GRID_INIT : GRID_STEP);
  g.diagnostics = init ? p->d.diagnostics : 0;
  g.init_from_g = init ? 1 : 0;
  g.phys_on = init ? 1 : 0;
  // A restored state needs only the Legendre inputs: the step's grid kernel (MFMA when eligible).
  SWX_HIP(init ? launch_grid<R>(g, s) : launch_grid_any<R>(p, g, s));
  if (init) {
    const size_t gbytes = (size_t)p->d.n_members * kNG * p->geo.J * (p->geo.M + 1) * 2 * p->rsz;
    SWX_HIP(hipMemcpyAsync(p->Gprev, p->G, gbytes, hipMemcpyDeviceToDevice, s));
    if (p->d.semi_implicit || p->d.raw_filter) {
      // A spectral initial state is band-limited: no residue.
      const int B = p->d.n_members;
      SWX_HIP(hipMemsetAsync(p->R0f, 0, gbytes, s));
      SWX_HIP(hipMemsetAsync(p->resid_c, 0, (size_t)B * 2 * 8, s));
      if (p->R0p) SWX_HIP(hipMemsetAsync(p->R0p, 0, (size_t)B * 3 * p->geo.J * p->geo.I * 8, s));
    }
  }
  return SWX_OK;
}

int ensure_adjoint(swx_plan* p) {
  if (p->Lam) return SWX_OK;
  const int B = p->d.n_members, J = p->geo.J, Mp = p->geo.M + 1, Np = p->geo.N + 1;
  int rc;
  if ((rc = dalloc(p, &p->Lam, (size_t)B * 3 * Mp * Np * 2 * 8)) ||
      (rc = dalloc(p, &p->Tbar, (size_t)B * Mp * kNT * J * 2 * p->rsz)) ||
      (rc = dalloc(p, &p->Gbar, (size_t)B * kNG * J * Mp * 2 * p->rsz)) || (rc = dalloc(p, &p->pgrad, (size_t)B * J * 3 * 8)) ||
      (rc = dalloc(p, &p->dgrad, (size_t)B * 2 * 8)))
    return rc;
  const size_t stc = p->use_adj_fold ? grid_adj_mfma_fold_cache_bytes(p->geo.M, B, J) : 0;
  if (stc > 0 && (rc = dalloc(p, &p->adj_stc, stc))) return rc;
  return SWX_OK;
}

template <typename R>
LegAdjArgs<R> leg_adj_args(swx_plan* p, int mode) {
  LegAdjArgs<R> a{};
  a.M = p->geo.M;
  a.N = p->geo.N;
  a.J = p->geo.J;
  a.B = p->d.n_members;
  a.mode = mode;
  a.diffflag = p->d.diffflag;
  a.Pf = (const R*)p->Pf;
  a.Hf = (const R*)p->Hf;
  a.Pi = (const R*)p->Pi;
  a.Hi = (const R*)p->Hi;
  a.coef = p->coef;
  a.ncoef = p->cl.ncoef;
  a.f2 = p->f2;
  a.Lam = p->Lam;
  a.Tbar = (R*)p->Tbar;
  a.G = (R*)p->G;
  a.Gbar = (const R*)p->Gbar;
  a.fwdP = p->fragP[0];
  a.fwdH = p->fragP[1];
  a.invP = p->fragP[2];
  a.invH = p->fragP[3];
  return a;
}

// Legendre adjoint launch: the matrix-core kernels for FP32 plans (unless the plain
// Legendre kernel is forced), else the plain kernel.
template <typename R>
hipError_t launch_leg_adj_any(swx_plan* p, const LegAdjArgs<R>& a, hipStream_t s) {
  if constexpr (sizeof(R) == 4) {
    if (p->use_leg_adj_mfma && !p->opt_plain_leg) return launch_leg_adj_mfma(a, s);
  }
  return launch_leg_adj<R>(a, s);
}

template <typename R>
GridAdjArgs<R> grid_adj_args(swx_plan* p, int mode) {
  GridAdjArgs<R> a{};
  a.M = p->geo.M;
  a.N = p->geo.N;
  a.J = p->geo.J;
  a.I = p->geo.I;
  a.B = p->d.n_members;
  a.mode = mode;
  a.forcflag = p->d.forcflag;
  a.test = p->d.test;
  a.tw = (const R*)p->tw;
  a.geo = p->geo_d;
  a.coef = p->coef;
  a.ncoef = p->cl.ncoef;
  a.G = (const R*)p->G;
  a.Tbar = (const R*)p->Tbar;
  a.Gbar = (R*)p->Gbar;
  a.uvic = p->uvic;
  a.pgrad = p->pgrad;
  a.stc = p->adj_stc;
  return a;
}

// accumulate = 0: Lam = invLeg^T(DFT^T(g)) (terminal); 1: Lam += (observation mid-sweep).
template <typename R>
int adjoint_terminal(swx_plan* p, const double* ge, const double* gd, const double* gp, const double* gu,
                     const double* gv, hipStream_t s, int accumulate = 0) {
  GridAdjArgs<R> g = grid_adj_args<R>(p, GADJ_TERM);
  g.g_eta = ge;
  g.g_delta = gd;
  g.g_Phi = gp;
  g.g_U = gu;
  g.g_V = gv;
  SWX_HIP(launch_grid_adj<R>(g, s));
  SWX_HIP(launch_leg_adj_any<R>(p, leg_adj_args<R>(p, accumulate ? LADJ_B : LADJ_TERM), s));
  return SWX_OK;
}

template <typename R>
int adjoint_step(swx_plan* p, const double* St, hipStream_t s) {
  LegAdjArgs<R> a = leg_adj_args<R>(p, LADJ_A);
  a.St = St;
  SWX_HIP(launch_leg_adj_any<R>(p, a, s));
  const GridAdjArgs<R> g = grid_adj_args<R>(p, GADJ_STEP);
  if constexpr (sizeof(R) == 4) {
    // The grid variant of the forward (SWX_OPT_GRID_FOLD) applies to the reverse step too; at T106
    // only the folded (two-pass) kernel fits the LDS.
    if (!p->opt_plain_grid && p->use_adj_fold && p->grid_fold == 2) {
      SWX_HIP(launch_grid_adj_mfma_fold(g, p->TinvH, p->TfwdH, s));
    } else if (!p->opt_plain_grid && p->use_adj_mfma) {
      SWX_HIP(launch_grid_adj_mfma(g, p->TinvF, p->TfwdF, s));
    } else {
      SWX_HIP(launch_grid_adj<R>(g, s));
    }
  } else {
    SWX_HIP(launch_grid_adj<R>(g, s));
  }
  SWX_HIP(launch_leg_adj_any<R>(p, leg_adj_args<R>(p, LADJ_B), s));
  return SWX_OK;
}

}  // namespace

// Legendre tile autotuner: time each instantiated member-group size on this plan's own
// buffers (zeroed Legendre inputs and state, 2 warm-up + 5 timed launches on the null
// stream) and keep the fastest.  The state is zeroed again afterwards.
int autotune_leg(swx_plan* p) {
  int mbs[3];
  const int nc = leg_mfma_candidates(p->geo.M, mbs);
  int maxlds = 0;
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device);
  const int B = p->d.n_members, J = p->geo.J, Mp = p->geo.M + 1, Np = p->geo.N + 1;
  const size_t sbytes = (size_t)B * 2 * 3 * Mp * Np * 2 * 8;
  SWX_HIP(hipMemset(p->T, 0, (size_t)B * Mp * kNT * J * 2 * p->rsz));
  hipEvent_t e0, e1;
  SWX_HIP(hipEventCreate(&e0));
  SWX_HIP(hipEventCreate(&e1));
  const LegArgs<float> l = leg_args<float>(p, LEG_STEP);
  float best = 0.f;
  for (int c = 0; c < nc; ++c) {
    p->leg_mb_ms[c] = 0.f;
    if (leg_mfma_lds_bytes(p->geo.M, mbs[c]) > (size_t)maxlds) continue;
    for (int w = 0; w < 2; ++w) SWX_HIP(launch_leg_mfma(l, mbs[c], nullptr));
    SWX_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < 5; ++r) SWX_HIP(launch_leg_mfma(l, mbs[c], nullptr));
    SWX_HIP(hipEventRecord(e1, nullptr));
    SWX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SWX_HIP(hipEventElapsedTime(&ms, e0, e1));
    p->leg_mb_ms[c] = ms / 5;
    if (best == 0.f || ms / 5 < best) {
      best = ms / 5;
      p->leg_mb = mbs[c];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SWX_HIP(hipMemset(p->S, 0, sbytes));
  return SWX_OK;
}

// Grid kernel variant: the half-period folded MFMA grid kernel wherever its LDS fits (it does half
// the matrix-core work: 34.5 against 74.7 us at T42 x 256), else the full-DFT one.  A fixed rule,
// not a timing, so that two plans of the same shape always round alike (the two variants agree
// only to FP32 rounding).
int choose_grid_variant(swx_plan* p) {
  p->grid_fold = p->has_fold ? 2 : 1;
  return SWX_OK;
}

// Two FP32 plans of different truncations stepped together: each step is one Legendre launch
// and one grid launch covering both plans' workgroups (MFMA pair kernels); a step where either
// plan needs a plain kernel falls back to one launch per plan.
// Plans stepped concurrently: plan 0 on the caller's stream, plan k on internal stream k.
constexpr int kMaxPipeline = 4;
struct swx_pipeline {
  int n = 0;
  swx_plan* p[kMaxPipeline] = {};
  hipStream_t s[kMaxPipeline] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxPipeline] = {};
};

struct swx_ensemble {
  swx_plan* p[2] = {nullptr, nullptr};  // p[0]->geo.M < p[1]->geo.M
  bool fuse_grid = true;                // one grid launch for both plans (timed at creation)
  float grid_ms[2] = {0, 0};            // fused, separate
  GraphCache graphs;                    // captured step sequences (swx_ensemble_run_graph)
  unsigned epoch[2] = {0, 0};           // the plans' opt_epoch when they were captured
};

namespace {

int ens_steps(swx_ensemble* e, int t0, int nsteps, bool save_last, hipStream_t s, const int* phys_from,
              const int* lineage_from, bool capturing = false) {
  for (int k = 0; k < nsteps; ++k) {
    const int t = t0 + k;
    const bool last = save_last && (k == nsteps - 1);
    LegArgs<float> l[2];
    GridArgs<float> g[2];
    for (int q = 0; q < 2; ++q) {
      swx_plan* p = e->p[q];
      if (last) {
        const size_t gbytes = (size_t)p->d.n_members * kNG * p->geo.J * (p->geo.M + 1) * 2 * p->rsz;
        SWX_HIP(launch_copy_words(p->Gprev, p->G, gbytes, s));
      }
      l[q] = leg_args<float>(p, LEG_STEP);
      l[q].do_ra = filter_on(p, t);
      l[q].lineage = k >= lineage_from[q] ? 1 : 0;
      g[q] = grid_args<float>(p, GRID_STEP);
      g[q].do_ra = filter_on(p, t);
      g[q].phys_on = k >= phys_from[q] ? 1 : 0;
      if (!capturing) {
        p->nstep[1] += 1;
        p->nstep[2] += l[q].lineage;
        p->nstep[3] += g[q].phys_on;
      }
      g[q].export_raw = last ? 1 : 0;
    }
    if (leg_uses_mfma(e->p[0], l[0]) && leg_uses_mfma(e->p[1], l[1])) {
      SWX_HIP(launch_leg_mfma_pair(l[0], l[1], s));
    } else {
      for (int q = 0; q < 2; ++q) SWX_HIP(launch_leg_any<float>(e->p[q], l[q], s));
    }
    if (e->fuse_grid && grid_uses_mfma(e->p[0], g[0]) && grid_uses_mfma(e->p[1], g[1])) {
      const swx_plan* q0 = e->p[0];
      const swx_plan* q1 = e->p[1];
      const bool f0 = q0->has_fold && q0->grid_fold == 2, f1 = q1->has_fold && q1->grid_fold == 2;
      SWX_HIP(launch_grid_mfma_pair(g[0], f0 ? q0->TinvH : q0->TinvF, f0 ? q0->TfwdH : q0->TfwdF, f0, g[1],
                                    f1 ? q1->TinvH : q1->TinvF, f1 ? q1->TfwdH : q1->TfwdF, f1, s));
    } else {
      for (int q = 0; q < 2; ++q) SWX_HIP(launch_grid_any<float>(e->p[q], g[q], s));
    }
  }
  return SWX_OK;
}

// Parameter checks the kernels rely on: the RAW physical carry recovers the filter
// increment from the filtered level, which needs alpha (1 - williams_alpha) != 1.
int check_params(const swx_plan_desc* d, const double* params) {
  if (!d->raw_filter) return SWX_OK;
  for (int b = 0; b < d->n_members; ++b) {
    const double al = params[(size_t)b * SWX_NPARAM + SWX_P_ALPHA];
    const double w = params[(size_t)b * SWX_NPARAM + SWX_P_WILLIAMS_ALPHA];
    if (!(std::fabs(1.0 - al * (1.0 - w)) > 1e-12))
      return fail(SWX_ERR_ARG, "raw_filter: alpha * (1 - williams_alpha) must differ from 1");
  }
  return SWX_OK;
}

// Grid step of both plans as one fused launch or as each plan's own launch: the same kernel
// bodies (identical results); the faster is kept.  The pair launch sizes its LDS for the
// larger of the two bodies, which can halve the occupancy of the smaller one.  Timed at
// ensemble creation (2 warm-up + 5 timed rounds, null stream) with the physical carry,
// filters and lag off, so the launches write only the Legendre inputs T, which are saved
// before and restored after.
int tune_ensemble(swx_ensemble* e) {
  GridArgs<float> g[2];
  void* saved[2] = {nullptr, nullptr};
  size_t tb[2];
  int rc = SWX_OK;
  for (int q = 0; q < 2 && rc == SWX_OK; ++q) {
    swx_plan* p = e->p[q];
    g[q] = grid_args<float>(p, GRID_STEP);
    g[q].phys_on = 0;
    g[q].do_ra = 0;
    g[q].si = 0;
    g[q].raw_filter = 0;
    tb[q] = (size_t)p->d.n_members * (p->geo.M + 1) * kNT * p->geo.J * 2 * p->rsz;
    if (hipMalloc(&saved[q], tb[q]) != hipSuccess ||
        hipMemcpy(saved[q], p->T, tb[q], hipMemcpyDeviceToDevice) != hipSuccess)
      rc = fail(SWX_ERR_HIP, "ensemble tuning: saving the Legendre inputs failed");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == SWX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(SWX_ERR_HIP, "hipEventCreate failed");
  for (int v = 0; v < 2 && rc == SWX_OK; ++v) {
    auto round = [&]() -> hipError_t {
      if (v == 0) {
        const swx_plan* q0 = e->p[0];
        const swx_plan* q1 = e->p[1];
        const bool f0 = q0->has_fold && q0->grid_fold == 2, f1 = q1->has_fold && q1->grid_fold == 2;
        return launch_grid_mfma_pair(g[0], f0 ? q0->TinvH : q0->TinvF, f0 ? q0->TfwdH : q0->TfwdF, f0, g[1],
                                     f1 ? q1->TinvH : q1->TinvF, f1 ? q1->TfwdH : q1->TfwdF, f1, nullptr);
      }
      hipError_t er = launch_grid_any<float>(e->p[0], g[0], nullptr);
      return er == hipSuccess ? launch_grid_any<float>(e->p[1], g[1], nullptr) : er;
    };
    hipError_t er = hipSuccess;
    for (int w = 0; w < 2 && er == hipSuccess; ++w) er = round();
    if (er == hipSuccess) er = hipEventRecord(e0, nullptr);
    for (int r = 0; r < 5 && er == hipSuccess; ++r) er = round();
    if (er == hipSuccess) er = hipEventRecord(e1, nullptr);
    if (er == hipSuccess) er = hipEventSynchronize(e1);
    float ms = 0.f;
    if (er == hipSuccess) er = hipEventElapsedTime(&ms, e0, e1);
    if (er != hipSuccess) rc = fail(SWX_ERR_HIP, std::string("ensemble tuning: ") + hipGetErrorString(er));
    e->grid_ms[v] = ms / 5;
  }
  if (rc == SWX_OK) e->fuse_grid = e->grid_ms[0] <= e->grid_ms[1];
  for (int q = 0; q < 2; ++q) {
    if (saved[q]) {
      if (rc == SWX_OK && hipMemcpy(e->p[q]->T, saved[q], tb[q], hipMemcpyDeviceToDevice) != hipSuccess)
        rc = fail(SWX_ERR_HIP, "ensemble tuning: restoring the Legendre inputs failed");
      (void)hipFree(saved[q]);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

extern "C" {

const char* swx_last_error(void) { return g_err.c_str(); }

int swx_measure_ceilings(double* out, size_t bytes, void* stream) {
  if (!out) return fail(SWX_ERR_ARG, "null output");
  if (bytes < ((size_t)256 << 20) || bytes % 16 != 0)
    return fail(SWX_ERR_ARG, "swx_measure_ceilings: bytes must be a multiple of 16 and at least 256 MiB");
  const hipError_t e = measure_ceilings(out, bytes, (hipStream_t)stream);
  if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("swx_measure_ceilings: ") + hipGetErrorString(e));
  return SWX_OK;
}

const char* swx_version(void) { return "swampe-hip 0.1.0 (gfx950)"; }

int swx_grid_size(int M, int* J, int* I) {
  if (!resolution(M, J, I, nullptr))
    return fail(SWX_ERR_ARG, "Unsupported value of M=" + std::to_string(M) + ". Only 42, 63, and 106 are supported.");
  return SWX_OK;
}

int swx_gauss_legendre(int J, double* mus, double* w) {
  if (J < 1) return fail(SWX_ERR_ARG, "gauss_legendre: J must be >= 1");
  if (!mus || !w) return fail(SWX_ERR_ARG, "gauss_legendre: null output");
  gauss_legendre(J, mus, w);
  return SWX_OK;
}

int swx_lambdas(int I, double* lambdas_out) {
  if (I < 1) return fail(SWX_ERR_ARG, "build_lambdas: I must be >= 1");
  if (!lambdas_out) return fail(SWX_ERR_ARG, "build_lambdas: null output");
  lambdas(I, lambdas_out);
  return SWX_OK;
}

int swx_legendre_basis(int J, int M, int N, const double* mus, int exact_basis, double* Pmn, double* Hmn) {
  if (J < 1 || M < 0 || N < 0) return fail(SWX_ERR_ARG, "PmnHmn: need J >= 1, M >= 0, N >= 0");
  if (!mus || !Pmn || !Hmn) return fail(SWX_ERR_ARG, "PmnHmn: null argument");
  for (int j = 0; j < J; ++j)
    if (!(std::fabs(mus[j]) <= 1.0))
      return fail(SWX_ERR_ARG, "PmnHmn: mus[" + std::to_string(j) + "] outside [-1, 1]");
  legendre_basis(M, N, J, mus, Pmn, Hmn, exact_basis == 0);
  return SWX_OK;
}

int swx_geometry(int M, int exact_basis, double* mus, double* w, double* lambdas_out, double* Pmn, double* Hmn) {
  Geometry g;
  if (!build_geometry(M, exact_basis != 0, g))
    return fail(SWX_ERR_ARG, "Unsupported value of M=" + std::to_string(M) + ". Only 42, 63, and 106 are supported.");
  if (mus) std::memcpy(mus, g.mus.data(), g.J * sizeof(double));
  if (w) std::memcpy(w, g.w.data(), g.J * sizeof(double));
  if (lambdas_out) std::memcpy(lambdas_out, g.lam.data(), g.I * sizeof(double));
  if (Pmn) std::memcpy(Pmn, g.P.data(), g.P.size() * sizeof(double));
  if (Hmn) std::memcpy(Hmn, g.H.data(), g.H.size() * sizeof(double));
  return SWX_OK;
}

int swx_plan_create(const swx_plan_desc* desc, const double* params, swx_plan** out) {
  if (!desc || !params || !out) return fail(SWX_ERR_ARG, "null argument");
  int rc0;
  *out = nullptr;
  if (desc->n_members < 1) return fail(SWX_ERR_ARG, "n_members must be >= 1");
  if (desc->precision != SWX_PREC_FP64 && desc->precision != SWX_PREC_FP32 && desc->precision != SWX_PREC_BF16X3)
    return fail(SWX_ERR_ARG, "unsupported precision " + std::to_string(desc->precision));
  if (!(desc->test == -1 || desc->test == 1 || desc->test == 2))
    return fail(SWX_ERR_ARG, "test must be -1 (None), 1 or 2");
  // model.py:1531-1537
  if (desc->semi_implicit && desc->expflag)
    return fail(SWX_ERR_ARG, "semi_implicit=True is incompatible with expflag=True (pick one scheme).");
  if (desc->raw_filter && desc->expflag)
    return fail(SWX_ERR_ARG,
                "raw_filter=True is incompatible with expflag=True (the explicit scheme reads the unfiltered "
                "spectral prev carry, so the RAW lineage change would alter its parity behavior).");
  if ((rc0 = check_params(desc, params))) return rc0;
  if ((desc->quad_mus == nullptr) != (desc->quad_w == nullptr))
    return fail(SWX_ERR_ARG, "quad_mus and quad_w must be given together");
  if (desc->quad_mus) {
    int Jq = 0;
    if (!resolution(desc->M, &Jq, nullptr, nullptr))
      return fail(SWX_ERR_ARG, "Unsupported value of M=" + std::to_string(desc->M) + ". Only 42, 63, and 106 are supported.");
    double wsum = 0.0;
    for (int j = 0; j < Jq; ++j) {
      const double x = desc->quad_mus[j], wj = desc->quad_w[j];
      if (!(x > -1.0 && x < 1.0) || !(wj > 0.0) || (j > 0 && !(x > desc->quad_mus[j - 1])))
        return fail(SWX_ERR_ARG, "quad_mus must ascend inside (-1, 1) and quad_w be positive");
      wsum += wj;
    }
    if (!(std::fabs(wsum - 2.0) < 1e-10)) return fail(SWX_ERR_ARG, "quad_w must sum to 2");
  }
  swx_plan* p = new swx_plan();
  p->d = *desc;
  p->d.quad_mus = nullptr;  // the plan keeps its own copy (geo)
  p->d.quad_w = nullptr;
  if (!build_geometry(desc->M, desc->exact_basis != 0, p->geo, desc->quad_mus, desc->quad_w)) {
    delete p;
    return fail(SWX_ERR_ARG, "Unsupported value of M=" + std::to_string(desc->M) + ". Only 42, 63, and 106 are supported.");
  }
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(SWX_ERR_HIP, "hipGetDevice failed");
  }
  const int M = p->geo.M, N = p->geo.N, J = p->geo.J, I = p->geo.I, B = desc->n_members;
  const int Mp = M + 1;
  p->rsz = desc->precision == SWX_PREC_FP64 ? 8 : 4;
  p->cl = CoefLayout::make(N, J);
  p->gl = GeoLayout::make(J, I);
  p->params.assign(params, params + (size_t)B * SWX_NPARAM);

  // Triangular basis blocks in both layouts, w folded into the Legendre inputs.
  const long long tri = tri_off(M + 1, N, J);
  std::vector<double> Pf(tri), Hf(tri), Pi(tri), Hi(tri);
  for (int m = 0; m <= M; ++m) {
    const long long off = tri_off(m, N, J);
    const int K = N + 1 - m;
    for (int j = 0; j < J; ++j)
      for (int np = 0; np < K; ++np) {
        const size_t src = ((size_t)j * Mp + m) * (N + 1) + (m + np);
        Pf[off + (long long)j * K + np] = p->geo.P[src];
        Hf[off + (long long)j * K + np] = p->geo.H[src];
        Pi[off + (long long)np * J + j] = p->geo.P[src];
        Hi[off + (long long)np * J + j] = p->geo.H[src];
      }
  }
  std::vector<double> tw(2 * I);
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < I; ++k) {
    tw[k] = std::cos(2.0 * pi * k / I);
    tw[I + k] = std::sin(2.0 * pi * k / I);
  }
  std::vector<double> geo(p->gl.n);
  for (int j = 0; j < J; ++j) {
    const double mu = p->geo.mus[j];
    geo[p->gl.mus + j] = mu;
    geo[p->gl.w + j] = p->geo.w[j];
    geo[p->gl.s1m + j] = std::sqrt(1 - mu * mu);
    geo[p->gl.cosphi + j] = std::cos(std::asin(mu));
    geo[p->gl.denomE + j] = 2.0 * (1.0 - mu * mu);
  }
  for (int i = 0; i < I; ++i) {
    const double l = p->geo.lam[i];
    geo[p->gl.cosl + i] = std::cos(l);
    geo[p->gl.day + i] = (l > -pi / 2 && l < pi / 2) ? 1.0 : 0.0;
    geo[p->gl.dcos + i] = geo[p->gl.day + i] * geo[p->gl.cosl + i];
  }
  std::vector<double> coef;
  fill_coef(p, coef);
  std::vector<double> f2((size_t)Mp * (N + 1), 1.0);
  if (!desc->exact_basis)
    for (int m = 0; m <= M; ++m)
      for (int n = m; n <= N; ++n) {
        const double f = reference_scale_factor(m, n);
        f2[(size_t)m * (N + 1) + n] = f * f;
      }

  int rc = SWX_OK;
  const size_t rs = p->rsz;
  const size_t cplxJM = (size_t)J * Mp * 2;
  if ((rc = dalloc(p, &p->Pf, tri * rs)) || (rc = dalloc(p, &p->Hf, tri * rs)) || (rc = dalloc(p, &p->Pi, tri * rs)) ||
      (rc = dalloc(p, &p->Hi, tri * rs)) || (rc = dalloc(p, &p->tw, 2 * I * rs)) ||
      (rc = dalloc(p, &p->geo_d, geo.size() * 8)) || (rc = dalloc(p, &p->coef, coef.size() * 8)) ||
      (rc = dalloc(p, &p->f2, f2.size() * 8)) ||
      (rc = dalloc(p, &p->S, (size_t)B * 2 * 3 * Mp * (N + 1) * 2 * 8)) ||
      (rc = dalloc(p, &p->T, (size_t)B * Mp * kNT * J * 2 * rs)) || (rc = dalloc(p, &p->G, (size_t)B * kNG * cplxJM * rs)) ||
      (rc = dalloc(p, &p->Gprev, (size_t)B * kNG * cplxJM * rs)) || (rc = dalloc(p, &p->raw, (size_t)B * kNRaw * cplxJM * rs)) ||
      (rc = dalloc(p, &p->lag, (size_t)B * 2 * cplxJM * rs)) || (rc = dalloc(p, &p->g0uv, (size_t)B * 2 * cplxJM * rs)) ||
      (rc = dalloc(p, &p->phys, (size_t)B * 8 * J * I * 8)) || (rc = dalloc(p, &p->uvic, (size_t)B * 2 * J * I * 8)) ||
      (rc = dalloc(p, &p->part_uv, (size_t)B * J * 2 * 8)) || (rc = dalloc(p, &p->dead, (size_t)B))) {
    swx_plan_destroy(p);
    return rc;
  }
  if ((desc->semi_implicit || desc->raw_filter) &&
      ((rc = dalloc(p, &p->R0f, (size_t)B * kNG * cplxJM * rs)) || (rc = dalloc(p, &p->resid_c, (size_t)B * 2 * 8)) ||
       (desc->semi_implicit && (rc = dalloc(p, &p->lagx, (size_t)B * 2 * cplxJM * rs))) ||
       (desc->raw_filter && (rc = dalloc(p, &p->R0p, (size_t)B * 3 * J * I * 8))))) {
    swx_plan_destroy(p);
    return rc;
  }
  if (desc->precision == SWX_PREC_FP32 || desc->precision == SWX_PREC_BF16X3) {
    int maxlds = 0;
    (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device);
    p->use_mfma = grid_mfma_lds_bytes(M) <= (size_t)maxlds;
    const size_t adj_lds = grid_adj_mfma_lds_bytes(M);
    p->use_adj_mfma = p->use_mfma && adj_lds > 0 && adj_lds <= (size_t)maxlds;
    p->leg_mb = leg_mfma_default_mb(M);
    p->use_leg_mfma = leg_mfma_lds_bytes(M, p->leg_mb) <= (size_t)maxlds;
    const size_t ladj_lds = leg_adj_mfma_lds_bytes(M);
    p->use_leg_adj_mfma = p->use_leg_mfma && ladj_lds > 0 && ladj_lds <= (size_t)maxlds;
    if (p->use_leg_mfma) {
      // Fragment-ordered copies of the basis blocks (layouts in LegArgs).
      const int KNP = leg_knp(N);
      const size_t per = (size_t)Mp * KNP * J;
      std::vector<std::vector<double>> fr(4, std::vector<double>(per, 0.0));
      for (int m = 0; m <= M; ++m) {
        const int K = N + 1 - m;
        for (int lane = 0; lane < 64; ++lane) {
          const int ln = lane & 15, lg = lane >> 4;
          for (int nt = 0; nt < KNP / 16; ++nt)
            for (int ks = 0; ks < J / 4; ++ks) {
              const int np = 16 * nt + ln, j = 4 * ks + lg;
              if (np >= K) continue;
              const size_t src = ((size_t)j * Mp + m) * (N + 1) + (m + np);
              const size_t dst = (((size_t)m * (KNP / 16) + nt) * 64 + lane) * (J / 4) + ks;
              fr[0][dst] = p->geo.P[src];
              fr[1][dst] = p->geo.H[src];
            }
          for (int jt = 0; jt < J / 16; ++jt)
            for (int ks = 0; ks < KNP / 4; ++ks) {
              const int np = 4 * ks + lg, j = 16 * jt + ln;
              if (np >= K) continue;
              const size_t src = ((size_t)j * Mp + m) * (N + 1) + (m + np);
              const size_t dst = (((size_t)m * (J / 16) + jt) * 64 + lane) * (KNP / 4) + ks;
              fr[2][dst] = p->geo.P[src];
              fr[3][dst] = p->geo.H[src];
            }
        }
      }
      for (int q = 0; q < 4 && rc == SWX_OK; ++q)
        if ((rc = dalloc(p, &p->fragP[q], per * 4)) == SWX_OK) rc = upload_real(p->fragP[q], fr[q], 4);
      // Equatorially folded tables (LegArgs ffwd*, finv*): the basis on the southern nodes j < J/2,
      // rows by degree parity.  They rely on mu_{J-1-j} = -mu_j and the basis parity
      // (checked here to 1e-12 of each block's largest value; otherwise the plan keeps the full
      // contractions).
      {
        const int J2 = J / 2, NTF = leg_fold_ntf(N), KSE = leg_fold_kse(N), KSO = leg_fold_kso(N);
        const int KSL = KSE + KSO;
        bool sym = J % 32 == 0;
        for (int j = 0; j < J2 && sym; ++j)
          sym = std::fabs(p->geo.mus[j] + p->geo.mus[J - 1 - j]) <= 1e-14;
        for (int m = 0; m <= M && sym; ++m)
          for (int n = m; n <= N && sym; ++n) {
            double amax = 0.0, err = 0.0;
            const double sp = ((n - m) & 1) ? -1.0 : 1.0;
            for (int j = 0; j < J; ++j) {
              const size_t a0 = ((size_t)j * Mp + m) * (N + 1) + n, a1 = ((size_t)(J - 1 - j) * Mp + m) * (N + 1) + n;
              amax = std::max(amax, std::max(std::fabs(p->geo.P[a0]), std::fabs(p->geo.H[a0])));
              err = std::max(err, std::fabs(p->geo.P[a1] - sp * p->geo.P[a0]));
              err = std::max(err, std::fabs(p->geo.H[a1] + sp * p->geo.H[a0]));
            }
            sym = err <= 1e-12 * amax;
          }
        if (sym) {
          const size_t perf = (size_t)Mp * NTF * 64 * (J / 8), peri = (size_t)Mp * (J / 32) * 64 * KSL;
          std::vector<std::vector<double>> ff(4);
          ff[0].assign(perf, 0.0);
          ff[1].assign(perf, 0.0);
          ff[2].assign(peri, 0.0);
          ff[3].assign(peri, 0.0);
          for (int m = 0; m <= M; ++m) {
            const int K = N + 1 - m, Ke = (K + 1) / 2, Ko = K / 2;
            const int nte = leg_fold_nte(K), kb = leg_fold_kb(K);
            for (int lane = 0; lane < 64; ++lane) {
              const int ln = lane & 15, lg = lane >> 4;
              for (int r = 0; r < NTF; ++r)
                for (int ks = 0; ks < J / 8; ++ks) {
                  const int i = 16 * (r < nte ? r : r - nte) + ln;
                  const int np = r < nte ? 2 * i : 2 * i + 1;
                  if (r < nte ? i >= Ke : i >= Ko) continue;
                  const int j = 4 * ks + lg;
                  const size_t src = ((size_t)j * Mp + m) * (N + 1) + (m + np);
                  const size_t dst = (((size_t)m * NTF + r) * 64 + lane) * (J / 8) + ks;
                  ff[0][dst] = p->geo.P[src];
                  ff[1][dst] = p->geo.H[src];
                }
              for (int jt = 0; jt < J2 / 16; ++jt)
                for (int slot = 0; slot < KSL; ++slot) {
                  int np = -1;
                  if (slot < KSE) {
                    const int row = 4 * slot + lg;
                    if (row < Ke) np = 2 * row;
                  } else {
                    const int row = 4 * (kb + slot - KSE) + lg;
                    if (row >= Ke && row < K) np = 2 * (row - Ke) + 1;
                  }
                  if (np < 0) continue;
                  const int j = 16 * jt + ln;
                  const size_t src = ((size_t)j * Mp + m) * (N + 1) + (m + np);
                  const size_t dst = (((size_t)m * (J2 / 16) + jt) * 64 + lane) * KSL + slot;
                  ff[2][dst] = p->geo.P[src];
                  ff[3][dst] = p->geo.H[src];
                }
            }
            (void)Ko;
          }
          for (int q = 0; q < 4 && rc == SWX_OK; ++q)
            if ((rc = dalloc(p, &p->fragF[q], ff[q].size() * 4)) == SWX_OK) rc = upload_real(p->fragF[q], ff[q], 4);
        }
      }
      if (rc) {
        swx_plan_destroy(p);
        return rc;
      }
    }
    if (p->use_mfma) {
      const int k1 = 2 * Mp, k2p = grid_mfma_k2p(M), kq1 = grid_mfma_kq1(M);
      auto tinv = [&](int i, int k) -> double {
        if (k >= k1) return 0.0;
        const int m = k / 2;
        const double ph = 2.0 * pi * (double)(((long long)m * i) % I) / I, wm = m == 0 ? 1.0 : 2.0;
        return (k & 1) ? -wm * std::sin(ph) : wm * std::cos(ph);
      };
      auto tfwd = [&](int kp, int i) -> double {
        if (kp >= k1) return 0.0;
        const int m = kp / 2;
        const double ph = 2.0 * pi * (double)(((long long)m * i) % I) / I;
        return (kp & 1) ? -std::sin(ph) / I : std::cos(ph) / I;
      };
      std::vector<double> ti((size_t)(I / 16) * 64 * kq1), tf((size_t)k2p * I);
      for (int lane = 0; lane < 64; ++lane) {
        const int ln = lane & 15, lg = lane >> 4;
        for (int it = 0; it < I / 16; ++it)
          for (int ks = 0; ks < kq1; ++ks) ti[((size_t)it * 64 + lane) * kq1 + ks] = tinv(16 * it + ln, 4 * ks + lg);
        for (int t = 0; t < k2p / 16; ++t)
          for (int g = 0; g < I / 16; ++g)
            for (int q = 0; q < 4; ++q)
              tf[(((size_t)t * (I / 16) + g) * 64 + lane) * 4 + q] = tfwd(16 * t + ln, 16 * g + 4 * q + lg);
      }
      if ((rc = dalloc(p, &p->TinvF, ti.size() * 4)) || (rc = dalloc(p, &p->TfwdF, tf.size() * 4)) ||
          (rc = upload_real(p->TinvF, ti, 4)) || (rc = upload_real(p->TfwdF, tf, 4))) {
        swx_plan_destroy(p);
        return rc;
      }
      FoldDims fd{};
      if (grid_mfma_fold_dims(M, &fd) && grid_mfma_fold_lds_bytes(M) <= (size_t)maxlds) {
        // Folded order of the k rows and of the output tile rows (swx_kernels.h FoldDims).
        auto kfold = [&](int kp) -> int {  // inverse row -> 2 m + c, or -1
          if (kp >= fd.K1) return -1;
          const int sl = kp / 2, c = kp & 1;
          const int m = sl < fd.KE4 / 2 ? 2 * sl : 2 * (sl - fd.KE4 / 2) + 1;
          return m < Mp ? 2 * m + c : -1;
        };
        auto ofold = [&](int t, int r) -> int {  // output tile row -> 2 m + c, or -1
          const int rr = 16 * (t < fd.TE ? t : t - fd.TE) + r;
          const int m = t < fd.TE ? 2 * (rr / 2) : 2 * (rr / 2) + 1;
          return m < Mp ? 2 * m + (rr & 1) : -1;
        };
        std::vector<double> hi((size_t)(fd.H / 16) * 64 * fd.KQ1, 0.0), hf((size_t)fd.NT2 * (fd.H / 16) * 64 * 4, 0.0);
        for (int lane = 0; lane < 64; ++lane) {
          const int ln = lane & 15, lg = lane >> 4;
          for (int it = 0; it < fd.H / 16; ++it)
            for (int ks = 0; ks < fd.KQ1; ++ks) {
              const int k = kfold(4 * ks + lg);
              if (k >= 0) hi[((size_t)it * 64 + lane) * fd.KQ1 + ks] = tinv(16 * it + ln, k);
            }
          for (int t = 0; t < fd.NT2; ++t) {
            const int k = ofold(t, ln);
            if (k < 0) continue;
            for (int g = 0; g < fd.H / 16; ++g)
              for (int q = 0; q < 4; ++q)
                hf[(((size_t)t * (fd.H / 16) + g) * 64 + lane) * 4 + q] = tfwd(k, 16 * g + 4 * q + lg);
          }
        }
        if ((rc = dalloc(p, &p->TinvH, hi.size() * 4)) || (rc = dalloc(p, &p->TfwdH, hf.size() * 4)) ||
            (rc = upload_real(p->TinvH, hi, 4)) || (rc = upload_real(p->TfwdH, hf, 4))) {
          swx_plan_destroy(p);
          return rc;
        }
        p->has_fold = true;
        const size_t adjf = grid_adj_mfma_fold_lds_bytes(M);
        p->use_adj_fold = adjf > 0 && adjf <= (size_t)maxlds;
      }
      if (desc->precision == SWX_PREC_BF16X3 && grid_bf16x3_lds_bytes(M) <= (size_t)maxlds) {
        // Fragment order of the split-BF16x3 kernel: per (tile, k-step, lane) 8 hi then 8 lo
        // values, k (or i) = 32 ks + 8 (lane / 16) + e.
        const int ks1 = grid_bf16x3_ks1(M);
        std::vector<uint16_t> bi((size_t)(I / 16) * ks1 * 64 * 16), bf((size_t)(k2p / 16) * (I / 32) * 64 * 16);
        for (int lane = 0; lane < 64; ++lane) {
          const int ln = lane & 15, lg = lane >> 4;
          for (int e = 0; e < 8; ++e) {
            for (int it = 0; it < I / 16; ++it)
              for (int ks = 0; ks < ks1; ++ks) {
                const size_t o = (((size_t)it * ks1 + ks) * 64 + lane) * 16;
                bf16_split(tinv(16 * it + ln, 32 * ks + 8 * lg + e), bi[o + e], bi[o + 8 + e]);
              }
            for (int t = 0; t < k2p / 16; ++t)
              for (int g = 0; g < I / 32; ++g) {
                const size_t o = (((size_t)t * (I / 32) + g) * 64 + lane) * 16;
                bf16_split(tfwd(16 * t + ln, 32 * g + 8 * lg + e), bf[o + e], bf[o + 8 + e]);
              }
          }
        }
        if ((rc = dalloc(p, &p->TinvB, bi.size() * 2)) || (rc = dalloc(p, &p->TfwdB, bf.size() * 2))) {
          swx_plan_destroy(p);
          return rc;
        }
        if (hipMemcpy(p->TinvB, bi.data(), bi.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->TfwdB, bf.data(), bf.size() * 2, hipMemcpyHostToDevice) != hipSuccess) {
          swx_plan_destroy(p);
          return fail(SWX_ERR_HIP, "hipMemcpy of the BF16x3 tables failed");
        }
        p->use_bf16x3 = true;
      }
    }
  }
  if ((rc = upload_real(p->Pf, Pf, rs)) || (rc = upload_real(p->Hf, Hf, rs)) || (rc = upload_real(p->Pi, Pi, rs)) ||
      (rc = upload_real(p->Hi, Hi, rs)) || (rc = upload_real(p->tw, tw, rs)) ||
      (rc = upload_real(p->geo_d, geo, 8)) || (rc = upload_real(p->coef, coef, 8)) ||
      (rc = upload_real(p->f2, f2, 8))) {
    swx_plan_destroy(p);
    return rc;
  }
  if (hipMemset(p->dead, 0, (size_t)B) != hipSuccess || hipMemset(p->part_uv, 0, (size_t)B * J * 2 * 8) != hipSuccess ||
      hipMemset(p->S, 0, (size_t)B * 2 * 3 * Mp * (N + 1) * 2 * 8) != hipSuccess) {
    swx_plan_destroy(p);
    return fail(SWX_ERR_HIP, "hipMemset failed");
  }
  if (p->use_mfma && (rc = choose_grid_variant(p)) != SWX_OK) {
    swx_plan_destroy(p);
    return rc;
  }
  if (p->use_leg_mfma && (rc = autotune_leg(p)) != SWX_OK) {
    swx_plan_destroy(p);
    return rc;
  }
  *out = p;
  return SWX_OK;
}

int swx_plan_set_params(swx_plan* p, const double* params, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!params) return fail(SWX_ERR_ARG, "null params");
  if ((rc = check_params(&p->d, params))) return rc;
  p->params.assign(params, params + (size_t)p->d.n_members * SWX_NPARAM);
  std::vector<double> coef;
  fill_coef(p, coef);
  hipStream_t s = (hipStream_t)stream;
  SWX_HIP(hipMemcpyAsync(p->coef, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, s));
  SWX_HIP(hipStreamSynchronize(s));
  return SWX_OK;
}

void swx_plan_destroy(swx_plan* p) {
  if (!p) return;
  p->graphs.clear();
  void* bufs[] = {p->Lam, p->Tbar, p->Gbar, p->pgrad, p->dgrad, p->adj_stc, p->Pf, p->Hf, p->Pi, p->Hi, p->fragP[0], p->fragP[1],
                  p->fragP[2], p->fragP[3], p->fragF[0], p->fragF[1], p->fragF[2], p->fragF[3], p->tw, p->TinvF, p->TfwdF, p->TinvH, p->TfwdH, p->TinvB, p->TfwdB, p->geo_d, p->f2, p->coef, p->S, p->T, p->G, p->Gprev,
                  p->raw, p->lag, p->g0uv, p->phys, p->uvic, p->part_uv, p->dead, p->R0f, p->R0p, p->resid_c, p->lagx,
                  p->Gsyn};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
  delete p;
}

int swx_init(swx_plan* p, const double* eta0, const double* delta0, const double* Phi0, const double* U0,
             const double* V0, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!eta0 || !delta0 || !Phi0) return fail(SWX_ERR_ARG, "eta0, delta0 and Phi0 are required");
  if ((U0 == nullptr) != (V0 == nullptr)) return fail(SWX_ERR_ARG, "Provide both U0 and V0, or neither.");
  hipStream_t s = (hipStream_t)stream;
  if (p->rsz == 8) return do_init<double>(p, eta0, delta0, Phi0, U0, V0, s);
  return do_init<float>(p, eta0, delta0, Phi0, U0, V0, s);
}

int swx_set_prev_level(swx_plan* p, const double* eta_prev, const double* delta_prev, const double* Phi_prev,
                       const unsigned char* dead, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->initialized) return fail(SWX_ERR_STATE, "swx_set_prev_level before swx_init");
  if (p->d.expflag || p->d.semi_implicit || p->d.raw_filter)
    return fail(SWX_ERR_ARG, "swx_set_prev_level: only the modified-Euler scheme reads no spectral prev level");
  if ((eta_prev == nullptr) != (delta_prev == nullptr) || (eta_prev == nullptr) != (Phi_prev == nullptr))
    return fail(SWX_ERR_ARG, "swx_set_prev_level: give all of eta_prev, delta_prev, Phi_prev or none");
  hipStream_t s = (hipStream_t)stream;
  const int B = p->d.n_members;
  const long long pl = (long long)p->geo.J * p->geo.I;
  if (eta_prev) {
    // The current levels and winds stay as the plan holds them.
    if ((rc = rebuild_phys(p, s))) return rc;
    const double* src[3] = {eta_prev, delta_prev, Phi_prev};
    // Physical planes 0, 2, 4 of [B][8][J][I] are the prev levels of eta, delta, Phi.
    for (int f = 0; f < 3; ++f) SWX_HIP(launch_copy_real<double>(p->phys + 2 * f * pl, 8 * pl, src[f], pl, pl, B, s));
  }
  if (dead) SWX_HIP(hipMemcpyAsync(p->dead, dead, (size_t)B, hipMemcpyDeviceToDevice, s));
  return SWX_OK;
}

int swx_run(swx_plan* p, int t0, int nsteps, double* hist_fields, double* hist_scalars, double* hist_phiext,
            unsigned char* hist_dead, int hist_T, int hist_k0, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->initialized) return fail(SWX_ERR_STATE, "swx_run before swx_init");
  if (nsteps < 0) return fail(SWX_ERR_ARG, "nsteps must be >= 0");
  const bool any_hist = hist_fields || hist_scalars || hist_phiext || hist_dead;
  if (any_hist && (hist_T <= 0 || hist_k0 < 0 || hist_k0 + nsteps > hist_T))
    return fail(SWX_ERR_ARG, "history window out of range");
  hipStream_t s = (hipStream_t)stream;
  const int lin = any_hist ? 0 : lineage_steps(p, nsteps);
  const int phys_from = lin ? nsteps : (any_hist ? 0 : std::max(0, nsteps - phys_window(p)));
  const int lineage_from = lin ? nsteps - lin : 1 << 30;
  // A call that writes the grid carry continues it from the rebuilt prev level.
  if (!lin && phys_from < nsteps && (rc = rebuild_phys(p, s))) return rc;
  if (p->rsz == 8)
    rc = do_steps<double>(p, t0, nsteps, hist_fields, hist_scalars, hist_phiext, hist_dead, hist_T, hist_k0, true, s,
                          phys_from, lineage_from);
  else
    rc = do_steps<float>(p, t0, nsteps, hist_fields, hist_scalars, hist_phiext, hist_dead, hist_T, hist_k0, true, s,
                         phys_from, lineage_from);
  if (rc == SWX_OK && lin && nsteps > 0) p->phys_stale = true;
  return rc;
}

int swx_run_graph(swx_plan* p, int t0, int nsteps, int graph_steps, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->initialized) return fail(SWX_ERR_STATE, "swx_run_graph before swx_init");
  if (nsteps <= 0) return SWX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (graph_steps > 0 && s == nullptr) return fail(SWX_ERR_ARG, "graph replay needs a non-default stream");
  // phys_on: from step phys_from of this call (the window that reaches the exported state), or
  // the spectral lineage from step lineage_from instead (lineage_steps).
  const int lin = lineage_steps(p, nsteps);
  const int phys_from = lin ? nsteps : std::max(0, nsteps - phys_window(p));
  const int lineage_from = lin ? nsteps - lin : 1 << 30;
  if (!lin && phys_from < nsteps && (rc = rebuild_phys(p, s))) return rc;
  auto steps = [&](int k0, int n, bool save_last) {
    const int pf = std::max(0, phys_from - k0);
    const int lf = std::max(0, lineage_from - k0);
    if (p->rsz == 8)
      return do_steps<double>(p, t0 + k0, n, nullptr, nullptr, nullptr, nullptr, 0, 0, save_last, s, pf, lf);
    return do_steps<float>(p, t0 + k0, n, nullptr, nullptr, nullptr, nullptr, 0, 0, save_last, s, pf, lf);
  };
  if (lin) p->phys_stale = true;
  // Steps with t <= 2 differ (no Robert-Asselin filter) and run eagerly.  The rest of the call
  // replays captured sequences: blocks of graph_steps steps while more than that many remain,
  // then the call's tail (at most graph_steps steps: the end of the window and the last step,
  // which saves the previous Fourier rows) as one sequence.  Each sequence is keyed by its shape
  // relative to its first step (GraphCache), so calls of the same length replay the same graphs.
  int k = 0;
  while (k < nsteps && t0 + k <= 2) {
    if ((rc = steps(k, 1, k == nsteps - 1))) return rc;
    ++k;
  }
  if (graph_steps <= 0) return k < nsteps ? steps(k, nsteps - k, true) : SWX_OK;
  const int shared = p->in_pipeline ? (p->pipe_same_m ? 2 : 1) : 0;
  auto replay = [&](int n, bool save_last) {
    const int pf = std::min(n, std::max(0, phys_from - k)), lf = std::min(n, std::max(0, lineage_from - k));
    const int key[GraphCache::kKey] = {n, pf, lf, n, n, save_last ? 1 : 0, shared};
    const int r = graph_replay(p->graphs, key, s, [&] {
      if (p->rsz == 8)
        return do_steps<double>(p, 3, n, nullptr, nullptr, nullptr, nullptr, 0, 0, save_last, s, pf, lf, true);
      return do_steps<float>(p, 3, n, nullptr, nullptr, nullptr, nullptr, 0, 0, save_last, s, pf, lf, true);
    });
    if (r) return r;
    p->nstep[0] += n;
    p->nstep[2] += n - lf;
    p->nstep[3] += n - pf;
    k += n;
    return SWX_OK;
  };
  while (nsteps - k > graph_steps)
    if ((rc = replay(graph_steps, false))) return rc;
  if (k < nsteps) return replay(nsteps - k, true);
  return SWX_OK;
}

int swx_ensemble_create(swx_plan* a, swx_plan* b, swx_ensemble** out) {
  if (!out) return fail(SWX_ERR_ARG, "null output");
  *out = nullptr;
  int rc = check_plan(a);
  if (rc == SWX_OK) rc = check_plan(b);
  if (rc) return rc;
  if (a->d.precision != SWX_PREC_FP32 || b->d.precision != SWX_PREC_FP32)
    return fail(SWX_ERR_ARG, "an ensemble needs FP32-transform plans");
  swx_plan* p0 = a->geo.M < b->geo.M ? a : b;
  swx_plan* p1 = a->geo.M < b->geo.M ? b : a;
  if (!mfma_pair_supported(p0->geo.M, p1->geo.M))
    return fail(SWX_ERR_ARG, "an ensemble needs two plans of different truncations among 42, 63, 106");
  if (!p0->use_mfma || !p0->use_leg_mfma || !p1->use_mfma || !p1->use_leg_mfma)
    return fail(SWX_ERR_ARG, "the MFMA kernels are unavailable for one of the plans on this device");
  swx_ensemble* e = new swx_ensemble();
  e->p[0] = p0;
  e->p[1] = p1;
  // The plans' pending work (their own streams) must not overlap the tuning launches.
  if (hipDeviceSynchronize() != hipSuccess) {
    delete e;
    return fail(SWX_ERR_HIP, "hipDeviceSynchronize failed");
  }
  if ((rc = tune_ensemble(e)) != SWX_OK) {
    delete e;
    return rc;
  }
  *out = e;
  return SWX_OK;
}

int swx_ensemble_run_graph(swx_ensemble* e, int t0, int nsteps, int graph_steps, void* stream) {
  if (!e) return fail(SWX_ERR_ARG, "null ensemble");
  int rc = check_plan(e->p[0]);
  if (rc == SWX_OK) rc = check_plan(e->p[1]);
  if (rc) return rc;
  if (!e->p[0]->initialized || !e->p[1]->initialized) return fail(SWX_ERR_STATE, "swx_ensemble_run_graph before init");
  if (nsteps <= 0) return SWX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (graph_steps > 0 && s == nullptr) return fail(SWX_ERR_ARG, "graph replay needs a non-default stream");
  int pf[2], lf[2];
  for (int q = 0; q < 2; ++q) {
    swx_plan* p = e->p[q];
    const int lin = lineage_steps(p, nsteps);
    pf[q] = lin ? nsteps : std::max(0, nsteps - phys_window(p));
    lf[q] = lin ? nsteps - lin : 1 << 30;
    if (!lin && pf[q] < nsteps && (rc = rebuild_phys(p, s))) return rc;
    if (lin) p->phys_stale = true;
  }
  auto steps = [&](int k0, int n, bool save_last) {
    const int pk[2] = {std::max(0, pf[0] - k0), std::max(0, pf[1] - k0)};
    const int lk[2] = {std::max(0, lf[0] - k0), std::max(0, lf[1] - k0)};
    return ens_steps(e, t0 + k0, n, save_last, s, pk, lk);
  };
  // As swx_run_graph: steps with t <= 2 eagerly, then blocks of graph_steps steps and the tail as
  // captured sequences.
  int k = 0;
  while (k < nsteps && t0 + k <= 2) {
    if ((rc = steps(k, 1, k == nsteps - 1))) return rc;
    ++k;
  }
  if (graph_steps <= 0) return k < nsteps ? steps(k, nsteps - k, true) : SWX_OK;
  // A plan option changed since the graphs were captured (kernel choices): capture again.
  if (e->epoch[0] != e->p[0]->opt_epoch || e->epoch[1] != e->p[1]->opt_epoch) {
    e->graphs.clear();
    e->epoch[0] = e->p[0]->opt_epoch;
    e->epoch[1] = e->p[1]->opt_epoch;
  }
  auto replay = [&](int n, bool save_last) {
    int pk[2], lk[2];
    for (int q = 0; q < 2; ++q) {
      pk[q] = std::min(n, std::max(0, pf[q] - k));
      lk[q] = std::min(n, std::max(0, lf[q] - k));
    }
    const int key[GraphCache::kKey] = {n, pk[0], lk[0], pk[1], lk[1], save_last ? 1 : 0, 0};
    const int r = graph_replay(e->graphs, key, s, [&] { return ens_steps(e, 3, n, save_last, s, pk, lk, true); });
    if (r) return r;
    for (int q = 0; q < 2; ++q) {
      e->p[q]->nstep[0] += n;
      e->p[q]->nstep[2] += n - lk[q];
      e->p[q]->nstep[3] += n - pk[q];
    }
    k += n;
    return SWX_OK;
  };
  while (nsteps - k > graph_steps)
    if ((rc = replay(graph_steps, false))) return rc;
  if (k < nsteps) return replay(nsteps - k, true);
  return SWX_OK;
}

void swx_ensemble_destroy(swx_ensemble* e) {
  if (!e) return;
  e->graphs.clear();
  delete e;
}

// Whether kernels on streams a and b run at the same time: one single-wave wait of ~200 us of
// device wall clock on each, both released by one event; on distinct hardware queues the pair
// takes about one wait, on a shared queue about two.  1 = overlap, 0 = serialized, -1 = error.
static int streams_overlap(hipStream_t a, hipStream_t b) {
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
  const unsigned long long ticks = (unsigned long long)khz / 5;  // 200 us
  hipEvent_t e[3] = {nullptr, nullptr, nullptr};
  int res = -1;
  float ta = 0.f, tb = 0.f;
  bool ok = true;
  for (auto& ev : e) ok = ok && hipEventCreate(&ev) == hipSuccess;
  // Warm both streams up first (a stream's first launch can carry set-up time).
  ok = ok && launch_wait_ticks(1, a) == hipSuccess && launch_wait_ticks(1, b) == hipSuccess &&
       hipStreamSynchronize(a) == hipSuccess && hipStreamSynchronize(b) == hipSuccess;
  ok = ok && hipEventRecord(e[0], a) == hipSuccess && hipStreamWaitEvent(b, e[0], 0) == hipSuccess &&
       launch_wait_ticks(ticks, a) == hipSuccess && launch_wait_ticks(ticks, b) == hipSuccess &&
       hipEventRecord(e[1], a) == hipSuccess && hipEventRecord(e[2], b) == hipSuccess &&
       hipEventSynchronize(e[1]) == hipSuccess && hipEventSynchronize(e[2]) == hipSuccess &&
       hipEventElapsedTime(&ta, e[0], e[1]) == hipSuccess && hipEventElapsedTime(&tb, e[0], e[2]) == hipSuccess;
  if (ok) res = (ta > tb ? ta : tb) < 0.3f ? 1 : 0;
  for (auto& ev : e)
    if (ev) (void)hipEventDestroy(ev);
  return res;
}

int swx_pipeline_create(swx_plan* const* plans, int n, swx_pipeline** out) {
  if (!out) return fail(SWX_ERR_ARG, "null output");
  *out = nullptr;
  if (!plans || n < 2 || n > kMaxPipeline)
    return fail(SWX_ERR_ARG, "a pipeline takes 2 to " + std::to_string(kMaxPipeline) + " plans");
  for (int k = 0; k < n; ++k) {
    int rc = check_plan(plans[k]);
    if (rc) return rc;
    for (int q = 0; q < k; ++q)
      if (plans[q] == plans[k]) return fail(SWX_ERR_ARG, "a pipeline needs distinct plans");
  }
  swx_pipeline* pl = new swx_pipeline();
  pl->n = n;
  bool ok = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; k < n && ok; ++k) {
    pl->p[k] = plans[k];
    // One default-priority stream per plan that does not synchronize with the null stream.
    // Every plan runs on a stream of its own, created here back to back: the runtime gives a
    // new stream its least-used hardware queue, so these land on distinct queues and the
    // plans overlap.  (The caller's stream may share a queue with one of them: a plan run on
    // it could then serialize behind the other.)
    ok = hipStreamCreateWithFlags(&pl->s[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&pl->join[k], hipEventDisableTiming) == hipSuccess;
  }
  // The runtime can still hand a new stream a queue that another of these streams uses.  Check
  // each pair; on a clash keep the clashing stream aside while a new one is created (the
  // runtime then prefers another queue) and try again, a few times at most.
  constexpr int kQueueTries = 4;
  hipStream_t spare[kMaxPipeline * kQueueTries] = {};
  int nspare = 0;
  for (int k = 1; k < n && ok; ++k) {
    for (int tries = 0; tries < kQueueTries && ok; ++tries) {
      int all = 1;
      for (int j = 0; j < k && all == 1; ++j) all = streams_overlap(pl->s[j], pl->s[k]);
      if (all < 0) ok = false;
      if (all != 0 || tries + 1 == kQueueTries) break;
      spare[nspare++] = pl->s[k];
      pl->s[k] = nullptr;
      ok = hipStreamCreateWithFlags(&pl->s[k], hipStreamNonBlocking) == hipSuccess;
    }
  }
  for (int q = 0; q < nspare; ++q) (void)hipStreamDestroy(spare[q]);
  if (!ok) {
    swx_pipeline_destroy(pl);
    return fail(SWX_ERR_HIP, "pipeline stream or event creation failed");
  }
  // The plans' graphs of the standalone mode stay cached; swx_run_graph keys its graphs by the mode.
  bool same_m = true;
  for (int k = 1; k < n; ++k) same_m = same_m && plans[k]->geo.M == plans[0]->geo.M;
  for (int k = 0; k < n; ++k) {
    plans[k]->in_pipeline = true;
    plans[k]->pipe_same_m = same_m;
  }
  *out = pl;
  return SWX_OK;
}

int swx_pipeline_run_graph(swx_pipeline* pl, int t0, int nsteps, int graph_steps, void* stream) {
  if (!pl) return fail(SWX_ERR_ARG, "null pipeline");
  // The caller's stream (the null stream too) only forks and joins: the plans replay their
  // graphs on the pipeline's own streams.
  hipStream_t s0 = (hipStream_t)stream;
  for (int k = 0; k < pl->n; ++k) {
    int rc = check_plan(pl->p[k]);
    if (rc) return rc;
    if (!pl->p[k]->initialized) return fail(SWX_ERR_STATE, "swx_pipeline_run_graph before init");
  }
  if (nsteps <= 0) return SWX_OK;
  // Fork: the internal streams start after the caller's prior work.
  SWX_HIP(hipEventRecord(pl->fork, s0));
  for (int k = 0; k < pl->n; ++k) SWX_HIP(hipStreamWaitEvent(pl->s[k], pl->fork, 0));
  for (int k = 0; k < pl->n; ++k) {
    int rc = swx_run_graph(pl->p[k], t0, nsteps, graph_steps, (void*)pl->s[k]);
    if (rc) return rc;
  }
  // Join: the caller's later work waits for every plan.
  for (int k = 0; k < pl->n; ++k) {
    SWX_HIP(hipEventRecord(pl->join[k], pl->s[k]));
    SWX_HIP(hipStreamWaitEvent(s0, pl->join[k], 0));
  }
  return SWX_OK;
}

void swx_pipeline_destroy(swx_pipeline* pl) {
  if (!pl) return;
  for (int k = 0; k < pl->n; ++k)
    if (pl->p[k]) {
      pl->p[k]->in_pipeline = false;
      pl->p[k]->pipe_same_m = false;
    }
  for (int k = 0; k < kMaxPipeline; ++k) {
    if (pl->s[k]) {
      (void)hipStreamSynchronize(pl->s[k]);
      (void)hipStreamDestroy(pl->s[k]);
    }
    if (pl->join[k]) (void)hipEventDestroy(pl->join[k]);
  }
  if (pl->fork) (void)hipEventDestroy(pl->fork);
  delete pl;
}

int swx_plan_set_option(swx_plan* p, int option, int value) {
  int rc = check_plan(p);
  if (rc) return rc;
  switch (option) {
    case SWX_OPT_GRID_KERNEL:
      if (value != 0 && value != 1) return fail(SWX_ERR_ARG, "SWX_OPT_GRID_KERNEL must be 0 (auto) or 1 (plain)");
      p->opt_plain_grid = value == 1;
      break;
    case SWX_OPT_FULL_PHYS_CARRY:
      p->opt_full_phys = value != 0;
      break;
    case SWX_OPT_LEG_KERNEL:
      if (value != 0 && value != 1) return fail(SWX_ERR_ARG, "SWX_OPT_LEG_KERNEL must be 0 (auto) or 1 (plain)");
      p->opt_plain_leg = value == 1;
      break;
    case SWX_OPT_LEG_MB: {
      int mbs[3];
      const int nc = leg_mfma_candidates(p->geo.M, mbs);
      bool ok = false;
      for (int c = 0; c < nc; ++c) ok = ok || mbs[c] == value;
      int maxlds = 0;
      (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device);
      if (!ok || leg_mfma_lds_bytes(p->geo.M, value) > (size_t)maxlds)
        return fail(SWX_ERR_ARG, "SWX_OPT_LEG_MB: not an available member-group size for this truncation");
      p->leg_mb = value;
      break;
    }
    case SWX_OPT_PHYS_LINEAGE:
      if (value < -1) return fail(SWX_ERR_ARG, "SWX_OPT_PHYS_LINEAGE must be >= -1");
      p->opt_lineage = value;
      break;
    case SWX_OPT_LEG_FOLD:
      if (value != 1 && value != 2) return fail(SWX_ERR_ARG, "SWX_OPT_LEG_FOLD must be 1 (full) or 2 (folded)");
      if (value == 2 && !p->fragF[0]) return fail(SWX_ERR_ARG, "SWX_OPT_LEG_FOLD: no folded tables for this plan");
      p->leg_fold = value;
      break;
    case SWX_OPT_ADJ_DIFFUSION:
      if (value != 0 && value != 1) return fail(SWX_ERR_ARG, "SWX_OPT_ADJ_DIFFUSION must be 0 or 1");
      p->opt_adj_diffusion = value == 1;
      break;
    case SWX_OPT_PHYS_STEPS:
      if (value < -1) return fail(SWX_ERR_ARG, "SWX_OPT_PHYS_STEPS must be >= -1");
      p->opt_phys_steps = value;
      break;
    case SWX_OPT_GRID_FOLD:
      if (value != 1 && value != 2) return fail(SWX_ERR_ARG, "SWX_OPT_GRID_FOLD must be 1 (full DFTs) or 2 (folded)");
      if (value == 2 && !p->has_fold)
        return fail(SWX_ERR_ARG, "SWX_OPT_GRID_FOLD: the folded grid kernel is not available for this plan");
      p->grid_fold = value;
      break;
    case SWX_OPT_CONCURRENT:
      p->in_pipeline = value != 0;
      if (!p->in_pipeline) p->pipe_same_m = false;
      break;
    default:
      return fail(SWX_ERR_ARG, "unknown option " + std::to_string(option));
  }
  p->graphs.clear();
  ++p->opt_epoch;
  return SWX_OK;
}

int swx_time_kernels(swx_plan* p, int t0, int nsteps, float* ms_leg, float* ms_grid, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->initialized) return fail(SWX_ERR_STATE, "swx_time_kernels before swx_init");
  if (nsteps <= 0 || !ms_leg || !ms_grid) return fail(SWX_ERR_ARG, "nsteps must be > 0 and outputs non-null");
  hipStream_t s = (hipStream_t)stream;
  // Same physical-carry mode as the steady steps of swx_run_graph.
  const int phys_on = phys_window(p) >= (1 << 30) ? 1 : 0;
  auto launch_grid_step = [&](int t) -> int {
    if (p->rsz == 8) {
      GridArgs<double> g = grid_args<double>(p, GRID_STEP);
      g.do_ra = filter_on(p, t);
      g.phys_on = phys_on;
      SWX_HIP(launch_grid<double>(g, s));
    } else {
      GridArgs<float> g = grid_args<float>(p, GRID_STEP);
      g.do_ra = filter_on(p, t);
      g.phys_on = phys_on;
      SWX_HIP(launch_grid_any<float>(p, g, s));
    }
    return SWX_OK;
  };
  auto launch_leg_step = [&](int t) -> int {
    if (p->rsz == 8) {
      LegArgs<double> l = leg_args<double>(p, LEG_STEP);
      l.do_ra = filter_on(p, t);
      SWX_HIP(launch_leg<double>(l, s));
    } else {
      LegArgs<float> l = leg_args<float>(p, LEG_STEP);
      l.do_ra = filter_on(p, t);
      SWX_HIP(launch_leg_any<float>(p, l, s));
    }
    return SWX_OK;
  };
  hipEvent_t ev[4];
  for (auto& e : ev) SWX_HIP(hipEventCreate(&e));
  double step_ms = 0.0, grid_ms = 0.0;
  if (!phys_on) {
    SWX_HIP(hipEventRecord(ev[0], s));
    for (int k = 0; k < nsteps && rc == SWX_OK; ++k)
      if ((rc = launch_leg_step(t0 + k)) == SWX_OK) rc = launch_grid_step(t0 + k);
    SWX_HIP(hipEventRecord(ev[1], s));
    // With the physical carry off the grid launch reads G and rewrites the same T:
    // repeating it leaves the state as it is.
    for (int k = 0; k < nsteps && rc == SWX_OK; ++k) rc = launch_grid_step(t0 + nsteps - 1);
    SWX_HIP(hipEventRecord(ev[2], s));
    SWX_HIP(hipEventSynchronize(ev[2]));
    float a = 0.f, b = 0.f;
    SWX_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
    SWX_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
    step_ms = a / nsteps;
    grid_ms = b / nsteps;
  } else {
    // Carry written every step (repeats would advance the RA filter): events around each launch.
    for (int k = 0; k < nsteps && rc == SWX_OK; ++k) {
      SWX_HIP(hipEventRecord(ev[0], s));
      if ((rc = launch_leg_step(t0 + k)) != SWX_OK) break;
      SWX_HIP(hipEventRecord(ev[1], s));
      if ((rc = launch_grid_step(t0 + k)) != SWX_OK) break;
      SWX_HIP(hipEventRecord(ev[2], s));
      SWX_HIP(hipEventSynchronize(ev[2]));
      float a = 0.f, b = 0.f;
      SWX_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
      SWX_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
      step_ms += (double)(a + b) / nsteps;
      grid_ms += (double)b / nsteps;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  *ms_grid = (float)grid_ms;
  *ms_leg = (float)(step_ms - grid_ms > 0.0 ? step_ms - grid_ms : 0.0);
  return SWX_OK;
}

int swx_plan_get_option(const swx_plan* p, int option, int* value) {
  if (!p || !value) return fail(SWX_ERR_ARG, "null plan or output");
  switch (option) {
    case SWX_OPT_GRID_KERNEL: *value = p->opt_plain_grid ? 1 : 0; break;
    case SWX_OPT_FULL_PHYS_CARRY: *value = p->opt_full_phys ? 1 : 0; break;
    case SWX_OPT_LEG_KERNEL: *value = p->opt_plain_leg ? 1 : 0; break;
    case SWX_OPT_PHYS_STEPS: *value = p->opt_phys_steps; break;
    case SWX_OPT_PHYS_LINEAGE: *value = p->opt_lineage; break;
    case SWX_OPT_LEG_FOLD: *value = p->leg_fold == 2 && p->fragF[0] ? 2 : 1; break;
    case SWX_OPT_LEG_MB: *value = p->leg_mb; break;
    case SWX_OPT_ADJ_DIFFUSION: *value = p->opt_adj_diffusion ? 1 : 0; break;
    case SWX_OPT_GRID_FOLD: *value = p->grid_fold; break;
    case SWX_OPT_CONCURRENT: *value = p->in_pipeline ? 1 : 0; break;
    default: return fail(SWX_ERR_ARG, "unknown option " + std::to_string(option));
  }
  return SWX_OK;
}

int swx_plan_step_counts(const swx_plan* p, long long* counts) {
  if (!p || !counts) return fail(SWX_ERR_ARG, "null plan or output");
  for (int k = 0; k < 4; ++k) counts[k] = p->nstep[k];
  return SWX_OK;
}

int swx_plan_device_bytes(const swx_plan* p, size_t* bytes) {
  if (!p || !bytes) return fail(SWX_ERR_ARG, "null plan or output");
  *bytes = p->dev_bytes;
  return SWX_OK;
}

int swx_init_spectral(swx_plan* p, const double* c, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!c) return fail(SWX_ERR_ARG, "null spectral state");
  hipStream_t s = (hipStream_t)stream;
  const int B = p->d.n_members;
  SWX_HIP(hipMemsetAsync(p->dead, 0, (size_t)B, s));
  if ((rc = copy_state(p, nullptr, c, true, true, s))) return rc;
  rc = p->rsz == 8 ? prime<double>(p, true, s) : prime<float>(p, true, s);
  if (rc) return rc;
  p->initialized = true;
  p->phys_stale = false;
  return SWX_OK;
}

int swx_state_get(swx_plan* p, double* out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!out) return fail(SWX_ERR_ARG, "null output");
  if (!p->initialized) return fail(SWX_ERR_STATE, "swx_state_get before initialization");
  if (p->d.semi_implicit || p->d.raw_filter)
    return fail(SWX_ERR_ARG, "spectral checkpoints hold one level: not available with semi_implicit / raw_filter");
  return copy_state(p, out, nullptr, false, false, (hipStream_t)stream);
}

int swx_state_set(swx_plan* p, const double* in, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!in) return fail(SWX_ERR_ARG, "null state");
  if (!p->initialized) return fail(SWX_ERR_STATE, "swx_state_set before initialization");
  if (p->d.semi_implicit || p->d.raw_filter)
    return fail(SWX_ERR_ARG, "spectral checkpoints hold one level: not available with semi_implicit / raw_filter");
  hipStream_t s = (hipStream_t)stream;
  if ((rc = copy_state(p, nullptr, in, true, true, s))) return rc;
  return p->rsz == 8 ? prime<double>(p, false, s) : prime<float>(p, false, s);
}

int swx_adjoint_begin(swx_plan* p, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (p->d.expflag || p->d.semi_implicit || p->d.raw_filter)
    return fail(SWX_ERR_ARG, "the adjoint covers the modified-Euler scheme (expflag = 0, no semi_implicit / raw_filter)");
  if ((rc = ensure_adjoint(p))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int B = p->d.n_members, J = p->geo.J, Mp = p->geo.M + 1, Np = p->geo.N + 1;
  SWX_HIP(hipMemsetAsync(p->Lam, 0, (size_t)B * 3 * Mp * Np * 2 * 8, s));
  SWX_HIP(hipMemsetAsync(p->pgrad, 0, (size_t)B * J * 3 * 8, s));
  SWX_HIP(hipMemsetAsync(p->dgrad, 0, (size_t)B * 2 * 8, s));
  return SWX_OK;
}

int swx_adjoint_terminal(swx_plan* p, const double* g_eta, const double* g_delta, const double* g_Phi,
                         const double* g_U, const double* g_V, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->Lam) return fail(SWX_ERR_STATE, "swx_adjoint_terminal before swx_adjoint_begin");
  hipStream_t s = (hipStream_t)stream;
  return p->rsz == 8 ? adjoint_terminal<double>(p, g_eta, g_delta, g_Phi, g_U, g_V, s)
                     : adjoint_terminal<float>(p, g_eta, g_delta, g_Phi, g_U, g_V, s);
}

int swx_adjoint_observe(swx_plan* p, const double* g_eta, const double* g_delta, const double* g_Phi,
                        const double* g_U, const double* g_V, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->Lam) return fail(SWX_ERR_STATE, "swx_adjoint_observe before swx_adjoint_begin");
  hipStream_t s = (hipStream_t)stream;
  return p->rsz == 8 ? adjoint_terminal<double>(p, g_eta, g_delta, g_Phi, g_U, g_V, s, 1)
                     : adjoint_terminal<float>(p, g_eta, g_delta, g_Phi, g_U, g_V, s, 1);
}

int swx_run_tape(swx_plan* p, int t0, int nsteps, double* tape, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->initialized) return fail(SWX_ERR_STATE, "swx_run_tape before initialization");
  if (nsteps < 0) return fail(SWX_ERR_ARG, "nsteps must be >= 0");
  if (nsteps > 0 && !tape) return fail(SWX_ERR_ARG, "null tape");
  if (p->d.semi_implicit || p->d.raw_filter)
    return fail(SWX_ERR_ARG, "spectral tapes hold one level: not available with semi_implicit / raw_filter");
  hipStream_t s = (hipStream_t)stream;
  const size_t per = (size_t)p->d.n_members * 3 * (p->geo.M + 1) * (p->geo.N + 1) * 2;  // doubles per state
  for (int k = 0; k < nsteps; ++k) {
    if ((rc = copy_state(p, tape + k * per, nullptr, false, false, s))) return rc;
    // Only the last step writes the physical grid carry (its current level is exact).
    const bool last = k == nsteps - 1;
    rc = p->rsz == 8 ? do_steps<double>(p, t0 + k, 1, nullptr, nullptr, nullptr, nullptr, 0, 0, last, s, last ? 0 : 1 << 30)
                     : do_steps<float>(p, t0 + k, 1, nullptr, nullptr, nullptr, nullptr, 0, 0, last, s, last ? 0 : 1 << 30);
    if (rc) return rc;
  }
  return SWX_OK;
}

int swx_adjoint_sweep(swx_plan* p, int nsteps, const double* tape, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->Lam) return fail(SWX_ERR_STATE, "swx_adjoint_sweep before swx_adjoint_begin");
  if (nsteps < 0) return fail(SWX_ERR_ARG, "nsteps must be >= 0");
  if (nsteps > 0 && !tape) return fail(SWX_ERR_ARG, "null tape");
  hipStream_t s = (hipStream_t)stream;
  const size_t per = (size_t)p->d.n_members * 3 * (p->geo.M + 1) * (p->geo.N + 1) * 2;
  const bool diffusion = p->opt_adj_diffusion && p->d.diffflag;
  for (int k = nsteps - 1; k >= 0; --k) {
    // The pair (dL/dS_{k+1}, S_{k+1}) inside the tape; the caller adds the one past its end
    // (swx_adjoint_diffusion) before the sweep.
    if (diffusion && k < nsteps - 1)
      SWX_HIP(launch_diffusion_partial(p->Lam, tape + (k + 1) * per, p->coef, p->cl.ncoef, p->geo.M, p->geo.N,
                                       p->geo.J, p->d.n_members, p->dgrad, s));
    rc = p->rsz == 8 ? adjoint_step<double>(p, tape + k * per, s) : adjoint_step<float>(p, tape + k * per, s);
    if (rc) return rc;
  }
  return SWX_OK;
}

int swx_adjoint_step(swx_plan* p, const double* state_t, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->Lam) return fail(SWX_ERR_STATE, "swx_adjoint_step before swx_adjoint_begin");
  if (!state_t) return fail(SWX_ERR_ARG, "null state");
  hipStream_t s = (hipStream_t)stream;
  return p->rsz == 8 ? adjoint_step<double>(p, state_t, s) : adjoint_step<float>(p, state_t, s);
}

int swx_adjoint_diffusion(swx_plan* p, const double* state, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->Lam) return fail(SWX_ERR_STATE, "swx_adjoint_diffusion before swx_adjoint_begin");
  if (!state) return fail(SWX_ERR_ARG, "null state");
  if (!p->d.diffflag) return SWX_OK;  // no diffusion factor: the partials stay zero
  SWX_HIP(launch_diffusion_partial(p->Lam, state, p->coef, p->cl.ncoef, p->geo.M, p->geo.N, p->geo.J,
                                   p->d.n_members, p->dgrad, (hipStream_t)stream));
  return SWX_OK;
}

int swx_adjoint_diffusion_result(swx_plan* p, double* out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->Lam) return fail(SWX_ERR_STATE, "swx_adjoint_diffusion_result before swx_adjoint_begin");
  if (!out) return fail(SWX_ERR_ARG, "null output");
  SWX_HIP(hipMemcpyAsync(out, p->dgrad, (size_t)p->d.n_members * 2 * 8, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return SWX_OK;
}

int swx_adjoint_result(swx_plan* p, double* lam_out, double* pgrad_out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->Lam) return fail(SWX_ERR_STATE, "swx_adjoint_result before swx_adjoint_begin");
  hipStream_t s = (hipStream_t)stream;
  const int B = p->d.n_members, J = p->geo.J, Mp = p->geo.M + 1, Np = p->geo.N + 1;
  if (lam_out)
    SWX_HIP(hipMemcpyAsync(lam_out, p->Lam, (size_t)B * 3 * Mp * Np * 2 * 8, hipMemcpyDeviceToDevice, s));
  if (pgrad_out) SWX_HIP(launch_pgrad_reduce(p->pgrad, pgrad_out, B, J, s));
  return SWX_OK;
}

}  // extern "C"

namespace {

// swx_export_state with the rows of a physical-carry rebuild in `rows` (null: the plan's own).
int export_state(swx_plan* p, const swx_state_out* o, hipStream_t s, void* rows) {
  int rc;
  const int B = p->d.n_members, J = p->geo.J, I = p->geo.I, Mp = p->geo.M + 1, Np = p->geo.N + 1;
  const long long cjm = (long long)J * Mp * 2;  // doubles per complex (J, M+1) field
  const long long pl = (long long)J * I;
  auto cp = [&](double* dst, const void* base, int nfield, int f, long long per) -> int {
    if (!dst) return SWX_OK;
    hipError_t e;
    if (p->rsz == 8)
      e = launch_copy_real<double>(dst, per, (const double*)base + f * per, nfield * per, per, B, s);
    else
      e = launch_copy_real<float>(dst, per, (const float*)base + f * per, nfield * per, per, B, s);
    if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("export copy: ") + hipGetErrorString(e));
    return SWX_OK;
  };
  auto cpd = [&](double* dst, const double* base, int nfield, int f, long long per) -> int {
    if (!dst) return SWX_OK;
    hipError_t e = launch_copy_real<double>(dst, per, base + f * per, nfield * per, per, B, s);
    if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("export copy: ") + hipGetErrorString(e));
    return SWX_OK;
  };
  auto cf = [&](double* dst, const void* base, int f) -> int {
    if (!dst) return SWX_OK;
    hipError_t e;
    if (p->rsz == 8)
      e = launch_copy_fourier<double>(dst, (const double*)base, f, B, J, Mp, s);
    else
      e = launch_copy_fourier<float>(dst, (const float*)base, f, B, J, Mp, s);
    if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("export copy: ") + hipGetErrorString(e));
    return SWX_OK;
  };
  const bool pinned = p->d.test == 1;
  if (p->d.semi_implicit || p->d.raw_filter) {
    // Both Fourier levels are filtered: each is the inverse Legendre transform of its
    // spectral level (into Gprev, free outside the modified-Euler export) plus the initial
    // fields' residue times that level's coefficient.
    for (int lvl = 1; lvl >= 0; --lvl) {
      double* dst[3] = {lvl ? o->etam_curr : o->etam_prev, lvl ? o->deltam_curr : o->deltam_prev,
                        lvl ? o->Phim_curr : o->Phim_prev};
      if (!dst[0] && !dst[1] && !dst[2]) continue;
      for (int f = 0; f < 3; ++f) {
        hipError_t e = hipSuccess;
        if (p->rsz == 8) {
          if (f == 0) {
            LegArgs<double> l = leg_args<double>(p, LEG_PRIME);
            l.diagnostics = 0;
            l.G = (double*)p->Gprev;
            l.prime_prev = lvl == 0;
            e = launch_leg<double>(l, s);
          }
          if (e == hipSuccess && dst[f])
            e = launch_copy_fourier<double>(dst[f], (const double*)p->Gprev, f, B, J, Mp, s, (const double*)p->R0f,
                                            p->resid_c, lvl);
        } else {
          if (f == 0) {
            LegArgs<float> l = leg_args<float>(p, LEG_PRIME);
            l.diagnostics = 0;
            l.G = (float*)p->Gprev;
            l.prime_prev = lvl == 0;
            e = launch_leg<float>(l, s);
          }
          if (e == hipSuccess && dst[f])
            e = launch_copy_fourier<float>(dst[f], (const float*)p->Gprev, f, B, J, Mp, s, (const float*)p->R0f,
                                           p->resid_c, lvl);
        }
        if (e != hipSuccess) return fail(SWX_ERR_HIP, std::string("export: ") + hipGetErrorString(e));
      }
    }
  } else if ((rc = cf(o->etam_curr, p->G, 0)) || (rc = cf(o->deltam_curr, p->G, 1)) ||
             (rc = cf(o->Phim_curr, p->G, 2)) || (rc = cf(o->etam_prev, p->Gprev, 0)) ||
             (rc = cf(o->deltam_prev, p->Gprev, 1)) || (rc = cf(o->Phim_prev, p->Gprev, 2))) {
    return rc;
  }
  // Semi-implicit plans move the lagged remainder every step; the exported one is in lagx.
  void* lagsrc = p->d.semi_implicit ? p->lagx : p->lag;
  if ((rc = pinned ? cp(o->Um_curr, p->g0uv, 2, 0, cjm) : cf(o->Um_curr, p->G, 3)) ||
      (rc = pinned ? cp(o->Vm_curr, p->g0uv, 2, 1, cjm) : cf(o->Vm_curr, p->G, 4)) ||
      (rc = cp(o->Am_curr, p->raw, kNRaw, 0, cjm)) || (rc = cp(o->Bm_curr, p->raw, kNRaw, 1, cjm)) ||
      (rc = cp(o->Cm_curr, p->raw, kNRaw, 2, cjm)) || (rc = cp(o->Dm_curr, p->raw, kNRaw, 3, cjm)) ||
      (rc = cp(o->Em_curr, p->raw, kNRaw, 4, cjm)) || (rc = cp(o->PhiFm_curr, p->raw, kNRaw, 5, cjm)) ||
      (rc = cp(o->Fm_curr, p->raw, kNRaw, 6, cjm)) || (rc = cp(o->Gm_curr, p->raw, kNRaw, 7, cjm)) ||
      (rc = cp(o->Rum_lag, lagsrc, 2, 0, cjm)) || (rc = cp(o->Rvm_lag, lagsrc, 2, 1, cjm)))
    return rc;
  if ((o->eta_prev || o->eta_curr || o->delta_prev || o->delta_curr || o->Phi_prev || o->Phi_curr || o->U_curr ||
       o->V_curr) &&
      (rc = rebuild_phys(p, s, rows)))
    return rc;
  if ((rc = cpd(o->eta_prev, p->phys, 8, 0, pl)) || (rc = cpd(o->eta_curr, p->phys, 8, 1, pl)) ||
      (rc = cpd(o->delta_prev, p->phys, 8, 2, pl)) || (rc = cpd(o->delta_curr, p->phys, 8, 3, pl)) ||
      (rc = cpd(o->Phi_prev, p->phys, 8, 4, pl)) || (rc = cpd(o->Phi_curr, p->phys, 8, 5, pl)) ||
      (rc = cpd(o->U_curr, p->phys, 8, 6, pl)) || (rc = cpd(o->V_curr, p->phys, 8, 7, pl)))
    return rc;
  if (o->spec) {
    // Current spectral level: [B][2][3][M+1][N+1] -> [B][3][M+1][N+1].
    const long long per = 3LL * Mp * Np * 2;
    if ((rc = cpd(o->spec, p->S + per, 2, 0, per))) return rc;
  }
  if (o->dead) SWX_HIP(hipMemcpyAsync(o->dead, p->dead, (size_t)B, hipMemcpyDeviceToDevice, s));
  return SWX_OK;
}

// Scratch of swx_export_state_noalloc: the Fourier rows of the spectral prev lineage that a rebuild
// of the physical carry writes ([B][5][M+1][J] complex in the transform precision), at most
// kScratchBytes.  One copy per GPU, resident from the library's load until the process ends; every
// call on a GPU shares it and it keeps the last call's contents (a rebuild overwrites every row it
// reads, so nothing is cleared).  Host code never names it: it uses the per-device address from
// hipGetSymbolAddress (scratch_address).
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
void* g_scratch_ptr[kMaxScratchDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on device `dev` (the current device), looked up once per device.
int scratch_address(int dev, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(SWX_ERR_ARG, "device index out of range for the scratch");
  if (!g_scratch_ptr[dev]) {
    void* q = nullptr;
    hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q)
      return fail(SWX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = q;
  }
  *out = g_scratch_ptr[dev];
  return SWX_OK;
}

}  // namespace

namespace {

// Shared checks of the standalone transforms: K batch rows of J latitudes; K is the launch's
// second grid dimension.
int transform_dims(const char* name, int K, int J) {
  if (K < 0) return fail(SWX_ERR_ARG, std::string(name) + ": negative batch count");
  if (K > 65535) return fail(SWX_ERR_ARG, std::string(name) + ": batch count above 65535");
  if (J < 1) return fail(SWX_ERR_ARG, std::string(name) + ": J must be >= 1");
  return SWX_OK;
}

int dft_length(const char* name, int I) {
  if (I < 1 || I > kMaxDftLength)
    return fail(SWX_ERR_ARG, std::string(name) + ": I = " + std::to_string(I) + " outside [1, " +
                                 std::to_string(kMaxDftLength) + "]");
  return SWX_OK;
}

}  // namespace

extern "C" {

int swx_export_state(swx_plan* p, const swx_state_out* o, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!o) return fail(SWX_ERR_ARG, "null output struct");
  if (!p->initialized) return fail(SWX_ERR_STATE, "export before swx_init");
  return export_state(p, o, (hipStream_t)stream, nullptr);
}

int swx_export_state_noalloc(swx_plan* p, const swx_state_out* o, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!o) return fail(SWX_ERR_ARG, "null output struct");
  if (!p->initialized) return fail(SWX_ERR_STATE, "export before swx_init");
  const size_t need = lineage_rows_bytes(p);
  if (need > kScratchBytes)
    return fail(SWX_ERR_ARG, "swx_export_state_noalloc: this plan's rebuild rows need " + std::to_string(need) +
                                 " bytes, above the " + std::to_string(kScratchBytes) +
                                 "-byte scratch (use swx_export_state, or fewer members per plan)");
  void* rows = nullptr;
  if ((rc = scratch_address(p->device, &rows))) return rc;
  return export_state(p, o, (hipStream_t)stream, rows);
}

static int phase_curve_dims(int n, int J, int I, int K) {
  if (n < 0) return fail(SWX_ERR_ARG, "phase curve: negative grid count");
  if (J < 1 || I < 1 || J > kPhaseCurveMaxI || I > kPhaseCurveMaxI)
    return fail(SWX_ERR_ARG, "phase curve: grid J x I = " + std::to_string(J) + " x " + std::to_string(I) +
                                 " outside 1.." + std::to_string(kPhaseCurveMaxI));
  if (K < 1 || K > kPhaseCurveMaxK)
    return fail(SWX_ERR_ARG, "phase curve: " + std::to_string(K) + " phases outside 1.." +
                                 std::to_string(kPhaseCurveMaxK));
  return SWX_OK;
}

int swx_phase_curve(const double* Phi, long long member_stride, int n, int J, int I, const double* mus,
                    const double* phases, int K, int normalize, double* flux, double* flux_raw, void* stream) {
  int rc = phase_curve_dims(n, J, I, K);
  if (rc) return rc;
  if (n == 0) return SWX_OK;
  if (!Phi || !mus || !phases || !flux) return fail(SWX_ERR_ARG, "phase curve: null pointer");
  if (member_stride < (long long)J * I) return fail(SWX_ERR_ARG, "phase curve: member stride below J * I");
  SWX_HIP(launch_phase_curve(Phi, member_stride, n, J, I, mus, phases, K, normalize, flux, flux_raw,
                             (hipStream_t)stream));
  return SWX_OK;
}

int swx_phase_curve_vjp(const double* g_flux, const double* flux_raw, int n, int J, int I, const double* mus,
                        const double* phases, int K, int normalize, double* g_Phi, void* stream) {
  int rc = phase_curve_dims(n, J, I, K);
  if (rc) return rc;
  if (n == 0) return SWX_OK;
  if (!g_flux || !flux_raw || !mus || !phases || !g_Phi) return fail(SWX_ERR_ARG, "phase curve vjp: null pointer");
  SWX_HIP(launch_phase_curve_vjp(g_flux, flux_raw, n, J, I, mus, phases, K, normalize, g_Phi, (hipStream_t)stream));
  return SWX_OK;
}

int swx_diagnostic_eta_delta(const double* Um, const double* Vm, const double* fmn, int n, int I, int J, int M, int N,
                             const double* Pmn, const double* Hmn, const double* w, const double* tstepcoeff,
                             const double* mJarray, double dt, double* eta, double* delta, double* etamn,
                             double* deltamn, void* stream) {
  int rc = transform_dims("diagnostic_eta_delta", n, J);
  if (rc) return rc;
  if (M < 0 || N < M) return fail(SWX_ERR_ARG, "diagnostic_eta_delta: need M >= 0 and N >= M");
  if ((eta || delta) && I < 2 * M + 1)
    return fail(SWX_ERR_ARG, "diagnostic_eta_delta: I = " + std::to_string(I) + " below 2 M + 1");
  if ((eta || delta) && (rc = dft_length("diagnostic_eta_delta", I))) return rc;
  if (!(dt != 0.0)) return fail(SWX_ERR_ARG, "diagnostic_eta_delta: dt must be nonzero");
  if (n == 0) return SWX_OK;
  if (!Um || !Vm || !fmn || !Pmn || !Hmn || !w || !tstepcoeff || !mJarray || !etamn || !deltamn)
    return fail(SWX_ERR_ARG, "diagnostic_eta_delta: null pointer");
  hipStream_t s = (hipStream_t)stream;
  SWX_HIP(launch_diag_eta_delta_spec(Um, Vm, fmn, n, J, M, N, Pmn, Hmn, w, tstepcoeff, mJarray, dt, etamn, deltamn,
                                     s));
  if (eta && delta) {
    SWX_HIP(launch_spec_to_grid(etamn, deltamn, n, J, I, M, N, Pmn, eta, delta, s));
  } else if (eta || delta) {
    SWX_HIP(launch_spec_to_grid(eta ? etamn : deltamn, nullptr, n, J, I, M, N, Pmn, eta ? eta : delta, nullptr, s));
  }
  return SWX_OK;
}

int swx_fwd_fft_trunc(const double* data, int is_complex, int K, int J, int C, int I, int M, double* out,
                      void* stream) {
  int rc = transform_dims("fwd_fft_trunc", K, J);
  if (rc || (rc = dft_length("fwd_fft_trunc", I))) return rc;
  if (C < 1 || M < 0) return fail(SWX_ERR_ARG, "fwd_fft_trunc: need C >= 1 columns and M >= 0");
  if (K == 0) return SWX_OK;
  if (!data || !out) return fail(SWX_ERR_ARG, "fwd_fft_trunc: null pointer");
  SWX_HIP(launch_fwd_dft_trunc(data, is_complex != 0, K, J, C, I, std::min(M + 1, I), out, (hipStream_t)stream));
  return SWX_OK;
}

int swx_invrs_fft(const double* X, int K, int J, int C, int I, double* out, void* stream) {
  int rc = transform_dims("invrs_fft", K, J);
  if (rc || (rc = dft_length("invrs_fft", I))) return rc;
  if (C < 1) return fail(SWX_ERR_ARG, "invrs_fft: need C >= 1 columns");
  if (K == 0) return SWX_OK;
  if (!X || !out) return fail(SWX_ERR_ARG, "invrs_fft: null pointer");
  SWX_HIP(launch_dft_synth(X, K, J, C, I, out, (hipStream_t)stream));
  return SWX_OK;
}

int swx_fwd_leg(const double* data, const double* basis, const double* w, int K, int J, int M, int N, double* out,
                void* stream) {
  int rc = transform_dims("fwd_leg", K, J);
  if (rc) return rc;
  if (M < 0 || N < 0) return fail(SWX_ERR_ARG, "fwd_leg: need M >= 0 and N >= 0");
  if (K == 0) return SWX_OK;
  if (!data || !basis || !out) return fail(SWX_ERR_ARG, "fwd_leg: null pointer");
  SWX_HIP(launch_fwd_leg(data, basis, w, K, J, M, N, out, (hipStream_t)stream));
  return SWX_OK;
}

int swx_invrs_leg(const double* S, const double* basis, int K, int I, int J, int M, int N, double* out,
                  void* stream) {
  int rc = transform_dims("invrs_leg", K, J);
  if (rc) return rc;
  if (M < 0 || N < 0) return fail(SWX_ERR_ARG, "invrs_leg: need M >= 0 and N >= 0");
  if (I < M + 1) return fail(SWX_ERR_ARG, "invrs_leg: I = " + std::to_string(I) + " below M + 1");
  if (M + 1 > 4096) return fail(SWX_ERR_ARG, "invrs_leg: M + 1 above 4096 (the kernel's LDS row)");
  if (K == 0) return SWX_OK;
  if (!S || !basis || !out) return fail(SWX_ERR_ARG, "invrs_leg: null pointer");
  SWX_HIP(launch_invrs_leg(S, basis, K, J, I, M, N, out, (hipStream_t)stream));
  return SWX_OK;
}

int swx_invrs_uv(const double* deltamn, const double* etamn, const double* fmn, int K, int I, int J, int M, int N,
                 const double* Pmn, const double* Hmn, const double* tstepcoeffmn, const double* marray, double* U,
                 double* V, double* Um, double* Vm, void* stream) {
  int rc = transform_dims("invrsUV", K, J);
  if (rc || (rc = dft_length("invrsUV", I))) return rc;
  if (M < 0 || N < M) return fail(SWX_ERR_ARG, "invrsUV: need M >= 0 and N >= M");
  if (I < 2 * M + 1) return fail(SWX_ERR_ARG, "invrsUV: I = " + std::to_string(I) + " below 2 M + 1");
  if (K == 0) return SWX_OK;
  if (!deltamn || !etamn || !fmn || !Pmn || !Hmn || !tstepcoeffmn || !marray)
    return fail(SWX_ERR_ARG, "invrsUV: null input");
  if (!U && !V && !Um && !Vm) return fail(SWX_ERR_ARG, "invrsUV: no output requested");
  SWX_HIP(launch_invrs_uv(deltamn, etamn, fmn, K, J, I, M, N, Pmn, Hmn, tstepcoeffmn, marray, U, V, Um, Vm,
                          (hipStream_t)stream));
  return SWX_OK;
}

int swx_plan_phase_curve(swx_plan* p, const double* phases, int K, int normalize, double* flux, double* flux_raw,
                         void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->initialized) return fail(SWX_ERR_STATE, "phase curve before swx_init");
  const int J = p->geo.J, I = p->geo.I;
  if ((rc = rebuild_phys(p, (hipStream_t)stream))) return rc;
  // Physical plane 5 of [B][8][J][I] is Phi_curr (the grid swx_export_state returns).
  return swx_phase_curve(p->phys + 5LL * J * I, 8LL * J * I, p->d.n_members, J, I, p->geo_d + p->gl.mus, phases,
                         K, normalize, flux, flux_raw, stream);
}

}  // extern "C"
