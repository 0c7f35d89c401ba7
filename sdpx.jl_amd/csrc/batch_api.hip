// C-ABI of the batched route: many independent single-block core-form SDPs
//   min c'x  s.t.  X = sum_i x_i A_i - C >= 0     (one PSD block, n_eq = 0)
// solved in lock step.  Every phase of SDPX's Newton step is one launch over
// the members still iterating (batch_kernels.hip; grouped MFMA GEMMs for the
// Schur Gram product and the Schur Cholesky); the per-member interior-point
// control (controller, stagnation, line-search outcome, collapse handling,
// restarts, best iterate) runs on the host in C++ over per-member scalars
// read back once per phase.
#include <algorithm>
#include <array>
#include <cmath>
#include <deque>
#include <string>
#include <vector>

#include "../../include/sdpx_mi355x.h"
#include "batch_kernels.hpp"

using namespace sdpx;
using namespace sdpx::batch;

namespace {

constexpr double kEps = 2.220446049250313e-16;
const double kSqrtEps = std::sqrt(kEps);
constexpr double kRefineMinDecrease = 0.5;  // kkt.jl:3634
constexpr int kRefineTolUlps = 64;          // kkt.jl:3637
constexpr double kNearSolutionMerit = 100.0;  // interior_point.jl:2703
constexpr double kCenteringBetaStep = 4.0;    // interior_point.jl:2706
constexpr double kCenteringBetaMax = 0.5;     // interior_point.jl:2709
constexpr int kPrecisionFloorUlps = 16;       // stagnation.jl:13
constexpr double kProjectionSlack = 10.0;     // stagnation.jl:26
constexpr int kPanelChunk = 16;               // constraints per Schur panel work-group
constexpr int kAlign = 32;                    // doubles: every member offset is 256-byte aligned
constexpr int kListRing = 64;                 // member-list staging slots between stream drains
constexpr int kParRing = 8;                   // parameter staging buffers (copies in flight without a drain)
constexpr int kTraceCols = SDPX_BATCH_TRACE_COLS;

long long align_up(long long v) { return (v + kAlign - 1) / kAlign * kAlign; }

// ---------------------------------------------------------------------------
// Interior-point control (types/workspaces.jl:101-228, adaptive_parameters.jl,
// stagnation.jl, interior_point.jl:1323-2420), one instance per member.
// ---------------------------------------------------------------------------
struct Params {
  double sigma = 0, pf = 0, df = 0, bt = 0, rtol = 0;
  int rmax = 0;
  double min_step = 0, restart_scale = 0, reg = 0;
  bool fallback = false;
  // fallback_reason: 0 none, 1 nonfinite_diagnostics, 2 degraded_factorization,
  // 3 repeated_instability, 4 previous_instability (adaptive_parameters.jl:253-280, 511-548)
  int reason = 0;
};

struct FixedPolicy {  // adaptive_parameters.jl:21-32, 112-131
  double sigma, pf, df, bt, rtol;
  int rmax;
  double min_step, restart_scale;
  static FixedPolicy from(const sdpx_batch_options_t& o) {
    const double tol = o.refine_tol > 0 ? o.refine_tol : kRefineTolUlps * kEps;
    const int count = o.refine_policy == 1 ? o.refine_steps : o.refine_max_steps;
    return {o.beta, o.gamma, o.gamma, o.gamma, tol, count, o.min_step, o.omega_step};
  }
  Params params(bool fallback, int reason = 0) const {
    Params p;
    p.sigma = sigma;
    p.pf = pf;
    p.df = df;
    p.bt = bt;
    p.rtol = rtol;
    p.rmax = rmax;
    p.min_step = min_step;
    p.restart_scale = restart_scale;
    p.fallback = fallback;
    p.reason = reason;
    return p;
  }
};

struct AdaptivePolicy {  // adaptive_parameters.jl:40-53, 133-158
  FixedPolicy fb;
  double sigma_min, sigma_max, fraction_min, fraction_max, bt_min, bt_max, centrality_floor, severe_floor,
      refinement_floor, requested;
  int instability_limit = 2;
  static AdaptivePolicy from(const sdpx_batch_options_t& o) {
    AdaptivePolicy p;
    p.fb = FixedPolicy::from(o);
    const double req = std::min(o.eps_gap, std::min(o.eps_primal, o.eps_dual));
    p.refinement_floor = std::max(kRefineTolUlps * kEps, req > 0 ? req * req : 0.0);
    p.sigma_max = o.adaptive_sigma_max > 0 ? std::max(o.adaptive_sigma_max, p.fb.sigma) : std::max(0.5, p.fb.sigma);
    p.sigma_min = std::min(1.0 / 50, p.fb.sigma);
    p.fraction_min = std::min(4.0 / 5, p.fb.pf);
    p.fraction_max = std::max(99.0 / 100, p.fb.pf);
    p.bt_min = std::min(11.0 / 20, p.fb.bt);
    p.bt_max = std::max(9.0 / 10, p.fb.bt);
    p.centrality_floor = 2.0 / 25;
    p.severe_floor = 1.0 / 5;
    p.requested = req;
    return p;
  }
};

struct Diag {  // IterationDiagnostics
  int iteration = 0;
  double primal_residual = INFINITY, dual_residual = INFINITY, relative_gap = INFINITY, mu = INFINITY,
         mu_aff = INFINITY, affine_primal_step = 0, affine_dual_step = 0, previous_primal_step = 1,
         previous_dual_step = 1;
  int backtracking_count = 0;
  double regularization = 0;
  int refinement_count = 0;
  double factorization_quality = 1, achieved_residual_reduction = 1, primal_psd_margin = 1, dual_psd_margin = 1;
  bool finite() const {
    const double v[] = {primal_residual, dual_residual, mu, mu_aff, affine_primal_step, affine_dual_step,
                        factorization_quality, primal_psd_margin, dual_psd_margin};
    for (double x : v)
      if (!std::isfinite(x)) return false;
    return mu > 0 && mu_aff >= 0;
  }
};

struct Hist {
  double sigma, primal_step, dual_step;
  int backtracking_count, refinement_count;
  double primal_residual, dual_residual, achieved;
  bool unstable, fallback;
};

double clampd(double v, double lo, double hi) { return std::min(std::max(v, lo), hi); }

double adaptive_fraction(const AdaptivePolicy& p, double affine_step, double previous_step, int bcount,
                         double ratio) {  // adaptive_parameters.jl:253-280
  double cand;
  if (affine_step < 1.0 / 20)
    cand = 21.0 / 25;
  else if (affine_step < 1.0 / 5)
    cand = 9.0 / 10;
  else if (affine_step < 7.0 / 10)
    cand = 19.0 / 20;
  else
    cand = 99.0 / 100;
  if (previous_step < 1.0 / 10) cand = std::min(cand, 9.0 / 10);
  if (bcount >= 4) cand = std::min(cand, 22.0 / 25);
  if (ratio < 1e-6) cand = std::min(cand, 9.0 / 10);
  return clampd(cand, p.fraction_min, p.fraction_max);
}

struct Controller {  // AdaptiveIPMController (adaptive_parameters.jl:440-782)
  bool adaptive = true;
  double default_beta = 0.1, default_gamma = 0.9, beta = 0.1, gamma = 0.9;
  FixedPolicy fixed{};
  AdaptivePolicy ap{};
  Params parameters;
  double previous_feasibility = NAN;
  int instability_count = 0;
  bool fallback = false;
  std::vector<Hist> history;

  void init(const sdpx_batch_options_t& o) {
    adaptive = o.adaptive != 0;
    default_beta = beta = o.beta;
    default_gamma = gamma = o.gamma;
    fixed = FixedPolicy::from(o);
    ap = AdaptivePolicy::from(o);
    parameters = fixed.params(false);
    previous_feasibility = NAN;
    instability_count = 0;
    fallback = false;
    history.clear();
  }
  bool running_adaptive() const { return adaptive && !fallback; }

  Params select_adaptive(const Diag& d) const {  // adaptive_parameters.jl:188-194, 253-453
    const AdaptivePolicy& p = ap;
    if (!d.finite()) return p.fb.params(true, 1);
    if (d.factorization_quality <= 0) return p.fb.params(true, 2);
    int unstable = 0;
    for (auto it = history.rbegin(); it != history.rend() && it->unstable; ++it) ++unstable;
    if (unstable >= p.instability_limit) return p.fb.params(true, 3);
    if (std::max(d.primal_residual, std::max(d.dual_residual, d.relative_gap)) >= 1) return p.fb.params(false);
    const double ratio = clampd(d.mu_aff / d.mu, 0.0, 1.0);
    const double amin = std::min(d.affine_primal_step, d.affine_dual_step);
    double sigma_c = std::pow(ratio, amin >= 3.0 / 4 ? 3 : 2);
    double floor;
    if (amin < 1.0 / 20 || d.achieved_residual_reduction > 5.0 / 4 || d.factorization_quality < kSqrtEps)
      floor = p.severe_floor;
    else if (amin < 1.0 / 4 || d.backtracking_count >= 4 ||
             d.refinement_count >= std::max(p.fb.rmax - 1, 1))
      floor = p.centrality_floor;
    else
      floor = 0.0;
    sigma_c = std::max(sigma_c, floor);
    if (d.iteration <= 1) sigma_c = std::max(sigma_c, std::min(p.fb.sigma, p.sigma_max));
    const double prev_sigma = history.empty() ? p.fb.sigma : history.back().sigma;
    Params r;
    r.sigma = clampd(4.0 / 5 * sigma_c + 1.0 / 5 * prev_sigma, p.sigma_min, p.sigma_max);
    r.pf = adaptive_fraction(p, d.affine_primal_step, d.previous_primal_step, d.backtracking_count, ratio);
    r.df = adaptive_fraction(p, d.affine_dual_step, d.previous_dual_step, d.backtracking_count, ratio);
    double bt;
    if (d.backtracking_count >= 6 || d.achieved_residual_reduction > 5.0 / 4)
      bt = 13.0 / 20;
    else if (d.backtracking_count >= 3)
      bt = 18.0 / 25;
    else if (amin > 4.0 / 5)
      bt = 17.0 / 20;
    else
      bt = 4.0 / 5;
    r.bt = clampd(bt, p.bt_min, p.bt_max);
    r.reg = d.regularization > 0 ? std::max(d.regularization / 10, kSqrtEps) : 0.0;
    r.rtol = std::max(p.refinement_floor, p.fb.rtol);
    const bool degraded = d.factorization_quality < kSqrtEps;
    r.rmax = (degraded || d.refinement_count > 0) ? p.fb.rmax : std::min(p.fb.rmax, 2);
    const double req = p.requested > 0 ? p.requested : kSqrtEps;
    const double floor_step = std::max(100 * kEps, std::min(p.fb.min_step, req / 100));
    r.min_step = (degraded || d.backtracking_count >= 8) ? p.fb.min_step : floor_step;
    const double fmin = std::max(std::min(d.primal_residual, d.dual_residual), req);
    const double fratio = std::max(d.primal_residual, d.dual_residual) / fmin;
    double rs = fratio > 1e6 ? 1000.0 : (fratio > 1000 ? 100.0 : 10.0);
    r.restart_scale = clampd(rs, 10.0, p.fb.restart_scale);
    r.fallback = false;
    return r;
  }

  Params select(const Diag& d) {
    Params prm;
    if (fallback) {
      FixedPolicy fx{default_beta, default_gamma, default_gamma, default_gamma, parameters.rtol, parameters.rmax,
                     parameters.min_step, parameters.restart_scale};
      prm = fx.params(true, 4);
    } else if (!adaptive) {
      prm = fixed.params(false);
    } else {
      prm = select_adaptive(d);
    }
    parameters = prm;
    beta = prm.sigma;
    gamma = std::min(prm.pf, prm.df);
    if (prm.fallback) {
      fallback = true;
      beta = default_beta;
      gamma = default_gamma;
    }
    return prm;
  }

  void record(int iteration, double comp_before, double comp_after, double p_res, double d_res, double tX, double tY,
              double p_after, double d_after, int bcount, int rcount, const Params& sel) {
    const double reduction = comp_before > 0 ? comp_after / comp_before : 1.0;
    const double feas = std::max(p_res, d_res);
    const double prev = previous_feasibility;
    const double fratio = (std::isfinite(prev) && prev > 0) ? feas / prev : 1.0;
    // adaptive_parameters.jl:550-586
    int reason = 0;  // 0 none, 1 nonfinite, 2 complementarity growth, 3 feasibility growth, 4 tiny, 5 stalled
    if (!(std::isfinite(reduction) && std::isfinite(fratio) && std::isfinite(tX) && std::isfinite(tY)))
      reason = 1;
    else if (iteration > 3 && reduction > 2)
      reason = 2;
    else if (iteration > 3 && fratio > 3)
      reason = 3;
    else if (iteration > 3 && tX < 1e-6 && tY < 1e-6 && fratio >= 1)
      reason = 4;
    else if (iteration > 5 && std::min(tX, tY) < 1.0 / 50 && fratio > 49.0 / 50 && reduction > 49.0 / 50)
      reason = 5;
    const bool unstable = reason != 0;
    if (adaptive && !fallback) {
      if (reason == 1 || reason == 2 || reason == 3)
        instability_count = 2;
      else if (unstable)
        instability_count += 1;
      else
        instability_count = std::max(instability_count - 1, 0);
      if (instability_count >= 2) {
        beta = default_beta;
        gamma = default_gamma;
        fallback = true;
      }
    }
    const double before = std::max(std::max(p_res, d_res), kEps);
    history.push_back({sel.sigma, tX, tY, bcount, rcount, p_res, d_res, std::max(p_after, d_after) / before,
                       unstable, fallback || sel.fallback});
    previous_feasibility = feas;
  }

  double last_primal_step() const { return history.empty() ? 1.0 : history.back().primal_step; }
  double last_dual_step() const { return history.empty() ? 1.0 : history.back().dual_step; }
  int last_backtracking() const { return history.empty() ? 0 : history.back().backtracking_count; }
  int last_refinement() const { return history.empty() ? 0 : history.back().refinement_count; }
};

int estimate_backtracking_count(double step, double gamma, bool backtrack) {  // adaptive_parameters.jl:500-508
  if (!backtrack || !std::isfinite(step) || step <= 0 || step >= 1 || gamma <= 0 || gamma >= 1) return 0;
  return std::max(0, (int)std::lround(std::log(step) / std::log(gamma)));
}

struct Stagnation {  // stagnation.jl:56-177
  int window = 15;
  double tolerance = 1e-3;
  std::deque<double> merits;
  std::deque<bool> floors;
  int reason = 0;  // 0 none, 1 no_progress, 2 too_slow, 3 precision_floor
  void init(int w, double tol) {
    window = std::max(w, 0);
    tolerance = std::max(tol, 0.0);
    merits.clear();
    floors.clear();
    reason = 0;
  }
  bool all_floor() const {
    for (bool f : floors)
      if (!f) return false;
    return true;
  }
  bool observe(double merit, bool floor, int remaining) {
    if (window <= 0) return false;
    merits.push_back(merit);
    floors.push_back(floor);
    if ((int)merits.size() > window + 1) {
      merits.pop_front();
      floors.pop_front();
    }
    if ((int)merits.size() <= window) return false;
    const double ref = merits.front();
    if (!(std::isfinite(ref) && std::isfinite(merit) && ref > 0 && merit > 0)) return false;
    const double progress = std::log(ref) - std::log(merit);
    const double rate = progress / window;
    if (!(progress > tolerance)) {
      reason = all_floor() ? 3 : 1;
      return true;
    }
    if (merit <= 1) return false;
    if (std::log(merit) / rate > kProjectionSlack * std::max(remaining, 1)) {
      reason = all_floor() ? 3 : 2;
      return true;
    }
    return false;
  }
};

enum Status { kRunning = -1, kOptimal = 0, kIterLimit = 1, kBreakdown = 2, kStalled = 3, kInsufficient = 4,
              kMaxRestarts = 5 };

struct MemberState {
  int status = kRunning;
  int iteration = 0, restarts = 0, centering = 0;
  double pobj = 0, dobj = 0, p_res = 0, d_res = 0, gap_rel = INFINITY, comp_cur = 0, mu = 0;
  double scale_p = 1, scale_d = 1;
  Controller ctrl;
  Stagnation stag;
  double best_merit = INFINITY, initial_merit = INFINITY;
  bool has_best = false;
  double best_pobj = 0, best_dobj = 0, best_pres = 0, best_dres = 0, best_gap = 0;
  // per-iteration scratch
  bool iter_adaptive = false;  // the iteration runs the adaptive branch (strategy adaptive, no fallback)
  double o_beta = 0;
  double p_it = 0, d_it = 0;
  int reg_attempts = 0;
  double reg = 0;
  bool factor_ok = false;
  bool has_ck = false;  // differentiation checkpoint taken (first iterate with <X,Y>/k <= adjoint_mu)
  double ck_mu = 0;
  Params prm;
  double comp = 0, mu_g = 0, mu_aff = 0;
  double ap = 1, ad = 1;
  int refine_steps = 0, refine_cap = 0;
  bool refine_fixed = false, refine_active = false;
  double refine_abstol = 0, refine_residual = 0, refine_rnorm = 0;
  // mixed precision: the Schur Gram product is assembled in FP32 while `mixed` holds; a step
  // whose reduced-precision direction fails a guard is redone in FP64 and FP64 is kept
  bool mixed = false;
  int mixed_iterations = 0;  // Newton steps taken with the reduced-precision Schur matrix
  int reason = 0;            // why the member stopped (sdpx_batch_get_reasons)
  double tX = 1, tY = 1;
  // one row per Newton step: iteration, pobj, dobj, p_res, d_res (of the step), tX, tY, sigma,
  // affine steps, regularization attempts, refinement steps, outcome (0 update, 1 centering,
  // 2 restart, 3 termination)
  std::vector<std::array<double, kTraceCols>> trace;
  void log(int outcome) {
    trace.push_back({(double)iteration, pobj, dobj, p_it, d_it, tX, tY, prm.sigma, ap, ad, (double)reg_attempts,
                     (double)refine_steps, (double)outcome});
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct sdpx_batch {
  int device = 0;
  int B = 0;
  hipStream_t stream = nullptr;
  std::vector<Member> mem;
  Arrays arr{};
  Member* dmem = nullptr;
  long long nK = 0, nM = 0, nAv = 0, nPh = 0, nS = 0, nD = 0, nG = 0;
  // sparse copy of COO constraints for the gather kernels (freed by set_dense / destroy)
  long long* sp_ptr = nullptr;
  int* sp_idx = nullptr;
  double* sp_val = nullptr;
  long long* sp_pptr = nullptr;
  int* sp_pcon = nullptr;
  double* sp_pval = nullptr;
  // grouped Schur Gram product
  GroupGemm* dgram = nullptr;
  GroupGramF32* dgram32 = nullptr;  // the same products on FP32 panels (mixed precision)
  GroupTile* dgram_tiles = nullptr;
  int ngram_tiles = 0;
  std::vector<GroupTile> gram_tiles;  // member-major tile list (the autotuner's candidates reorder it)
  int gram_order = 0;                 // 0 K-descending, 1 member-major, 2 XCD bins (sdpx_batch_autotune)
  // grouped Schur Cholesky: per 64-column step, the members it touches and the
  // panel / trailing GEMM tiles
  CholDesc* dchol = nullptr;
  int* dchol_list = nullptr;
  GroupGemm* dfac = nullptr;
  GroupTile* dfac_tiles = nullptr;
  std::vector<int> step_list_off, step_list_cnt, step_panel_off, step_panel_cnt, step_trail_off, step_trail_cnt;
  // Schur panel work items
  int2* dpanel = nullptr;
  int npanel = 0;
  int panel_off[5] = {};  // panel work items of kp = 64, 48, 32, 16 (offsets)
  // member lists (ring of kListRing slots of B ints, host pinned + device)
  int* dlist = nullptr;
  int* hlist = nullptr;
  int ring_pos = 0;
  int* dall = nullptr;  // 0..B-1
  // results packed member after member (unpadded) for one device-to-host copy:
  // offsets of each member's k x k block and m-vector, and the staging buffer
  long long* dpackK = nullptr;
  long long* dpackM = nullptr;
  double* dpack = nullptr;
  long long packK = 0, packM = 0;
  // host mirrors of the scalar arrays (pinned)
  double* hsc = nullptr;
  int* hsi = nullptr;
  double* hpar = nullptr;
  double* hpar_ring = nullptr;  // kParRing pinned copies of hpar, one per copy in flight
  hipEvent_t par_done[kParRing] = {};
  bool par_busy[kParRing] = {};
  int par_pos = 0;
  std::vector<double> scale_p, scale_d;
  std::vector<MemberState> st;
  bool have_data = false, have_start = false, solved = false;
  int precision = 0;  // assembly precision of sdpx_batch_schur_at_start / autotune (sdpx_batch_set_precision)
  long long bytes = 0;
  std::vector<void*> allocs;
  // phase timing (device events around whole phases, summed over a solve)
  double t_schur = 0, t_factor = 0, t_total = 0;
  int last_iterations = 0;
};

namespace {

int balloc(sdpx_batch_t* h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_last_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return kOutOfMemory;
  }
  h->allocs.push_back(*p);
  h->bytes += (long long)bytes;
  return kOk;
}

template <typename T>
int balloc_n(sdpx_batch_t* h, T** p, long long count) {
  return balloc(h, reinterpret_cast<void**>(p), (size_t)count * sizeof(T));
}

int check_batch(sdpx_batch_t* h) {
  if (!h) {
    set_last_error("null batch handle");
    return kInvalidArgument;
  }
  SDPX_HIP(hipSetDevice(h->device));
  return kOk;
}

void clear_sparse(sdpx_batch_t* h) {
  void* bufs[] = {h->sp_ptr, h->sp_idx, h->sp_val, h->sp_pptr, h->sp_pcon, h->sp_pval};
  bool any = false;
  for (void* b : bufs) any = any || b;
  if (any) hipStreamSynchronize(h->stream);
  for (void* b : bufs)
    if (b) hipFree(b);
  h->sp_ptr = h->sp_pptr = nullptr;
  h->sp_idx = h->sp_pcon = nullptr;
  h->sp_val = h->sp_pval = nullptr;
  h->arr.sp_ptr = h->arr.sp_pptr = nullptr;
  h->arr.sp_idx = h->arr.sp_pcon = nullptr;
  h->arr.sp_val = h->arr.sp_pval = nullptr;
}

// Sparse copy of COO constraints (upper-triangle entries per constraint, in
// member order) for the gather forms of Av x and Av' w: per constraint the
// duplicates are summed in entry order (as the densified panel sums them) and
// every off-diagonal entry is stored at both mirror positions, so the
// positions of one constraint are unique.  Offsets follow the m-vector layout.
int build_sparse(sdpx_batch_t* h, const long long* ptr, const int* row, const int* col, const double* val) {
  clear_sparse(h);
  std::vector<long long> sptr((size_t)h->nM + 1, 0);
  std::vector<int> sidx;
  std::vector<double> sval;
  long long ncons = 0;
  for (int p = 0; p < h->B; ++p) ncons += h->mem[p].m;
  sidx.reserve((size_t)(2 * ptr[ncons] + 16));
  sval.reserve((size_t)(2 * ptr[ncons] + 16));
  std::vector<int> stamp(KMAX * KMAX, -1), slot(KMAX * KMAX, 0);
  std::vector<int> upos;
  std::vector<double> uval;
  long long j = 0, t = 0;
  int gen = 0;
  for (int p = 0; p < h->B; ++p) {
    const Member& M = h->mem[p];
    for (; t < M.oM; ++t) sptr[(size_t)t] = (long long)sidx.size();
    for (int i = 0; i < M.m; ++i, ++j, ++gen) {
      sptr[(size_t)(t++)] = (long long)sidx.size();
      upos.clear();
      uval.clear();
      for (long long e = ptr[j]; e < ptr[j + 1]; ++e) {
        const int pos = row[e] + col[e] * M.k;
        if (stamp[pos] == gen) {
          uval[slot[pos]] += val[e];
        } else {
          stamp[pos] = gen;
          slot[pos] = (int)upos.size();
          upos.push_back(pos);
          uval.push_back(val[e]);
        }
      }
      for (size_t u = 0; u < upos.size(); ++u) {
        const int r = upos[u] % M.k, c = upos[u] / M.k;
        sidx.push_back(upos[u]);
        sval.push_back(uval[u]);
        if (r != c) {
          sidx.push_back(c + r * M.k);
          sval.push_back(uval[u]);
        }
      }
    }
  }
  for (; t <= h->nM; ++t) sptr[(size_t)t] = (long long)sidx.size();
  // position-ordered copy: per member a counting sort of its entries by position,
  // stable in constraint order
  const size_t ne = std::max<size_t>(sidx.size(), 1);
  std::vector<long long> pptr((size_t)h->nK + 1, 0);
  std::vector<int> pcon(ne, 0);
  std::vector<double> pval(ne, 0.0);
  {
    long long u = 0;  // running position offset in the nK layout
    for (int p = 0; p < h->B; ++p) {
      const Member& M = h->mem[p];
      const int k2 = M.k * M.k;
      for (; u < M.oK; ++u) pptr[(size_t)u] = sptr[(size_t)M.oM];
      const long long e0 = sptr[(size_t)M.oM], e1 = sptr[(size_t)(M.oM + M.m)];
      std::vector<long long> cnt((size_t)k2 + 1, 0);
      for (long long e = e0; e < e1; ++e) ++cnt[(size_t)sidx[(size_t)e] + 1];
      for (int q = 0; q < k2; ++q) cnt[(size_t)q + 1] += cnt[(size_t)q];
      for (int q = 0; q < k2; ++q) pptr[(size_t)(M.oK + q)] = e0 + cnt[(size_t)q];
      u = M.oK + k2;
      for (int i = 0; i < M.m; ++i)
        for (long long e = sptr[(size_t)(M.oM + i)]; e < sptr[(size_t)(M.oM + i + 1)]; ++e) {
          const long long slotq = e0 + cnt[(size_t)sidx[(size_t)e]]++;
          pcon[(size_t)slotq] = i;
          pval[(size_t)slotq] = sval[(size_t)e];
        }
    }
    for (; u <= h->nK; ++u) pptr[(size_t)u] = (long long)sidx.size();
  }
  if (hipMalloc(&h->sp_ptr, sizeof(long long) * sptr.size()) != hipSuccess ||
      hipMalloc(&h->sp_idx, sizeof(int) * ne) != hipSuccess || hipMalloc(&h->sp_val, sizeof(double) * ne) != hipSuccess ||
      hipMalloc(&h->sp_pptr, sizeof(long long) * pptr.size()) != hipSuccess ||
      hipMalloc(&h->sp_pcon, sizeof(int) * ne) != hipSuccess ||
      hipMalloc(&h->sp_pval, sizeof(double) * ne) != hipSuccess) {
    clear_sparse(h);
    set_last_error("sdpx_batch_set_coo: hipMalloc of the sparse copy failed");
    return kOutOfMemory;
  }
  const size_t n = sidx.size();
  if (hipMemcpy(h->sp_ptr, sptr.data(), sizeof(long long) * sptr.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->sp_pptr, pptr.data(), sizeof(long long) * pptr.size(), hipMemcpyHostToDevice) != hipSuccess ||
      (n > 0 && (hipMemcpy(h->sp_idx, sidx.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
                 hipMemcpy(h->sp_val, sval.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
                 hipMemcpy(h->sp_pcon, pcon.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
                 hipMemcpy(h->sp_pval, pval.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess))) {
    clear_sparse(h);
    set_last_error("sdpx_batch_set_coo: upload of the sparse copy failed");
    return kHipError;
  }
  h->arr.sp_ptr = h->sp_ptr;
  h->arr.sp_idx = h->sp_idx;
  h->arr.sp_val = h->sp_val;
  h->arr.sp_pptr = h->sp_pptr;
  h->arr.sp_pcon = h->sp_pcon;
  h->arr.sp_pval = h->sp_pval;
  return kOk;
}

// Upload a member list and return its device copy.  Lists go to a ring of
// pinned staging slots (an asynchronous copy reads its slot when the stream
// reaches it, so a slot is reused only after the stream has been drained).
const int* put_list(sdpx_batch_t* h, const std::vector<int>& l) {
  if (h->ring_pos == kListRing) {
    hipStreamSynchronize(h->stream);
    h->ring_pos = 0;
  }
  const size_t off = (size_t)h->ring_pos * h->B;
  h->ring_pos += 1;
  if (l.empty()) return h->dlist + off;
  std::copy(l.begin(), l.end(), h->hlist + off);
  hipMemcpyAsync(h->dlist + off, h->hlist + off, sizeof(int) * l.size(), hipMemcpyHostToDevice, h->stream);
  return h->dlist + off;
}

// The host keeps editing hpar after a push, so each push stages a snapshot in the next of
// kParRing pinned buffers and copies that one: the stream is not drained (kernels launched before
// the copy read the earlier values, stream order); only a buffer whose copy is still in flight is
// waited for before it is reused.
int push_params(sdpx_batch_t* h) {
  const size_t n = (size_t)h->B * NPARAM;
  const int pos = h->par_pos;
  h->par_pos = (pos + 1) % kParRing;
  if (h->par_busy[pos]) SDPX_HIP(hipEventSynchronize(h->par_done[pos]));
  double* stage = h->hpar_ring + (size_t)pos * n;
  std::copy(h->hpar, h->hpar + n, stage);
  SDPX_HIP(hipMemcpyAsync(h->arr.par, stage, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipEventRecord(h->par_done[pos], h->stream));
  h->par_busy[pos] = true;
  return kOk;
}

int pull_scalars(sdpx_batch_t* h, bool ints = false) {
  SDPX_HIP(hipMemcpyAsync(h->hsc, h->arr.sc, sizeof(double) * (size_t)h->B * NSLOT, hipMemcpyDeviceToHost,
                          h->stream));
  if (ints)
    SDPX_HIP(hipMemcpyAsync(h->hsi, h->arr.si, sizeof(int) * (size_t)h->B * NISLOT, hipMemcpyDeviceToHost,
                            h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  h->ring_pos = 0;
  return kOk;
}

inline double& par(sdpx_batch_t* h, int p, int slot) { return h->hpar[(size_t)p * NPARAM + slot]; }
inline double sc(sdpx_batch_t* h, int p, int slot) { return h->hsc[(size_t)p * NSLOT + slot]; }
inline int si(sdpx_batch_t* h, int p, int slot) { return h->hsi[(size_t)p * NISLOT + slot]; }

// Visiting orders of the grouped Schur Gram launch (work-groups are dispatched round-robin over
// the 8 XCDs in launch order):
//   0  tiles of members with the largest K = k^2 first (stable: member order, then tile order);
//   1  member-major (the creation order);
//   2  members dealt to 8 XCD bins, largest work first (work = tiles x K, each member to the
//      lightest bin), every bin in K-descending order, bins interleaved so that launch position
//      8 j + x holds the j-th tile of bin x while every bin lasts: a member's tiles, which share
//      its P-hat panels, run on one XCD and its L2.
std::vector<GroupTile> gram_tile_order(const sdpx_batch_t* h, int mode) {
  std::vector<GroupTile> t = h->gram_tiles;
  auto K = [&](int p) { return (long long)h->mem[p].k * h->mem[p].k; };
  if (mode == 0) {
    std::stable_sort(t.begin(), t.end(), [&](const GroupTile& x, const GroupTile& y) { return K(x.g) > K(y.g); });
    return t;
  }
  if (mode == 1) return t;
  const int B = h->B;
  std::vector<long long> ntile(B, 0);
  for (const GroupTile& x : t) ntile[x.g] += 1;
  std::vector<int> order(B);
  for (int p = 0; p < B; ++p) order[p] = p;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ntile[x] * K(x) > ntile[y] * K(y); });
  std::vector<long long> load(kNumXcd, 0);
  std::vector<int> bin_of(B, 0);
  for (int p : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    bin_of[p] = x;
    load[x] += ntile[p] * K(p);
  }
  std::stable_sort(t.begin(), t.end(), [&](const GroupTile& x, const GroupTile& y) { return K(x.g) > K(y.g); });
  std::vector<std::vector<GroupTile>> bins(kNumXcd);
  for (const GroupTile& x : t) bins[bin_of[x.g]].push_back(x);
  std::vector<GroupTile> out;
  out.reserve(t.size());
  std::vector<size_t> pos(kNumXcd, 0);
  while (out.size() < t.size()) {
    for (int x = 0; x < kNumXcd; ++x) {
      int src = x;
      if (pos[src] >= bins[src].size()) {  // an exhausted bin's slot takes the fullest remaining bin's tile
        size_t best = 0;
        for (int y = 0; y < kNumXcd; ++y)
          if (bins[y].size() - pos[y] > best) {
            best = bins[y].size() - pos[y];
            src = y;
          }
        if (best == 0) break;
      }
      out.push_back(bins[src][pos[src]++]);
    }
  }
  return out;
}

int set_gram_order(sdpx_batch_t* h, int mode) {
  const std::vector<GroupTile> t = gram_tile_order(h, mode);
  SDPX_HIP(hipStreamSynchronize(h->stream));  // no launch in flight reads the old list
  SDPX_HIP(hipMemcpy(h->dgram_tiles, t.data(), sizeof(GroupTile) * t.size(), hipMemcpyHostToDevice));
  h->gram_order = mode;
  return kOk;
}

// the line-search kernels the listed members need (one per block-size class)
unsigned step_classes(const sdpx_batch_t* h, const std::vector<int>& l) {
  unsigned c = 0;
  for (int p : l) c |= step_class(h->mem[p].k);
  return c;
}

// I_SKIP = 0 for the members of l64 (FP64 assembly), I_SKIP32 = 0 for those of l32 (FP32
// assembly), 1 otherwise
int set_active(sdpx_batch_t* h, const std::vector<int>& l64, const std::vector<int>& l32) {
  SDPX_CHECK(launch_set_int(h->arr, h->dall, h->B, I_SKIP, 1, h->stream));
  SDPX_CHECK(launch_set_int(h->arr, h->dall, h->B, I_SKIP32, 1, h->stream));
  if (!l64.empty()) SDPX_CHECK(launch_set_int(h->arr, put_list(h, l64), (int)l64.size(), I_SKIP, 0, h->stream));
  if (!l32.empty()) SDPX_CHECK(launch_set_int(h->arr, put_list(h, l32), (int)l32.size(), I_SKIP32, 0, h->stream));
  return kOk;
}

// Schur assembly for the listed members: panels then the grouped Gram product, in FP32 for the
// members whose state is `mixed` and in FP64 for the others
int schur_build(sdpx_batch_t* h, const std::vector<int>& l) {
  if (l.empty()) return kOk;
  std::vector<int> l64, l32;
  for (int p : l) (h->st[p].mixed ? l32 : l64).push_back(p);
  SDPX_CHECK(set_active(h, l64, l32));
  SDPX_CHECK(launch_panel(h->arr, h->dpanel, h->panel_off, kPanelChunk, h->stream));
  if (!l64.empty())
    SDPX_CHECK(gemm_f64_grouped(h->dgram, h->dgram_tiles, h->ngram_tiles, true, false, true, h->stream));
  if (!l32.empty()) SDPX_CHECK(gram_f32_grouped(h->dgram32, h->dgram_tiles, h->ngram_tiles, h->stream));
  return kOk;
}

// rho = r - S dx with S applied matrix-free in FP64, (S dx)_i = <A_i, X^-1 (sum_j dx_j A_j) Y>;
// S_RHO = ||rho||_inf, S_RNORM = ||r||_inf (dX and W2 are scratch)
int exact_residual(sdpx_batch_t* h, const std::vector<int>& l) {
  if (l.empty()) return kOk;
  const int* dl = put_list(h, l);
  const int n = (int)l.size();
  SDPX_CHECK(launch_gemv(h->arr, dl, n, 3, h->stream));
  SDPX_CHECK(launch_exact_block(h->arr, dl, n, h->stream));
  SDPX_CHECK(launch_gemvt(h->arr, dl, n, 3, h->stream));
  return kOk;
}

// Refinement of dx against the matrix-free FP64 operator (the mixed-precision direction's
// _refine_mixed_direction!, mixed_precision_kkt.jl:1634-1713): delta = Sf^-1 rho with the factor of
// the reduced-precision Schur matrix, dx += delta, until ||rho||_inf <= rel * max(||r||_inf, 1) or
// `cap` steps; a step that raises the residual is rolled back and ends the member's refinement.
// Sets refine_residual, refine_rnorm, refine_abstol and refine_steps of the listed members.
int refine_exact(sdpx_batch_t* h, const std::vector<int>& l, double rel, int cap) {
  if (l.empty()) return kOk;
  SDPX_CHECK(exact_residual(h, l));
  SDPX_CHECK(pull_scalars(h));
  for (int p : l) {
    MemberState& s = h->st[p];
    s.refine_residual = sc(h, p, S_RHO);
    s.refine_rnorm = sc(h, p, S_RNORM);
    s.refine_abstol = rel * std::max(s.refine_rnorm, 1.0);
    s.refine_steps = 0;
    s.refine_active = !(s.refine_residual <= s.refine_abstol);
  }
  for (int it = 0; it < cap; ++it) {
    std::vector<int> act;
    for (int p : l)
      if (h->st[p].refine_active) act.push_back(p);
    if (act.empty()) break;
    const int* dl = put_list(h, act);
    const int na = (int)act.size();
    SDPX_CHECK(launch_vec_op(h->arr, dl, na, 0, h->stream));
    SDPX_CHECK(launch_solve(h->arr, dl, na, 1, h->stream));
    SDPX_CHECK(launch_vec_op(h->arr, dl, na, 1, h->stream));
    SDPX_CHECK(exact_residual(h, act));
    SDPX_CHECK(pull_scalars(h));
    std::vector<int> back;
    for (int p : act) {
      MemberState& s = h->st[p];
      const double corrected = sc(h, p, S_RHO);
      if (!std::isfinite(corrected) || corrected > s.refine_residual) {
        back.push_back(p);
        s.refine_active = false;
        continue;
      }
      const double contraction = corrected / s.refine_residual;
      s.refine_steps += 1;
      s.refine_residual = corrected;
      if (corrected <= s.refine_abstol) {
        s.refine_active = false;
      } else if (contraction > 0.0) {
        // the refinement budget of _predicted_mixed_refinement_steps (mixed_precision_kkt.jl:394-415)
        // with the measured contraction in place of cond * eps * 8: a member that cannot reach
        // its target within `cap` steps stops now (its step is redone in FP64) instead of
        // holding every member's refinement rounds open until the cap
        const double more = std::ceil(std::log(s.refine_abstol / corrected) / std::log(contraction));
        if (!(s.refine_steps + more <= cap)) s.refine_active = false;
      }
    }
    if (!back.empty()) {
      SDPX_CHECK(launch_vec_op(h->arr, put_list(h, back), (int)back.size(), 2, h->stream));
      SDPX_CHECK(exact_residual(h, back));
    }
  }
  return kOk;
}

// One grouped Cholesky of Sf for the listed members (others skipped through info).
int factor_once(sdpx_batch_t* h, const std::vector<int>& l) {
  SDPX_CHECK(launch_set_int(h->arr, h->dall, h->B, I_INFOS, -1, h->stream));
  SDPX_CHECK(push_params(h));
  SDPX_CHECK(launch_shift_copy(h->arr, put_list(h, l), (int)l.size(), h->stream));
  for (size_t j = 0; j < h->step_list_cnt.size(); ++j) {
    SDPX_CHECK(potrf_diag_grouped(h->dchol, h->dchol_list + h->step_list_off[j], h->step_list_cnt[j],
                                  (int)j * kCholNb, h->stream));
    SDPX_CHECK(gemm_f64_grouped(h->dfac, h->dfac_tiles + h->step_panel_off[j], h->step_panel_cnt[j], false, true,
                                false, h->stream));
    SDPX_CHECK(gemm_f64_grouped(h->dfac, h->dfac_tiles + h->step_trail_off[j], h->step_trail_cnt[j], false, true,
                                false, h->stream));
  }
  return kOk;
}

// factorize! with the sqrt(eps) 10^(a-1) ladder (kkt.jl:1430-1446) per member
int schur_factor(sdpx_batch_t* h, const std::vector<int>& l) {
  std::vector<int> todo = l;
  for (int p : l) {
    h->st[p].reg_attempts = 0;
    h->st[p].reg = 0.0;
    h->st[p].factor_ok = false;
    par(h, p, P_REG) = 0.0;
  }
  while (!todo.empty()) {
    SDPX_CHECK(factor_once(h, todo));
    SDPX_CHECK(pull_scalars(h, true));
    std::vector<int> again;
    for (int p : todo) {
      MemberState& s = h->st[p];
      if (si(h, p, I_INFOS) == 0) {
        s.factor_ok = true;
        continue;
      }
      if (s.reg_attempts >= 6) continue;  // gives up: reported by the caller
      s.reg_attempts += 1;
      s.reg = s.reg_attempts == 1 ? kSqrtEps : s.reg * 10.0;
      par(h, p, P_REG) = s.reg;
      again.push_back(p);
    }
    todo.swap(again);
  }
  return kOk;
}

double gap_of(double pobj, double dobj) { return std::fabs(pobj - dobj) / std::max(1.0, (std::fabs(pobj) + std::fabs(dobj)) / 2); }

double merit_of(const sdpx_batch_options_t& o, double p_res, double d_res, double gap, double comp, double sp,
                double sd, double objective_scale) {  // stagnation.jl merit
  auto safe = [](double e) { return e > 0 ? e : kEps; };
  double m = std::max((p_res / sp) / safe(o.eps_primal), std::max((d_res / sd) / safe(o.eps_dual), gap / safe(o.eps_gap)));
  if (std::isfinite(comp) && comp > 0) m = std::max(m, comp / (safe(o.eps_gap) * std::max(1.0, objective_scale)));
  return m;
}

bool at_floor(double p_res, double d_res, double gap, double sp, double sd) {
  const double f = kPrecisionFloorUlps * kEps;
  return p_res / sp <= f && d_res / sd <= f && gap <= f;
}

// The Newton direction of the listed members (newton_step!, step.jl:643-1006, from the Schur
// build to dX, dY): members that end here in breakdown get their status, members whose
// reduced-precision step failed a guard move to `retry` (the caller redoes them in FP64), the
// others stay in `ok` with their direction and step parameters.
int newton_direction(sdpx_batch_t* h, const sdpx_batch_options_t& o, std::vector<int>& ok, std::vector<int>& retry) {
  auto& S = h->st;
  SDPX_CHECK(schur_build(h, ok));
  SDPX_CHECK(schur_factor(h, ok));
  // members whose reduced-precision step is abandoned move from ok to retry
  auto to_retry = [&](const std::vector<int>& bad) {
    if (bad.empty()) return;
    std::vector<int> keep;
    for (int p : ok)
      if (std::find(bad.begin(), bad.end(), p) == bad.end()) keep.push_back(p);
    ok.swap(keep);
    retry.insert(retry.end(), bad.begin(), bad.end());
  };
  {
    std::vector<int> keep, bad;
    for (int p : ok) {
      if (!S[p].factor_ok) {
        if (S[p].mixed) {
          bad.push_back(p);  // reduced-precision Schur matrix not positive definite
          continue;
        }
        S[p].status = kBreakdown;
        S[p].reason = 3;
        S[p].p_res = S[p].p_it;
        S[p].d_res = S[p].d_it;
      } else {
        keep.push_back(p);
      }
    }
    ok.swap(keep);
    retry.insert(retry.end(), bad.begin(), bad.end());
  }
  if (ok.empty()) return kOk;
  std::vector<int> mixed;
  for (int p : ok)
    if (S[p].mixed) mixed.push_back(p);
  if (!mixed.empty() && o.cond_limit > 0) {
    // condition guard: the pivots d_i = L_ii^2 of a Cholesky factor lie in [lambda_min, lambda_max],
    // so (max L_ii / min L_ii)^2 bounds cond_2(S) from below; a member above the limit cannot
    // keep the reduced-precision Schur matrix (the refinement guards below catch the rest)
    SDPX_CHECK(launch_factor_extrema(h->arr, put_list(h, mixed), (int)mixed.size(), h->stream));
    SDPX_CHECK(pull_scalars(h));
    std::vector<int> bad, keep;
    for (int p : mixed) {
      const double mn = sc(h, p, S_SFMIN), mx = sc(h, p, S_SFMAX);
      const double est = (mx / mn) * (mx / mn);
      if (!(est <= o.cond_limit)) bad.push_back(p);
      else keep.push_back(p);
    }
    mixed.swap(keep);
    to_retry(bad);
    if (ok.empty()) return kOk;
  }
  const int mixed_cap = o.mixed_max_refine;
  // predictor (step.jl:699-726)
  {
    const int* l = put_list(h, ok);
    const int n0 = (int)ok.size();
    SDPX_CHECK(launch_rhs_block(h->arr, l, n0, h->stream));
    SDPX_CHECK(launch_gemvt(h->arr, l, n0, 1, h->stream));
    SDPX_CHECK(launch_solve(h->arr, l, n0, 0, h->stream));
    if (!mixed.empty()) {
      // predictor guard (mixed_precision_kkt.jl:1475-1586): residual <= 1e-8 max(||r||, 1)
      SDPX_CHECK(refine_exact(h, mixed, 1e-8, mixed_cap));
      std::vector<int> bad, keep;
      for (int p : mixed) (S[p].refine_residual <= S[p].refine_abstol ? keep : bad).push_back(p);
      mixed.swap(keep);
      to_retry(bad);
      if (ok.empty()) return kOk;
    }
    l = put_list(h, ok);
    const int n1 = (int)ok.size();
    SDPX_CHECK(launch_gemv(h->arr, l, n1, 1, h->stream));
    SDPX_CHECK(launch_dir_block(h->arr, l, n1, h->stream));
  }
  const int n = (int)ok.size();
  std::vector<int> fp64;
  for (int p : ok)
    if (!S[p].mixed) fp64.push_back(p);
  // affine predictor diagnostics (step.jl:546-625): the fraction to the boundary of the affine
  // direction, then the complementarity at those steps (read by the comp pass from the
  // line-search slots: no read-back in between) and the factor's diagonal extrema
  std::vector<int> adapt;
  for (int p : ok) {
    if (S[p].iter_adaptive) {
      adapt.push_back(p);
      par(h, p, P_PF) = 1.0;
      par(h, p, P_DF) = 1.0;
      par(h, p, P_A) = -1.0;
      par(h, p, P_B) = -1.0;
    } else {
      par(h, p, P_A) = 1.0;
      par(h, p, P_B) = 1.0;
    }
  }
  SDPX_CHECK(push_params(h));
  if (!adapt.empty())
    SDPX_CHECK(launch_ftb(h->arr, put_list(h, adapt), (int)adapt.size(), step_classes(h, adapt), h->stream));
  {
    const int* l = put_list(h, ok);
    SDPX_CHECK(launch_comp(h->arr, l, n, h->stream));
    SDPX_CHECK(launch_factor_extrema(h->arr, l, n, h->stream));
    SDPX_CHECK(pull_scalars(h));
  }
  for (int p : ok) {
    MemberState& s = S[p];
    if (s.iter_adaptive) {
      s.ap = sc(h, p, S_TX);
      s.ad = sc(h, p, S_TY);
    } else {
      s.ap = s.ad = 1.0;
    }
    par(h, p, P_A) = s.ap;
    par(h, p, P_B) = s.ad;
  }
  auto quality = [](double mn, double mx) {
    return (mx > 0.0 && std::isfinite(mx)) ? std::min(std::max(mn / mx, 0.0), 1.0) : 0.0;
  };
  for (int p : ok) {
    MemberState& s = S[p];
    const int k = h->mem[p].k;
    const double comp = sc(h, p, S_COMP), compt = sc(h, p, S_COMPT);
    const bool uniform = s.iter_adaptive && std::isfinite(comp);
    s.comp = comp;
    s.mu_g = comp / k;
    s.mu_aff = std::max(0.0, compt / k);
    const double cur = std::max(s.p_it / std::max(s.scale_p, 1.0), s.d_it / std::max(s.scale_d, 1.0));
    double achieved = 1.0;
    if (!s.ctrl.history.empty()) {
      const Hist& row = s.ctrl.history.back();
      const double prev = std::max(std::max(row.primal_residual / std::max(s.scale_p, 1.0),
                                            row.dual_residual / std::max(s.scale_d, 1.0)),
                                   kEps);
      achieved = cur / prev;
    }
    Diag d;
    d.iteration = s.iteration + 1;
    d.primal_residual = s.p_it / std::max(s.scale_p, 1.0);
    d.dual_residual = s.d_it / std::max(s.scale_d, 1.0);
    d.relative_gap = s.gap_rel;
    d.mu = s.mu_g;
    d.mu_aff = s.mu_aff;
    d.affine_primal_step = s.ap;
    d.affine_dual_step = s.ad;
    d.previous_primal_step = s.ctrl.last_primal_step();
    d.previous_dual_step = s.ctrl.last_dual_step();
    d.backtracking_count = s.ctrl.last_backtracking();
    d.regularization = s.reg_attempts <= 0 ? 0.0 : kSqrtEps * std::pow(10.0, s.reg_attempts - 1);
    d.refinement_count = s.ctrl.last_refinement();
    d.factorization_quality = quality(sc(h, p, S_SFMIN), sc(h, p, S_SFMAX));
    d.achieved_residual_reduction = achieved;
    d.primal_psd_margin = std::min(1.0, quality(sc(h, p, S_LXMIN), sc(h, p, S_LXMAX)));
    d.dual_psd_margin = std::min(1.0, quality(sc(h, p, S_LYMIN), sc(h, p, S_LYMAX)));
    s.prm = s.ctrl.select(d);
    // corrector target (threaded.jl:839-854, 943-1003)
    if (s.iter_adaptive && !s.prm.fallback) {
      par(h, p, P_MODE) = uniform ? 2.0 : 1.0;
      par(h, p, P_SIGMA) = s.prm.sigma;
      par(h, p, P_MUG) = s.mu_g;
    } else {
      par(h, p, P_MODE) = 0.0;
      par(h, p, P_MU) = s.mu;
      par(h, p, P_BETA) = s.o_beta;
    }
    // refinement policy
    s.refine_fixed = o.refine_policy == 1;
    double rtol;
    if (s.iter_adaptive && !s.prm.fallback) {
      rtol = s.prm.rtol;
      s.refine_cap = s.prm.rmax;
    } else {
      rtol = o.refine_tol;
      s.refine_cap = o.refine_max_steps;
    }
    if (s.refine_fixed) s.refine_cap = o.refine_steps;
    s.refine_abstol = rtol > 0 ? rtol : kRefineTolUlps * kEps;  // scaled by ||r|| below
  }
  SDPX_CHECK(push_params(h));
  // corrector (step.jl:846-936)
  {
    const int* l = put_list(h, ok);
    SDPX_CHECK(launch_corrector_block(h->arr, l, n, h->stream));
    SDPX_CHECK(launch_rhs_block(h->arr, l, n, h->stream));
    SDPX_CHECK(launch_gemvt(h->arr, l, n, 1, h->stream));
    SDPX_CHECK(launch_solve(h->arr, l, n, 0, h->stream));
    if (!fp64.empty()) SDPX_CHECK(launch_residual(h->arr, put_list(h, fp64), (int)fp64.size(), h->stream));
    SDPX_CHECK(pull_scalars(h));
  }
  if (!mixed.empty()) {
    // _refine_mixed_direction! (mixed_precision_kkt.jl:1634-1713) to the FP64 target
    SDPX_CHECK(refine_exact(h, mixed, o.refine_target > 0 ? o.refine_target : 1e-10, mixed_cap));
    std::vector<int> bad, keep;
    for (int p : mixed) {
      MemberState& s = S[p];
      if (s.refine_residual <= s.refine_abstol) {
        s.refine_steps = std::min(s.refine_steps, 1);  // the controller sees a native-equivalent count
        keep.push_back(p);
      } else {
        bad.push_back(p);
      }
    }
    mixed.swap(keep);
    to_retry(bad);
    if (ok.empty()) return kOk;
  }
  // refine! (kkt_backend.jl:476-485 -> kkt.jl:3891-3963), member by member in rounds
  int max_cap = 0;
  for (int p : fp64) {
    MemberState& s = S[p];
    s.refine_steps = 0;
    s.refine_residual = sc(h, p, S_RHO);
    s.refine_abstol = s.refine_fixed ? 0.0 : s.refine_abstol * std::max(sc(h, p, S_RNORM), 1.0);
    s.refine_active = s.refine_cap > 0;
    max_cap = std::max(max_cap, s.refine_cap);
  }
  bool stale = false;  // a roll-back in the last round recomputed S_RHO after its read-back
  for (int it = 0; it < max_cap; ++it) {
    std::vector<int> act;
    for (int p : fp64) {
      MemberState& s = S[p];
      if (!s.refine_active) continue;
      if (it >= s.refine_cap || (!s.refine_fixed && s.refine_residual <= s.refine_abstol)) {
        s.refine_active = false;
        continue;
      }
      act.push_back(p);
    }
    if (act.empty()) break;
    const int* l = put_list(h, act);
    const int na = (int)act.size();
    SDPX_CHECK(launch_vec_op(h->arr, l, na, 0, h->stream));
    SDPX_CHECK(launch_solve(h->arr, l, na, 1, h->stream));
    SDPX_CHECK(launch_vec_op(h->arr, l, na, 1, h->stream));
    SDPX_CHECK(launch_residual(h->arr, l, na, h->stream));
    SDPX_CHECK(pull_scalars(h));
    std::vector<int> back;
    for (int p : act) {
      MemberState& s = S[p];
      const double corrected = sc(h, p, S_RHO);
      if (!std::isfinite(corrected) || corrected > s.refine_residual) {
        back.push_back(p);
        s.refine_active = false;
        continue;
      }
      s.refine_steps += 1;
      const bool small = corrected > s.refine_residual * kRefineMinDecrease;
      s.refine_residual = corrected;
      if (!s.refine_fixed && small) s.refine_active = false;
    }
    stale = !back.empty();
    if (stale) {
      const int* lb = put_list(h, back);
      SDPX_CHECK(launch_vec_op(h->arr, lb, (int)back.size(), 2, h->stream));
      SDPX_CHECK(launch_residual(h->arr, lb, (int)back.size(), h->stream));
    }
  }
  // S_RHO / S_RNORM on the host are the residuals of every member's final dx: from the corrector's
  // pass, or from the last refinement round it took part in, unless that round rolled it back
  if (stale) SDPX_CHECK(pull_scalars(h));
  {
    std::vector<int> keep;
    for (int p : ok) {
      MemberState& s = S[p];
      const double res = s.mixed ? s.refine_residual : sc(h, p, S_RHO);
      const double r_norm = s.mixed ? s.refine_rnorm : sc(h, p, S_RNORM);
      const double tol = std::max(kSqrtEps, std::max(o.eps_primal, o.eps_dual)) * std::max(r_norm, 1.0);
      if (!std::isfinite(res) || res > tol) {
        s.status = kBreakdown;
        s.reason = 4;
        s.p_res = s.p_it;
        s.d_res = s.d_it;
      } else {
        keep.push_back(p);
      }
    }
    ok.swap(keep);
  }
  if (ok.empty()) return kOk;
  const int n2 = (int)ok.size();
  {
    const int* l = put_list(h, ok);
    SDPX_CHECK(launch_gemv(h->arr, l, n2, 1, h->stream));
    SDPX_CHECK(launch_dir_block(h->arr, l, n2, h->stream));
  }
  return kOk;
}

int run_solve(sdpx_batch_t* h, const sdpx_batch_options_t& o, int* iterations_out) {
  const int B = h->B;
  std::vector<int> all(B);
  for (int p = 0; p < B; ++p) all[p] = p;
  auto& S = h->st;
  // ---- start (interior_point.jl:1528-1650 warm start, ipm loop set-up) ------------------
  for (int p = 0; p < B; ++p) {
    S[p] = MemberState();
    S[p].ctrl.init(o);
    S[p].stag.init(o.stall_iterations, o.stall_tolerance);
    S[p].scale_p = h->scale_p[p];
    S[p].scale_d = h->scale_d[p];
    S[p].mixed = o.precision != 0;
    par(h, p, P_A) = 0.0;
    par(h, p, P_B) = 0.0;
  }
  SDPX_CHECK(push_params(h));
  const int* dl = put_list(h, all);
  SDPX_CHECK(launch_comp(h->arr, dl, B, h->stream));
  SDPX_CHECK(launch_objectives(h->arr, dl, B, h->stream));
  SDPX_CHECK(launch_gemv(h->arr, dl, B, 0, h->stream));
  SDPX_CHECK(launch_gemvt(h->arr, dl, B, 0, h->stream));
  SDPX_CHECK(pull_scalars(h));
  for (int p = 0; p < B; ++p) {
    MemberState& s = S[p];
    const int k = h->mem[p].k;
    s.comp_cur = sc(h, p, S_COMP);
    s.mu = o.beta * s.comp_cur / k;
    s.pobj = sc(h, p, S_POBJ);
    s.dobj = sc(h, p, S_DOBJ);
    s.p_res = sc(h, p, S_PRES);
    s.d_res = sc(h, p, S_DRES);
  }
  int launched = 0;
  while (true) {
    // ---- loop top: termination, stagnation, best iterate --------------------------------------
    std::vector<int> run;
    for (int p = 0; p < B; ++p)
      if (S[p].status == kRunning) run.push_back(p);
    if (run.empty()) break;
    std::vector<int> need_sol;
    for (int p : run) {
      MemberState& s = S[p];
      s.gap_rel = gap_of(s.pobj, s.dobj);
      const bool term = s.p_res / s.scale_p <= o.eps_primal && s.d_res / s.scale_d <= o.eps_dual;
      if (term && s.gap_rel <= o.eps_gap) need_sol.push_back(p);
    }
    if (!need_sol.empty()) {
      const int* l = put_list(h, need_sol);
      SDPX_CHECK(launch_gemv(h->arr, l, (int)need_sol.size(), 2, h->stream));
      SDPX_CHECK(launch_gemvt(h->arr, l, (int)need_sol.size(), 2, h->stream));
      SDPX_CHECK(pull_scalars(h));
      for (int p : need_sol) {
        S[p].p_res = sc(h, p, S_PSOL);
        S[p].d_res = sc(h, p, S_DSOL);
      }
    }
    std::vector<int> save_best, step;
    for (int p : run) {
      MemberState& s = S[p];
      const bool term_ok = s.p_res / s.scale_p <= o.eps_primal && s.d_res / s.scale_d <= o.eps_dual;
      const bool gap_ok = s.gap_rel <= o.eps_gap;
      const double merit = merit_of(o, s.p_res, s.d_res, s.gap_rel, s.comp_cur, s.scale_p, s.scale_d,
                                    std::max(std::fabs(s.pobj), std::fabs(s.dobj)));
      const bool floor = at_floor(s.p_res, s.d_res, s.gap_rel, s.scale_p, s.scale_d);
      const bool stagnated = s.stag.observe(merit, floor, o.iter_max - s.iteration);
      if (std::isfinite(merit) && !std::isfinite(s.initial_merit)) s.initial_merit = merit;
      if (std::isfinite(merit) && merit < s.best_merit) {
        s.best_merit = merit;
        if (!(term_ok && gap_ok)) {
          s.has_best = true;
          s.best_pobj = s.pobj;
          s.best_dobj = s.dobj;
          s.best_pres = s.p_res;
          s.best_dres = s.d_res;
          s.best_gap = s.gap_rel;
          save_best.push_back(p);
        }
      }
      if (term_ok && gap_ok) {
        s.status = kOptimal;
        s.reason = 1;
      } else if (s.iteration >= o.iter_max) {
        s.status = kIterLimit;
        s.reason = 9;
      } else if (stagnated) {
        s.status = s.stag.reason == 3 ? kInsufficient : kStalled;
        s.reason = 10 + s.stag.reason;
      } else {
        step.push_back(p);
      }
    }
    if (!save_best.empty()) SDPX_CHECK(launch_best(h->arr, put_list(h, save_best), (int)save_best.size(), 0, h->stream));
    if (step.empty()) continue;
    launched += 1;
    // ---- newton_step! (step.jl:643-1006) ------------------------------------------------------
    for (int p : step) {
      MemberState& s = S[p];
      // controller_options (adaptive_parameters.jl:771-782)
      s.iter_adaptive = s.ctrl.running_adaptive();
      s.o_beta = s.ctrl.beta;
      par(h, p, P_MU) = s.mu;
      par(h, p, P_AFF) = s.iter_adaptive ? 1.0 : 0.0;
    }
    SDPX_CHECK(push_params(h));
    {
      const int* l = put_list(h, step);
      const int n = (int)step.size();
      SDPX_CHECK(launch_gemv(h->arr, l, n, 0, h->stream));
      SDPX_CHECK(launch_gemvt(h->arr, l, n, 0, h->stream));
      SDPX_CHECK(launch_block_factor(h->arr, l, n, h->stream));
      SDPX_CHECK(pull_scalars(h, true));
    }
    std::vector<int> ok;
    for (int p : step) {
      MemberState& s = S[p];
      s.p_it = sc(h, p, S_PRES);
      s.d_it = sc(h, p, S_DRES);
      if (si(h, p, I_INFOX) != 0 || si(h, p, I_INFOY) != 0) {
        s.status = kBreakdown;
        s.reason = 2;
        s.p_res = s.p_it;
        s.d_res = s.d_it;
      } else {
        ok.push_back(p);
      }
    }
    {
      std::vector<int> retry;
      SDPX_CHECK(newton_direction(h, o, ok, retry));
      if (!retry.empty()) {
        // the reduced-precision direction failed a guard: redo the step in FP64 and stay there
        // (R = -X Y (+ mu I) and the block factors are recomputed, bitwise as before)
        for (int p : retry) S[p].mixed = false;
        SDPX_CHECK(push_params(h));
        SDPX_CHECK(launch_block_factor(h->arr, put_list(h, retry), (int)retry.size(), h->stream));
        std::vector<int> none;
        SDPX_CHECK(newton_direction(h, o, retry, none));
        ok.insert(ok.end(), retry.begin(), retry.end());
        std::sort(ok.begin(), ok.end());
      }
    }
    if (ok.empty()) continue;
    for (int p : ok)
      if (S[p].mixed) S[p].mixed_iterations += 1;
    // ---- line search (interior_point.jl:1962; threaded.jl:1589-1600) -------------------------------
    std::vector<int> ftb, bt;
    for (int p : ok) {
      MemberState& s = S[p];
      s.p_res = s.p_it;
      s.d_res = s.d_it;
      const bool use_ftb = o.step_rule == 2 || (o.step_rule == 0 && h->mem[p].k <= 2);
      if (use_ftb) {
        ftb.push_back(p);
        par(h, p, P_PF) = s.prm.pf;
        par(h, p, P_DF) = s.prm.df;
      } else {
        bt.push_back(p);
        par(h, p, P_GAMMA) = s.prm.bt;
        par(h, p, P_MINSTEP) = s.prm.min_step;
        par(h, p, P_RATIO) = o.adaptive ? kSqrtEps : 0.0;
      }
    }
    SDPX_CHECK(push_params(h));
    if (!ftb.empty()) SDPX_CHECK(launch_ftb(h->arr, put_list(h, ftb), (int)ftb.size(), step_classes(h, ftb), h->stream));
    if (!bt.empty())
      SDPX_CHECK(launch_backtrack(h->arr, put_list(h, bt), (int)bt.size(), step_classes(h, bt), h->stream));
    SDPX_CHECK(pull_scalars(h));
    // collapse handling (interior_point.jl:2010-2217)
    std::vector<int> upd, center, restart;
    for (int p : ok) {
      MemberState& s = S[p];
      const bool backtrack = !(o.step_rule == 2 || (o.step_rule == 0 && h->mem[p].k <= 2));
      double tX = sc(h, p, S_TX), tY = sc(h, p, S_TY);
      const int bcount = estimate_backtracking_count(tX, s.prm.bt, backtrack) +
                         estimate_backtracking_count(tY, s.prm.bt, backtrack);
      const bool pf_ok = s.p_res / s.scale_p <= o.eps_primal;
      const bool df_ok = s.d_res / s.scale_d <= o.eps_dual;
      bool x_stuck = tX < s.prm.min_step, y_stuck = tY < s.prm.min_step;
      if (x_stuck && pf_ok && !y_stuck) {
        tX = 0.0;
        x_stuck = false;
      }
      if (y_stuck && df_ok && !x_stuck) {
        tY = 0.0;
        y_stuck = false;
      }
      s.tX = tX;
      s.tY = tY;
      s.refine_cap = bcount;  // reused below as the backtracking count of this iteration
      if (x_stuck || y_stuck) {
        const double near = 1000.0;
        if (s.gap_rel < near * o.eps_gap && s.p_res / s.scale_p < near * o.eps_primal &&
            s.d_res / s.scale_d < near * o.eps_dual) {
          s.status = kStalled;
          s.reason = 5;
          s.log(3);
          continue;
        }
        if (o.stall_iterations > 0 && std::isfinite(s.initial_merit) && s.best_merit < s.initial_merit * 1e-3 &&
            s.best_merit < kNearSolutionMerit) {
          s.status = kInsufficient;
          s.reason = 6;
          s.log(3);
          continue;
        }
        if (s.centering < o.max_centering && s.ctrl.beta < kCenteringBetaMax) {
          s.centering += 1;
          s.ctrl.beta = std::min(s.ctrl.beta * kCenteringBetaStep, kCenteringBetaMax);
          center.push_back(p);
          s.log(1);
          continue;
        }
        const bool rescale_X = x_stuck && !pf_ok;
        const bool rescale_Y = y_stuck && !df_ok;
        if (o.restart && s.restarts < o.max_restarts && (rescale_X || rescale_Y)) {
          s.restarts += 1;
          par(h, p, P_SCALE) = s.prm.restart_scale;
          par(h, p, P_A) = rescale_X ? 1.0 : 0.0;  // which blocks to rescale (host bookkeeping only)
          par(h, p, P_B) = rescale_Y ? 1.0 : 0.0;
          restart.push_back(p);
          s.log(2);
          continue;
        }
        if (o.restart && s.restarts < o.max_restarts)
          s.status = kStalled;
        else
          s.status = s.restarts >= o.max_restarts ? kMaxRestarts : kBreakdown;
        s.reason = 7;
        s.log(3);
        continue;
      }
      par(h, p, P_A) = tX;
      par(h, p, P_B) = tY;
      upd.push_back(p);
    }
    if (!restart.empty()) {
      std::vector<int> rx, ry;
      for (int p : restart) {
        if (par(h, p, P_A) != 0.0) rx.push_back(p);
        if (par(h, p, P_B) != 0.0) ry.push_back(p);
      }
      SDPX_CHECK(push_params(h));
      if (!rx.empty()) SDPX_CHECK(launch_scale(h->arr, put_list(h, rx), (int)rx.size(), 1, h->stream));
      if (!ry.empty()) SDPX_CHECK(launch_scale(h->arr, put_list(h, ry), (int)ry.size(), 2, h->stream));
    }
    std::vector<int> recomp = center;
    recomp.insert(recomp.end(), restart.begin(), restart.end());
    for (int p : recomp) {
      par(h, p, P_A) = 0.0;
      par(h, p, P_B) = 0.0;
    }
    SDPX_CHECK(push_params(h));
    if (!recomp.empty()) SDPX_CHECK(launch_comp(h->arr, put_list(h, recomp), (int)recomp.size(), h->stream));
    // ---- accepted-step update ----------------------------------------------------------------------
    if (!upd.empty()) {
      const int* l = put_list(h, upd);
      const int nu = (int)upd.size();
      SDPX_CHECK(launch_update(h->arr, l, nu, h->stream));
      SDPX_CHECK(launch_accepted(h->arr, l, nu, h->stream));
      SDPX_CHECK(launch_objectives(h->arr, l, nu, h->stream));
    }
    SDPX_CHECK(pull_scalars(h));
    for (int p : recomp) {
      MemberState& s = S[p];
      const double comp = sc(h, p, S_COMP);
      if (std::find(restart.begin(), restart.end(), p) != restart.end()) s.comp_cur = comp;
      s.mu = s.ctrl.beta * comp / h->mem[p].k;
    }
    std::vector<int> ck;
    for (int p : upd) {
      MemberState& s = S[p];
      const double comp_after = sc(h, p, S_COMP);
      if (o.adjoint_mu > 0 && !s.has_ck && comp_after / h->mem[p].k <= o.adjoint_mu) {
        s.has_ck = true;
        s.ck_mu = comp_after / h->mem[p].k;
        ck.push_back(p);
      }
      const bool finite = std::isfinite(sc(h, p, S_XMAX)) && std::isfinite(sc(h, p, S_XBMAX)) &&
                          std::isfinite(sc(h, p, S_YBMAX));
      if (!(finite && std::isfinite(s.mu))) {
        s.status = kBreakdown;
        s.reason = 8;
        continue;
      }
      const double p_after = std::fabs(1.0 - s.tX) * s.p_it;
      const double d_after = sc(h, p, S_DACC);
      s.ctrl.record(s.iteration + 1, s.comp, comp_after, s.p_it, s.d_it, s.tX, s.tY, p_after, d_after, s.refine_cap,
                    s.refine_steps, s.prm);
      s.comp_cur = comp_after;
      s.mu = s.ctrl.beta * comp_after / h->mem[p].k;
      s.pobj = sc(h, p, S_POBJ);
      s.dobj = sc(h, p, S_DOBJ);
      s.iteration += 1;
      s.p_res = p_after;
      s.d_res = d_after;
      s.log(0);
    }
    if (!ck.empty()) SDPX_CHECK(launch_best(h->arr, put_list(h, ck), (int)ck.size(), 2, h->stream));
  }
  // ---- results: the best iterate replaces a non-optimal final one ---------------------------------
  std::vector<int> restore;
  for (int p = 0; p < B; ++p) {
    MemberState& s = S[p];
    if (s.status != kOptimal && s.has_best) {
      restore.push_back(p);
      s.pobj = s.best_pobj;
      s.dobj = s.best_dobj;
      s.p_res = s.best_pres;
      s.d_res = s.best_dres;
      s.gap_rel = s.best_gap;
    }
  }
  if (!restore.empty()) SDPX_CHECK(launch_best(h->arr, put_list(h, restore), (int)restore.size(), 1, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  h->last_iterations = launched;
  if (iterations_out) *iterations_out = launched;
  return kOk;
}

}  // namespace

extern "C" {

int sdpx_batch_default_options(sdpx_batch_options_t* o, double tol) {
  if (!o) {
    set_last_error("sdpx_batch_default_options: null options");
    return kInvalidArgument;
  }
  *o = sdpx_batch_options_t{};
  o->beta = 0.1;
  o->gamma = 0.9;
  o->eps_gap = o->eps_primal = o->eps_dual = tol;
  o->min_step = 1e-10;
  o->omega_step = 1e5;
  o->stall_tolerance = 1e-3;
  o->refine_tol = 0.0;
  o->adaptive_sigma_max = 0.0;
  o->iter_max = 200;
  o->max_restarts = 5;
  o->max_centering = 4;
  o->stall_iterations = 15;
  o->restart = 1;
  o->refine_policy = 0;
  o->refine_steps = 1;
  o->refine_max_steps = 8;
  o->adaptive = 1;
  o->step_rule = 0;
  o->precision = 0;
  o->cond_limit = 1e6;
  o->refine_target = 1e-10;
  o->mixed_max_refine = 2;  // measured fastest on the batched workloads (rounds hold the whole batch)
  o->adjoint_mu = 1e-6;
  return kOk;
}

int sdpx_batch_create(int device, int B, const int* k, const int* m, sdpx_batch_t** out) {
  if (!out || B <= 0 || !k || !m) {
    set_last_error("sdpx_batch_create: invalid arguments");
    return kInvalidArgument;
  }
  *out = nullptr;
  for (int p = 0; p < B; ++p) {
    if (k[p] < 1 || k[p] > KMAX || m[p] < 1 || m[p] > MMAX) {
      set_last_error("sdpx_batch_create: member " + std::to_string(p) + " needs 1 <= k <= " + std::to_string(KMAX) +
                     " and 1 <= m <= " + std::to_string(MMAX));
      return kInvalidArgument;
    }
  }
  if (hipSetDevice(device) != hipSuccess) {
    set_last_error("sdpx_batch_create: hipSetDevice failed");
    return kHipError;
  }
  auto* h = new sdpx_batch();
  h->device = device;
  h->B = B;
  auto fail = [&](int s) {
    sdpx_batch_destroy(h);
    return s;
  };
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    set_last_error("hipStreamCreate failed");
    return fail(kHipError);
  }
  // ---- member layout -----------------------------------------------------------------------
  h->mem.resize(B);
  for (int p = 0; p < B; ++p) {
    Member& M = h->mem[p];
    M.k = k[p];
    M.m = m[p];
    M.nblk = ceil_div(m[p], kCholNb);
    M.kp = (k[p] + 15) / 16 * 16;
    const long long k2 = (long long)k[p] * k[p];
    M.ldP = k2 + (k2 & 1);
    M.ldPf = (k2 + 3) / 4 * 4;  // the FP32 panel (4 bytes each) fits in the FP64 one: ldPf <= 2 ldP
    M.oAv = h->nAv;
    h->nAv += align_up(k2 * m[p]);
    M.oPh = h->nPh;
    h->nPh += align_up(M.ldP * m[p]);
    M.oK = h->nK;
    h->nK += align_up(k2);
    M.oM = h->nM;
    h->nM += align_up(m[p]);
    M.oS = h->nS;
    h->nS += align_up((long long)m[p] * m[p]);
    M.oD = h->nD;
    h->nD += (long long)M.nblk * kCholNb * kCholNb;
    const int nch = ceil_div(m[p], kColChunk);
    M.oG = h->nG;
    h->nG += align_up((long long)nch * k2);
    h->arr.max_chunks = std::max(h->arr.max_chunks, nch);
    h->arr.max_m = std::max(h->arr.max_m, m[p]);
  }
  int s;
  Arrays& a = h->arr;
  if ((s = balloc_n(h, &h->dmem, B))) return fail(s);
  a.mem = h->dmem;
  if ((s = balloc_n(h, &a.Av, h->nAv))) return fail(s);
  if ((s = balloc_n(h, &a.Ph, h->nPh))) return fail(s);
  double** kb[] = {&a.C, &a.X, &a.Y, &a.LX, &a.LY, &a.LXinv, &a.Xinv, &a.P, &a.R, &a.dX, &a.dY, &a.W1, &a.W2,
                   &a.bX, &a.bY, &a.G, &a.ckX, &a.ckY};
  for (double** q : kb)
    if ((s = balloc_n(h, q, h->nK))) return fail(s);
  double** mb[] = {&a.c, &a.x, &a.dx, &a.dxb, &a.d, &a.v, &a.r, &a.rho, &a.delta, &a.bx, &a.ckx};
  for (double** q : mb)
    if ((s = balloc_n(h, q, h->nM))) return fail(s);
  if ((s = balloc_n(h, &a.S, h->nS))) return fail(s);
  if ((s = balloc_n(h, &a.Sf, h->nS))) return fail(s);
  if ((s = balloc_n(h, &a.dinv, h->nD))) return fail(s);
  if ((s = balloc_n(h, &a.gpart, h->nG))) return fail(s);
  if ((s = balloc_n(h, &a.sc, (long long)B * NSLOT))) return fail(s);
  if ((s = balloc_n(h, &a.si, (long long)B * NISLOT))) return fail(s);
  if ((s = balloc_n(h, &a.par, (long long)B * NPARAM))) return fail(s);
  if ((s = balloc_n(h, &h->dlist, (long long)B * kListRing))) return fail(s);
  if ((s = balloc_n(h, &h->dall, B))) return fail(s);
  if ((s = balloc_n(h, &h->dpackK, B)) || (s = balloc_n(h, &h->dpackM, B))) return fail(s);
  {
    std::vector<long long> ok(B), om(B);
    for (int p = 0; p < B; ++p) {
      ok[p] = h->packK;
      om[p] = h->packM;
      h->packK += (long long)k[p] * k[p];
      h->packM += m[p];
    }
    if ((s = balloc_n(h, &h->dpack, std::max(h->packK, h->packM)))) return fail(s);
    if (hipMemcpy(h->dpackK, ok.data(), sizeof(long long) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->dpackM, om.data(), sizeof(long long) * B, hipMemcpyHostToDevice) != hipSuccess) {
      set_last_error("hipMemcpy(pack offsets) failed");
      return fail(kHipError);
    }
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&h->hlist), sizeof(int) * B * kListRing) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&h->hsc), sizeof(double) * B * NSLOT) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&h->hsi), sizeof(int) * B * NISLOT) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&h->hpar), sizeof(double) * B * NPARAM) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&h->hpar_ring), sizeof(double) * B * NPARAM * kParRing) != hipSuccess) {
    set_last_error("hipHostMalloc failed");
    return fail(kOutOfMemory);
  }
  for (auto& e : h->par_done)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      set_last_error("hipEventCreate failed");
      return fail(kHipError);
    }
  std::fill(h->hpar, h->hpar + (size_t)B * NPARAM, 0.0);
  // zero the panels once: padding stays finite, Av starts empty for COO input
  if (hipMemset(a.Av, 0, sizeof(double) * h->nAv) != hipSuccess ||
      hipMemset(a.Ph, 0, sizeof(double) * h->nPh) != hipSuccess ||
      hipMemset(a.sc, 0, sizeof(double) * B * NSLOT) != hipSuccess ||
      hipMemset(a.si, 0, sizeof(int) * B * NISLOT) != hipSuccess) {
    set_last_error("hipMemset failed");
    return fail(kHipError);
  }
  {
    std::vector<int> all(B);
    for (int p = 0; p < B; ++p) all[p] = p;
    if (hipMemcpy(h->dall, all.data(), sizeof(int) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->dmem, h->mem.data(), sizeof(Member) * B, hipMemcpyHostToDevice) != hipSuccess) {
      set_last_error("hipMemcpy(member table) failed");
      return fail(kHipError);
    }
  }
  // ---- Schur panel work items and the grouped Gram product ----------------------------------
  {
    std::vector<int2> work;
    for (int p = 0; p < B; ++p)
      for (int i0 = 0; i0 < m[p]; i0 += kPanelChunk) work.push_back(make_int2(p, i0));
    // big blocks first (better tail balance), grouped by padded block size for the per-size kernels
    std::stable_sort(work.begin(), work.end(), [&](const int2& x, const int2& y) { return k[x.x] > k[y.x]; });
    h->npanel = (int)work.size();
    for (int c = 0; c < 5; ++c) h->panel_off[c] = 0;
    for (const int2& wi : work) {
      const int kp = h->mem[wi.x].kp;
      const int c = kp > 48 ? 0 : kp > 32 ? 1 : kp > 16 ? 2 : 3;
      h->panel_off[c + 1] += 1;
    }
    for (int c = 1; c < 5; ++c) h->panel_off[c] += h->panel_off[c - 1];
    if ((s = balloc_n(h, &h->dpanel, h->npanel))) return fail(s);
    if (hipMemcpy(h->dpanel, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail(kHipError);
    std::vector<GroupGemm> g(B);
    std::vector<GroupTile> tiles;
    for (int p = 0; p < B; ++p) {
      const Member& M = h->mem[p];
      GroupGemm& q = g[p];
      q = GroupGemm{};
      q.A = a.Ph + M.oPh;
      q.B = a.Ph + M.oPh;
      q.C = a.S + M.oS;
      q.lda = q.ldb = M.ldP;
      q.ldc = M.m;
      q.M = q.N = M.m;
      q.K = M.k * M.k;
      q.flags = kGemmLowerOut;
      q.alpha = 1.0;
      q.beta = 0.0;
      q.skip = a.si + (long long)p * NISLOT + I_SKIP;
      const int t = ceil_div(M.m, 64);
      for (int ti = 0; ti < t; ++ti)
        for (int tj = 0; tj <= ti; ++tj) tiles.push_back(GroupTile{p, ti, tj, 0});
    }
    h->gram_tiles = tiles;
    tiles = gram_tile_order(h, 0);
    h->ngram_tiles = (int)tiles.size();
    std::vector<GroupGramF32> g32(B);
    for (int p = 0; p < B; ++p) {
      const Member& M = h->mem[p];
      g32[p] = GroupGramF32{reinterpret_cast<const float*>(a.Ph + M.oPh), a.S + M.oS, M.ldPf, (long long)M.m, M.m,
                            (int)M.ldPf, a.si + (long long)p * NISLOT + I_SKIP32};
    }
    if ((s = balloc_n(h, &h->dgram, B))) return fail(s);
    if ((s = balloc_n(h, &h->dgram32, B))) return fail(s);
    if ((s = balloc_n(h, &h->dgram_tiles, h->ngram_tiles))) return fail(s);
    if (hipMemcpy(h->dgram32, g32.data(), sizeof(GroupGramF32) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->dgram, g.data(), sizeof(GroupGemm) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->dgram_tiles, tiles.data(), sizeof(GroupTile) * tiles.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
      return fail(kHipError);
  }
  // ---- grouped Schur Cholesky schedule ---------------------------------------------------------
  {
    std::vector<CholDesc> cd(B);
    int maxblk = 0;
    for (int p = 0; p < B; ++p) {
      const Member& M = h->mem[p];
      cd[p] = CholDesc{a.Sf + M.oS, (long long)M.m, M.m, a.dinv + M.oD, a.si + (long long)p * NISLOT + I_INFOS};
      maxblk = std::max(maxblk, M.nblk);
    }
    std::vector<int> lists;
    std::vector<GroupGemm> g;
    std::vector<GroupTile> tiles;
    for (int J = 0; J < maxblk; ++J) {
      const int j = J * kCholNb;
      h->step_list_off.push_back((int)lists.size());
      for (int p = 0; p < B; ++p)
        if (h->mem[p].nblk > J) lists.push_back(p);
      h->step_list_cnt.push_back((int)lists.size() - h->step_list_off.back());
      std::vector<GroupTile> panel, trail;
      for (int p = 0; p < B; ++p) {
        const Member& M = h->mem[p];
        if (M.nblk <= J) continue;
        const int jb = std::min(kCholNb, M.m - j), rest = M.m - j - jb;
        if (rest <= 0) continue;
        double* Sf = a.Sf + M.oS;
        int* info = a.si + (long long)p * NISLOT + I_INFOS;
        GroupGemm pq{};
        pq.A = Sf + j + jb + (long long)j * M.m;
        pq.lda = M.m;
        pq.B = a.dinv + M.oD + (long long)J * kCholNb * kCholNb;
        pq.ldb = kCholNb;
        pq.C = Sf + j + jb + (long long)j * M.m;
        pq.ldc = M.m;
        pq.M = rest;
        pq.N = jb;
        pq.K = jb;
        pq.alpha = 1.0;
        pq.beta = 0.0;
        pq.skip = info;
        const int gp = (int)g.size();
        g.push_back(pq);
        for (int ti = 0; ti < ceil_div(rest, 64); ++ti) panel.push_back(GroupTile{gp, ti, 0, 0});
        GroupGemm tq{};
        tq.A = Sf + j + jb + (long long)j * M.m;
        tq.lda = M.m;
        tq.B = tq.A;
        tq.ldb = M.m;
        tq.C = Sf + j + jb + (long long)(j + jb) * M.m;
        tq.ldc = M.m;
        tq.M = tq.N = rest;
        tq.K = jb;
        tq.flags = kGemmLowerOut;
        tq.alpha = -1.0;
        tq.beta = 1.0;
        tq.skip = info;
        const int gt = (int)g.size();
        g.push_back(tq);
        const int t = ceil_div(rest, 64);
        for (int ti = 0; ti < t; ++ti)
          for (int tj = 0; tj <= ti; ++tj) trail.push_back(GroupTile{gt, ti, tj, 0});
      }
      h->step_panel_off.push_back((int)tiles.size());
      h->step_panel_cnt.push_back((int)panel.size());
      tiles.insert(tiles.end(), panel.begin(), panel.end());
      h->step_trail_off.push_back((int)tiles.size());
      h->step_trail_cnt.push_back((int)trail.size());
      tiles.insert(tiles.end(), trail.begin(), trail.end());
    }
    if ((s = balloc_n(h, &h->dchol, B))) return fail(s);
    if ((s = balloc_n(h, &h->dchol_list, std::max<long long>(1, (long long)lists.size())))) return fail(s);
    if ((s = balloc_n(h, &h->dfac, std::max<long long>(1, (long long)g.size())))) return fail(s);
    if ((s = balloc_n(h, &h->dfac_tiles, std::max<long long>(1, (long long)tiles.size())))) return fail(s);
    if (hipMemcpy(h->dchol, cd.data(), sizeof(CholDesc) * B, hipMemcpyHostToDevice) != hipSuccess ||
        (!lists.empty() &&
         hipMemcpy(h->dchol_list, lists.data(), sizeof(int) * lists.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        (!g.empty() && hipMemcpy(h->dfac, g.data(), sizeof(GroupGemm) * g.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        (!tiles.empty() &&
         hipMemcpy(h->dfac_tiles, tiles.data(), sizeof(GroupTile) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess))
      return fail(kHipError);
  }
  h->st.resize(B);
  h->scale_p.assign(B, 1.0);
  h->scale_d.assign(B, 1.0);
  *out = h;
  return kOk;
}

int sdpx_batch_destroy(sdpx_batch_t* h) {
  if (!h) return kOk;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  clear_sparse(h);
  for (void* p : h->allocs) hipFree(p);
  if (h->hlist) hipHostFree(h->hlist);
  if (h->hsc) hipHostFree(h->hsc);
  if (h->hsi) hipHostFree(h->hsi);
  if (h->hpar) hipHostFree(h->hpar);
  if (h->hpar_ring) hipHostFree(h->hpar_ring);
  for (auto& e : h->par_done)
    if (e) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return kOk;
}

int sdpx_batch_memory(sdpx_batch_t* h, long long* bytes) {
  SDPX_CHECK(check_batch(h));
  if (bytes) *bytes = h->bytes;
  return kOk;
}

namespace {
// scale_p = 1 + max |C|, scale_d = 1 + max |c| (interior_point.jl)
void set_scales(sdpx_batch_t* h, const double* C, const double* c) {
  long long oc = 0, om = 0;
  for (int p = 0; p < h->B; ++p) {
    const Member& M = h->mem[p];
    const long long k2 = (long long)M.k * M.k;
    double mc = 0.0, mv = 0.0;
    for (long long e = 0; e < k2; ++e) mc = std::max(mc, std::fabs(C[oc + e]));
    for (int i = 0; i < M.m; ++i) mv = std::max(mv, std::fabs(c[om + i]));
    h->scale_p[p] = 1.0 + mc;
    h->scale_d[p] = 1.0 + mv;
    oc += k2;
    om += M.m;
  }
}

// inverse of pack_blocks_kernel: an unpadded member-after-member buffer to the
// members' aligned device offsets
__global__ __launch_bounds__(256) void unpack_blocks_kernel(const Member* __restrict__ mem,
                                                           const double* __restrict__ src,
                                                           const long long* __restrict__ off, int kblock,
                                                           double* __restrict__ dst) {
  const Member M = mem[blockIdx.x];
  const long long n = kblock ? (long long)M.k * M.k : M.m;
  const double* s = src + off[blockIdx.x];
  double* d = dst + (kblock ? M.oK : M.oM);
  for (long long i = threadIdx.x; i < n; i += 256) d[i] = s[i];
}

// host (members concatenated, unpadded) -> device (aligned member offsets): one
// copy into the staging buffer, then scattered on the device
int upload_blocks(sdpx_batch_t* h, double* dst, const double* src, bool kblock) {
  if (h->B <= 0) return kOk;
  SDPX_HIP(hipMemcpyAsync(h->dpack, src, sizeof(double) * (kblock ? h->packK : h->packM), hipMemcpyHostToDevice,
                          h->stream));
  hipLaunchKernelGGL(unpack_blocks_kernel, dim3(h->B), dim3(256), 0, h->stream, h->dmem, h->dpack,
                     kblock ? h->dpackK : h->dpackM, kblock ? 1 : 0, dst);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

// members' k x k blocks (or m-vectors) from their aligned device offsets into
// one unpadded buffer, member after member (one work-group per member)
__global__ __launch_bounds__(256) void pack_blocks_kernel(const Member* __restrict__ mem, const double* __restrict__ src,
                                                         const long long* __restrict__ off, int kblock,
                                                         double* __restrict__ dst) {
  const Member M = mem[blockIdx.x];
  const long long n = kblock ? (long long)M.k * M.k : M.m;
  const double* s = src + (kblock ? M.oK : M.oM);
  double* d = dst + off[blockIdx.x];
  for (long long i = threadIdx.x; i < n; i += 256) d[i] = s[i];
}

// device (aligned member offsets) -> host (members concatenated, unpadded):
// packed on the device, then one copy (a copy per member costs ~4 us each)
int download_blocks(sdpx_batch_t* h, double* dst, const double* src, bool kblock) {
  if (h->B <= 0) return kOk;
  hipLaunchKernelGGL(pack_blocks_kernel, dim3(h->B), dim3(256), 0, h->stream, h->dmem, src,
                     kblock ? h->dpackK : h->dpackM, kblock ? 1 : 0, h->dpack);
  SDPX_HIP(hipGetLastError());
  SDPX_HIP(hipMemcpyAsync(dst, h->dpack, sizeof(double) * (kblock ? h->packK : h->packM), hipMemcpyDeviceToHost,
                          h->stream));
  return kOk;
}
}  // namespace

int sdpx_batch_set_dense(sdpx_batch_t* h, const double* Av, const double* C, const double* c) {
  SDPX_CHECK(check_batch(h));
  h->have_data = false;  // set again only when the whole upload succeeds
  if (!Av || !C || !c) {
    set_last_error("sdpx_batch_set_dense: null data");
    return kInvalidArgument;
  }
  clear_sparse(h);  // dense constraints: the gather kernels no longer apply
  long long o = 0;
  for (int p = 0; p < h->B; ++p) {
    const Member& M = h->mem[p];
    const long long n = (long long)M.k * M.k * M.m;
    SDPX_HIP(hipMemcpyAsync(h->arr.Av + M.oAv, Av + o, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    o += n;
  }
  SDPX_CHECK(upload_blocks(h, h->arr.C, C, true));
  SDPX_CHECK(upload_blocks(h, h->arr.c, c, false));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  set_scales(h, C, c);
  h->have_data = true;
  return kOk;
}

int sdpx_batch_set_coo(sdpx_batch_t* h, const long long* ptr, const int* row, const int* col, const double* val,
                       const double* C, const double* c) {
  SDPX_CHECK(check_batch(h));
  h->have_data = false;  // set again only when the whole upload succeeds
  if (!ptr || !C || !c) {
    set_last_error("sdpx_batch_set_coo: null data");
    return kInvalidArgument;
  }
  long long ncons = 0;
  for (int p = 0; p < h->B; ++p) ncons += h->mem[p].m;
  if (ptr[0] != 0) {
    set_last_error("sdpx_batch_set_coo: ptr[0] must be 0");
    return kInvalidArgument;
  }
  for (long long q = 0; q < ncons; ++q) {
    if (ptr[q + 1] < ptr[q]) {
      set_last_error("sdpx_batch_set_coo: ptr must be non-decreasing (constraint " + std::to_string(q) + ")");
      return kInvalidArgument;
    }
  }
  const long long nnz = ptr[ncons];
  if (nnz > 0 && (!row || !col || !val)) {
    set_last_error("sdpx_batch_set_coo: null row, col or val with entries present");
    return kInvalidArgument;
  }
  // column offsets and block sides per constraint; entries checked on the host
  std::vector<long long> cons(2 * ncons);
  long long j = 0;
  for (int p = 0; p < h->B; ++p) {
    const Member& M = h->mem[p];
    for (int i = 0; i < M.m; ++i, ++j) {
      cons[2 * j] = M.oAv + (long long)i * M.k * M.k;
      cons[2 * j + 1] = M.k;
      for (long long e = ptr[j]; e < ptr[j + 1]; ++e) {
        if (row[e] < 0 || col[e] < row[e] || col[e] >= M.k) {
          set_last_error("sdpx_batch_set_coo: entry outside the upper triangle of its block");
          return kInvalidArgument;
        }
      }
    }
  }
  long long* dptr = nullptr;
  long long* dcons = nullptr;
  int *drow = nullptr, *dcol = nullptr;
  double* dval = nullptr;
  auto release = [&]() {
    if (dptr) hipFree(dptr);
    if (dcons) hipFree(dcons);
    if (drow) hipFree(drow);
    if (dcol) hipFree(dcol);
    if (dval) hipFree(dval);
  };
  if (hipMalloc(&dptr, sizeof(long long) * (ncons + 1)) != hipSuccess ||
      hipMalloc(&dcons, sizeof(long long) * 2 * ncons) != hipSuccess ||
      hipMalloc(&drow, sizeof(int) * std::max<long long>(nnz, 1)) != hipSuccess ||
      hipMalloc(&dcol, sizeof(int) * std::max<long long>(nnz, 1)) != hipSuccess ||
      hipMalloc(&dval, sizeof(double) * std::max<long long>(nnz, 1)) != hipSuccess) {
    release();
    set_last_error("sdpx_batch_set_coo: hipMalloc failed");
    return kOutOfMemory;
  }
  int st = kOk;
  if (hipMemcpy(dptr, ptr, sizeof(long long) * (ncons + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dcons, cons.data(), sizeof(long long) * 2 * ncons, hipMemcpyHostToDevice) != hipSuccess ||
      (nnz > 0 && (hipMemcpy(drow, row, sizeof(int) * nnz, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(dcol, col, sizeof(int) * nnz, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(dval, val, sizeof(double) * nnz, hipMemcpyHostToDevice) != hipSuccess)) ||
      hipMemsetAsync(h->arr.Av, 0, sizeof(double) * h->nAv, h->stream) != hipSuccess) {
    set_last_error("sdpx_batch_set_coo: upload failed");
    st = kHipError;
  }
  if (st == kOk) st = launch_coo_densify(h->arr, dptr, drow, dcol, dval, dcons, (int)ncons, h->stream);
  if (st == kOk) st = build_sparse(h, ptr, row, col, val);
  if (st == kOk) st = upload_blocks(h, h->arr.C, C, true);
  if (st == kOk) st = upload_blocks(h, h->arr.c, c, false);
  if (st == kOk && hipStreamSynchronize(h->stream) != hipSuccess) st = kHipError;
  release();
  if (st != kOk) return st;
  set_scales(h, C, c);
  h->have_data = true;
  return kOk;
}

// SparseBlockCOO as stored (types/constraints.jl:59-109): positions 1..m with 1-based Int32
// ptr (ptr[0] = 1), every stored entry of the symmetric A_i (both triangles), 1-based row /
// col and lin = (col - 1) k + row.  Converted to this ABI's COO: 0-based, row <= col, in
// stored order; a lower entry must mirror an upper entry of the same position with the same
// value (checked), and is then dropped.
static int block_coo_convert(int k, int m, const int* ptr, const int* lin, const int* row, const int* col,
                             const double* val, std::vector<long long>& optr, std::vector<int>& orow,
                             std::vector<int>& ocol, std::vector<double>& oval, const char* who) {
  if (k < 1 || m < 1 || !ptr) {
    set_last_error(std::string(who) + ": invalid block side, count or null ptr");
    return kInvalidArgument;
  }
  if (ptr[0] != 1) {
    set_last_error(std::string(who) + ": ptr[1] must be 1 (1-based positions)");
    return kInvalidArgument;
  }
  for (int i = 0; i < m; ++i)
    if (ptr[i + 1] < ptr[i]) {
      set_last_error(std::string(who) + ": ptr must be non-decreasing");
      return kInvalidArgument;
    }
  const long long total = (long long)ptr[m] - 1;
  if (total > 0 && (!lin || !row || !col || !val)) {
    set_last_error(std::string(who) + ": null lin, row, col or val with entries present");
    return kInvalidArgument;
  }
  for (int i = 0; i < m; ++i) {
    optr.push_back((long long)orow.size());
    std::vector<std::pair<long long, double>> upper, lower;  // (0-based lin of the upper twin, value)
    for (long long e = (long long)ptr[i] - 1; e < (long long)ptr[i + 1] - 1; ++e) {
      const int r = row[e], c = col[e];
      if (r < 1 || r > k || c < 1 || c > k || lin[e] != (c - 1) * k + r) {
        set_last_error(std::string(who) + ": entry " + std::to_string(e + 1) + " of position " +
                       std::to_string(i + 1) + " has row/col outside the block or lin != (col-1)k+row");
        return kInvalidArgument;
      }
      if (r <= c) {
        orow.push_back(r - 1);
        ocol.push_back(c - 1);
        oval.push_back(val[e]);
        if (r < c) upper.push_back({(long long)(c - 1) * k + (r - 1), val[e]});
      } else {
        lower.push_back({(long long)(r - 1) * k + (c - 1), val[e]});
      }
    }
    // each stored lower entry has an upper twin with the same value (sums of duplicates compared)
    auto merge = [](std::vector<std::pair<long long, double>>& v) {
      std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      std::vector<std::pair<long long, double>> out;
      for (const auto& x : v) {
        if (!out.empty() && out.back().first == x.first)
          out.back().second += x.second;
        else
          out.push_back(x);
      }
      v.swap(out);
    };
    merge(upper);
    merge(lower);
    if (upper.size() != lower.size()) {
      set_last_error(std::string(who) + ": position " + std::to_string(i + 1) +
                     " does not store both triangles of a symmetric matrix");
      return kInvalidArgument;
    }
    for (size_t q = 0; q < upper.size(); ++q)
      if (upper[q].first != lower[q].first || upper[q].second != lower[q].second) {
        set_last_error(std::string(who) + ": position " + std::to_string(i + 1) + " is not symmetric");
        return kInvalidArgument;
      }
  }
  return kOk;
}

int sdpx_block_coo_convert(int k, int m, const int* ptr, const int* lin, const int* row, const int* col,
                           const double* val, long long* out_ptr, int* out_row, int* out_col, double* out_val,
                           long long* out_nnz) {
  std::vector<long long> optr;
  std::vector<int> orow, ocol;
  std::vector<double> oval;
  SDPX_CHECK(block_coo_convert(k, m, ptr, lin, row, col, val, optr, orow, ocol, oval, "sdpx_block_coo_convert"));
  optr.push_back((long long)orow.size());
  if (out_nnz) *out_nnz = (long long)orow.size();
  if (out_ptr) std::copy(optr.begin(), optr.end(), out_ptr);
  if (out_row) std::copy(orow.begin(), orow.end(), out_row);
  if (out_col) std::copy(ocol.begin(), ocol.end(), out_col);
  if (out_val) std::copy(oval.begin(), oval.end(), out_val);
  return kOk;
}

int sdpx_batch_set_block_coo(sdpx_batch_t* h, const int* ptr, const int* lin, const int* row, const int* col,
                             const double* val, const double* C, const double* c) {
  SDPX_CHECK(check_batch(h));
  h->have_data = false;
  if (!ptr || !C || !c) {
    set_last_error("sdpx_batch_set_block_coo: null data");
    return kInvalidArgument;
  }
  std::vector<long long> optr;
  std::vector<int> orow, ocol;
  std::vector<double> oval;
  long long pofs = 0, eofs = 0;
  for (int p = 0; p < h->B; ++p) {
    const Member& M = h->mem[p];
    const int* pp = ptr + pofs;
    const long long ne = (long long)pp[M.m] - 1;
    SDPX_CHECK(block_coo_convert(M.k, M.m, pp, lin ? lin + eofs : nullptr, row ? row + eofs : nullptr,
                                 col ? col + eofs : nullptr, val ? val + eofs : nullptr, optr, orow, ocol, oval,
                                 "sdpx_batch_set_block_coo"));
    pofs += M.m + 1;
    eofs += ne;
  }
  optr.push_back((long long)orow.size());
  return sdpx_batch_set_coo(h, optr.data(), orow.data(), ocol.data(), oval.data(), C, c);
}

int sdpx_batch_set_start(sdpx_batch_t* h, const double* x, const double* X, const double* Y) {
  SDPX_CHECK(check_batch(h));
  if (!x || !X || !Y) {
    set_last_error("sdpx_batch_set_start: null data");
    return kInvalidArgument;
  }
  SDPX_CHECK(upload_blocks(h, h->arr.x, x, false));
  SDPX_CHECK(upload_blocks(h, h->arr.X, X, true));
  SDPX_CHECK(upload_blocks(h, h->arr.Y, Y, true));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  h->have_start = true;
  return kOk;
}

int sdpx_batch_solve(sdpx_batch_t* h, const sdpx_batch_options_t* o, int* iterations) {
  SDPX_CHECK(check_batch(h));
  if (!o || !h->have_data || !h->have_start) {
    set_last_error("sdpx_batch_solve: options, data (set_dense / set_coo) and a start point are required");
    return kInvalidArgument;
  }
  if ((o->precision != 0 && o->precision != 1) || o->mixed_max_refine < 0) {
    set_last_error("sdpx_batch_solve: precision 0 (plain FP64) or 1 (FP32 Schur assembly, FP64 refinement), "
                   "mixed_max_refine >= 0");
    return kInvalidArgument;
  }
  if (!(o->beta > 0 && o->beta < 1) || !(o->gamma > 0 && o->gamma < 1) || o->step_rule < 0 || o->step_rule > 2) {
    set_last_error("sdpx_batch_solve: beta and gamma must lie in (0, 1), step_rule in 0..2");
    return kInvalidArgument;
  }
  const int s = run_solve(h, *o, iterations);
  h->solved = s == kOk;
  return s;
}

int sdpx_batch_get_result(sdpx_batch_t* h, int* status, int* iterations, double* pobj, double* dobj, double* x,
                          double* X, double* Y) {
  SDPX_CHECK(check_batch(h));
  if (!h->solved) {
    set_last_error("sdpx_batch_get_result: no solve has completed");
    return kInvalidArgument;
  }
  for (int p = 0; p < h->B; ++p) {
    const MemberState& s = h->st[p];
    if (status) status[p] = s.status;
    if (iterations) iterations[p] = s.iteration;
    if (pobj) pobj[p] = s.pobj;
    if (dobj) dobj[p] = s.dobj;
  }
  if (x) SDPX_CHECK(download_blocks(h, x, h->arr.x, false));
  if (X) SDPX_CHECK(download_blocks(h, X, h->arr.X, true));
  if (Y) SDPX_CHECK(download_blocks(h, Y, h->arr.Y, true));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

namespace {
// L_X, L_Y, X^-1 and the factored Schur matrix at the current iterate for every
// member; members whose blocks or Schur matrix do not factor are dropped from `ok`.
int factor_at_iterate(sdpx_batch_t* h, double mu_scale, std::vector<int>& ok) {
  std::vector<int> all(h->B);
  for (int p = 0; p < h->B; ++p) all[p] = p;
  SDPX_CHECK(launch_comp(h->arr, put_list(h, all), h->B, h->stream));
  SDPX_CHECK(pull_scalars(h));
  for (int p = 0; p < h->B; ++p) {
    par(h, p, P_MU) = mu_scale * sc(h, p, S_COMP) / h->mem[p].k;
    par(h, p, P_AFF) = 0.0;
  }
  SDPX_CHECK(push_params(h));
  const int* l = put_list(h, all);
  SDPX_CHECK(launch_gemv(h->arr, l, h->B, 0, h->stream));
  SDPX_CHECK(launch_gemvt(h->arr, l, h->B, 0, h->stream));
  SDPX_CHECK(launch_block_factor(h->arr, l, h->B, h->stream));
  SDPX_CHECK(pull_scalars(h, true));
  ok.clear();
  for (int p = 0; p < h->B; ++p)
    if (si(h, p, I_INFOX) == 0 && si(h, p, I_INFOY) == 0) ok.push_back(p);
  SDPX_CHECK(schur_build(h, ok));
  SDPX_CHECK(schur_factor(h, ok));
  std::vector<int> keep;
  for (int p : ok)
    if (h->st[p].factor_ok) keep.push_back(p);
  ok.swap(keep);
  return kOk;
}
}  // namespace

int sdpx_batch_restore_checkpoint(sdpx_batch_t* h, int* restored, double* mu) {
  SDPX_CHECK(check_batch(h));
  std::vector<int> l;
  for (int p = 0; p < h->B; ++p) {
    if (h->st[p].has_ck) l.push_back(p);
    if (mu) mu[p] = h->st[p].has_ck ? h->st[p].ck_mu : NAN;
  }
  if (!l.empty()) SDPX_CHECK(launch_best(h->arr, put_list(h, l), (int)l.size(), 3, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  if (restored) *restored = (int)l.size();
  return kOk;
}

int sdpx_batch_center(sdpx_batch_t* h, int iterations) { return sdpx_batch_center_at(h, iterations, nullptr); }

int sdpx_batch_center_at(sdpx_batch_t* h, int iterations, const double* target) {
  SDPX_CHECK(check_batch(h));
  if (!h->have_data || !h->have_start || iterations < 0) {
    set_last_error("sdpx_batch_center: data and an iterate are required, iterations >= 0");
    return kInvalidArgument;
  }
  if (target) {
    for (int p = 0; p < h->B; ++p) {
      if (!(std::isnan(target[p]) || (target[p] > 0.0 && std::isfinite(target[p])))) {
        set_last_error("sdpx_batch_center_at: mu[" + std::to_string(p) + "] must be positive and finite, or NaN");
        return kInvalidArgument;
      }
    }
  }
  // mu fixed at the member's target, or at its value on entry
  std::vector<int> all(h->B);
  for (int p = 0; p < h->B; ++p) {
    all[p] = p;
    par(h, p, P_A) = par(h, p, P_B) = 0.0;
  }
  SDPX_CHECK(push_params(h));
  SDPX_CHECK(launch_comp(h->arr, put_list(h, all), h->B, h->stream));
  SDPX_CHECK(pull_scalars(h));
  std::vector<double> mu(h->B);
  for (int p = 0; p < h->B; ++p)
    mu[p] = (target && !std::isnan(target[p])) ? target[p] : sc(h, p, S_COMP) / h->mem[p].k;
  for (int it = 0; it < iterations; ++it) {
    std::vector<int> ok;
    SDPX_CHECK(factor_at_iterate(h, 1.0, ok));
    for (int p = 0; p < h->B; ++p) par(h, p, P_MU) = mu[p];
    SDPX_CHECK(push_params(h));
    // R = mu I - X Y with the fixed mu, then one Newton step on the central-path equations
    const int n = (int)ok.size();
    const int* l = put_list(h, ok);
    SDPX_CHECK(launch_block_factor(h->arr, l, n, h->stream));
    SDPX_CHECK(launch_rhs_block(h->arr, l, n, h->stream));
    SDPX_CHECK(launch_gemvt(h->arr, l, n, 1, h->stream));
    SDPX_CHECK(launch_solve(h->arr, l, n, 0, h->stream));
    SDPX_CHECK(launch_gemv(h->arr, l, n, 1, h->stream));
    SDPX_CHECK(launch_dir_block(h->arr, l, n, h->stream));
    for (int p : ok) {
      par(h, p, P_GAMMA) = 0.9;
      par(h, p, P_MINSTEP) = 1e-10;
      par(h, p, P_RATIO) = 0.0;
    }
    SDPX_CHECK(push_params(h));
    SDPX_CHECK(launch_backtrack(h->arr, put_list(h, ok), n, step_classes(h, ok), h->stream));
    SDPX_CHECK(pull_scalars(h));
    for (int p : ok) {
      par(h, p, P_A) = sc(h, p, S_TX);
      par(h, p, P_B) = sc(h, p, S_TY);
    }
    SDPX_CHECK(push_params(h));
    SDPX_CHECK(launch_update(h->arr, put_list(h, ok), n, h->stream));
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_batch_adjoint(sdpx_batch_t* h, const double* G, const double* hv, double* dC, double* db) {
  SDPX_CHECK(check_batch(h));
  if (!h->have_data || !h->have_start || !G || !hv) {
    set_last_error("sdpx_batch_adjoint: data, an iterate, G and h are required");
    return kInvalidArgument;
  }
  std::vector<int> ok;
  SDPX_CHECK(factor_at_iterate(h, 0.0, ok));
  SDPX_CHECK(upload_blocks(h, h->arr.G, G, true));
  SDPX_CHECK(upload_blocks(h, h->arr.v, hv, false));
  const int n = (int)ok.size();
  const int* l = put_list(h, ok);
  SDPX_CHECK(launch_adj_w(h->arr, l, n, h->stream));
  SDPX_CHECK(launch_gemvt(h->arr, l, n, 1, h->stream));    // r = h - Av' vec(W)
  SDPX_CHECK(launch_solve(h->arr, l, n, 0, h->stream));    // dx = lambda = S^-1 r
  SDPX_CHECK(launch_gemv(h->arr, l, n, 3, h->stream));     // dX = A_lambda
  SDPX_CHECK(launch_adj_final(h->arr, l, n, h->stream));   // W2 = dL/dC_std
  if (dC) SDPX_CHECK(download_blocks(h, dC, h->arr.W2, true));
  if (db) SDPX_CHECK(download_blocks(h, db, h->arr.dx, false));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  // members that could not be factored get NaN
  std::vector<char> good(h->B, 0);
  for (int p : ok) good[p] = 1;
  long long oc = 0, om = 0;
  for (int p = 0; p < h->B; ++p) {
    const long long k2 = (long long)h->mem[p].k * h->mem[p].k;
    if (!good[p]) {
      if (dC) std::fill(dC + oc, dC + oc + k2, NAN);
      if (db) std::fill(db + om, db + om + h->mem[p].m, NAN);
    }
    oc += k2;
    om += h->mem[p].m;
  }
  return n == h->B ? kOk : kNotPositiveDefinite;
}

int sdpx_batch_get_trace(sdpx_batch_t* h, int member, double* rows, int cap, int* n) {
  SDPX_CHECK(check_batch(h));
  if (member < 0 || member >= h->B) {
    set_last_error("sdpx_batch_get_trace: invalid member");
    return kInvalidArgument;
  }
  const auto& tr = h->st[member].trace;
  const int cnt = (int)tr.size();
  if (n) *n = cnt;
  if (rows)
    for (int i = 0; i < std::min(cnt, cap); ++i)
      for (int j = 0; j < kTraceCols; ++j) rows[(size_t)i * kTraceCols + j] = tr[i][j];
  return kOk;
}

int sdpx_batch_get_schur(sdpx_batch_t* h, int member, double* S) {
  SDPX_CHECK(check_batch(h));
  if (member < 0 || member >= h->B || !S) {
    set_last_error("sdpx_batch_get_schur: invalid member");
    return kInvalidArgument;
  }
  const Member& M = h->mem[member];
  SDPX_HIP(hipMemcpy(S, h->arr.S + M.oS, sizeof(double) * M.m * M.m, hipMemcpyDeviceToHost));
  return kOk;
}

int sdpx_batch_schur_at_start(sdpx_batch_t* h) {
  SDPX_CHECK(check_batch(h));
  if (!h->have_data || !h->have_start) {
    set_last_error("sdpx_batch_schur_at_start: data and a start point are required");
    return kInvalidArgument;
  }
  std::vector<int> all(h->B);
  for (int p = 0; p < h->B; ++p) {
    all[p] = p;
    par(h, p, P_MU) = 0.0;
    par(h, p, P_AFF) = 1.0;
    h->st[p].mixed = h->precision != 0;
  }
  SDPX_CHECK(push_params(h));
  SDPX_CHECK(launch_block_factor(h->arr, put_list(h, all), h->B, h->stream));
  SDPX_CHECK(pull_scalars(h, true));
  for (int p = 0; p < h->B; ++p)
    if (si(h, p, I_INFOX) != 0 || si(h, p, I_INFOY) != 0) {
      set_last_error("sdpx_batch_schur_at_start: start point not positive definite");
      return kNotPositiveDefinite;
    }
  SDPX_CHECK(schur_build(h, all));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_batch_set_precision(sdpx_batch_t* h, int precision) {
  SDPX_CHECK(check_batch(h));
  if (precision != 0 && precision != 1) {
    set_last_error("sdpx_batch_set_precision: precision 0 (FP64) or 1 (FP32)");
    return kInvalidArgument;
  }
  h->precision = precision;
  return kOk;
}

int sdpx_batch_get_mixed_iterations(sdpx_batch_t* h, int* out) {
  SDPX_CHECK(check_batch(h));
  if (!h->solved || !out) {
    set_last_error("sdpx_batch_get_mixed_iterations: no solve has completed, or null output");
    return kInvalidArgument;
  }
  for (int p = 0; p < h->B; ++p) out[p] = h->st[p].mixed_iterations;
  return kOk;
}

int sdpx_batch_get_reasons(sdpx_batch_t* h, int* out) {
  SDPX_CHECK(check_batch(h));
  if (!h->solved || !out) {
    set_last_error("sdpx_batch_get_reasons: no solve has completed, or null output");
    return kInvalidArgument;
  }
  for (int p = 0; p < h->B; ++p) out[p] = h->st[p].reason;
  return kOk;
}

int sdpx_batch_set_gram_order(sdpx_batch_t* h, int mode) {
  SDPX_CHECK(check_batch(h));
  if (mode < 0 || mode > 2) {
    set_last_error("sdpx_batch_set_gram_order: mode 0 (K-descending), 1 (member-major) or 2 (XCD bins)");
    return kInvalidArgument;
  }
  return set_gram_order(h, mode);
}

int sdpx_batch_autotune(sdpx_batch_t* h, int repeats, double* out) {
  SDPX_CHECK(check_batch(h));
  if (repeats < 1 || !out) {
    set_last_error("sdpx_batch_autotune: repeats >= 1 and out (4) required");
    return kInvalidArgument;
  }
  // panels of every member at the start point, then each candidate order's Gram launch timed
  SDPX_CHECK(sdpx_batch_schur_at_start(h));
  hipEvent_t e0, e1;
  SDPX_HIP(hipEventCreate(&e0));
  SDPX_HIP(hipEventCreate(&e1));
  double best = INFINITY;
  int best_mode = 0;
  int rc = kOk;
  for (int mode = 0; mode < 3 && rc == kOk; ++mode) {
    rc = set_gram_order(h, mode);
    std::vector<float> ms;
    for (int r = 0; r <= repeats && rc == kOk; ++r) {  // the first launch of an order is a warm-up
      if (hipEventRecord(e0, h->stream) != hipSuccess) rc = kHipError;
      if (rc == kOk)
        rc = h->precision != 0 ? gram_f32_grouped(h->dgram32, h->dgram_tiles, h->ngram_tiles, h->stream)
                               : gemm_f64_grouped(h->dgram, h->dgram_tiles, h->ngram_tiles, true, false, true, h->stream);
      if (rc == kOk && (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = kHipError;
      float f = 0.f;
      if (rc == kOk && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = kHipError;
      if (rc == kOk && r > 0) ms.push_back(f);
    }
    if (rc != kOk) break;
    std::sort(ms.begin(), ms.end());
    out[1 + mode] = ms[ms.size() / 2];
    if (out[1 + mode] < best) {
      best = out[1 + mode];
      best_mode = mode;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != kOk) {
    if (rc == kHipError) set_last_error("sdpx_batch_autotune: timing the Gram launch failed");
    return rc;
  }
  SDPX_CHECK(set_gram_order(h, best_mode));
  out[0] = best_mode;
  return kOk;
}

/* Host-only hooks on the per-member controller (no device call). */
int sdpx_batch_policy_select(const sdpx_batch_options_t* o, const double* diag, int nhist, const double* hist,
                             int preset_fallback, double* out) {
  if (!o || !diag || !out || nhist < 0 || (nhist > 0 && !hist)) {
    set_last_error("sdpx_batch_policy_select: invalid arguments");
    return kInvalidArgument;
  }
  Controller c;
  c.init(*o);
  for (int i = 0; i < nhist; ++i) {
    Hist r{};
    r.sigma = hist[2 * i];
    r.unstable = hist[2 * i + 1] != 0.0;
    c.history.push_back(r);
  }
  c.fallback = preset_fallback != 0;
  Diag d;
  d.iteration = (int)diag[0];
  d.primal_residual = diag[1];
  d.dual_residual = diag[2];
  d.relative_gap = diag[3];
  d.mu = diag[4];
  d.mu_aff = diag[5];
  d.affine_primal_step = diag[6];
  d.affine_dual_step = diag[7];
  d.previous_primal_step = diag[8];
  d.previous_dual_step = diag[9];
  d.backtracking_count = (int)diag[10];
  d.regularization = diag[11];
  d.refinement_count = (int)diag[12];
  d.factorization_quality = diag[13];
  d.achieved_residual_reduction = diag[14];
  d.primal_psd_margin = diag[15];
  d.dual_psd_margin = diag[16];
  const Params p = c.select(d);
  const double v[] = {p.sigma, p.pf, p.df, p.bt, p.rtol, (double)p.rmax, p.min_step, p.restart_scale, p.reg,
                      p.fallback ? 1.0 : 0.0, (double)p.reason, c.fallback ? 1.0 : 0.0, c.beta, c.gamma,
                      c.ap.sigma_max, c.ap.sigma_min};
  for (int i = 0; i < 16; ++i) out[i] = v[i];
  return kOk;
}

int sdpx_batch_policy_record(const sdpx_batch_options_t* o, const double* rec, double* out) {
  if (!o || !rec || !out) {
    set_last_error("sdpx_batch_policy_record: invalid arguments");
    return kInvalidArgument;
  }
  Controller c;
  c.init(*o);
  c.record((int)rec[0], rec[1], rec[2], rec[3], rec[4], rec[5], rec[6], rec[7], rec[8], (int)rec[9], (int)rec[10],
           c.parameters);
  const Hist& h = c.history.back();
  out[0] = h.achieved;
  out[1] = h.unstable ? 1.0 : 0.0;
  out[2] = c.fallback ? 1.0 : 0.0;
  out[3] = (double)c.instability_count;
  return kOk;
}

}  // extern "C"
