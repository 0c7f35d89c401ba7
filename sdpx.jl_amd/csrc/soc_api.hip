// Lorentz (second-order) cone kernels: many cones of arbitrary dimension
// resident in HBM, one thread per cone.
//
// Layout: the cones' coordinates are concatenated into vectors of
// total = sum_c q_c doubles (cone c occupies [ptr[c], ptr[c+1]), head first).
// A handle owns kSocSlots such vectors (the caller decides what each slot
// holds: slack, dual, scaling point, lambda, directions, scratch) plus the
// per-cone NT scalars eta, eta^2 and status flags.
//
// Every kernel is bandwidth-bound: it streams its input slots once from HBM
// (a cone's repeated passes hit L1/L2) and writes its outputs once.  The
// arithmetic follows the reference's operation order exactly and the file is
// compiled without floating-point contraction, so results are reproducible
// operation for operation.
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/sdpx_mi355x.h"
#include "common.hpp"

#pragma clang fp contract(off)

namespace sdpx {
namespace {

constexpr int kSocSlots = SDPX_SOC_SLOTS;
constexpr int kStrictCandidates = 25;
constexpr int kFixedQMax = 8;  // uniform cone dimensions with register-resident kernels: 2-6, 8  // one full step + 24 halvings (soc_native.jl:2705-2745)

struct Cone {
  const long long* ptr;
  int ncones;
};

__device__ __forceinline__ void cone_range(const long long* ptr, int c, long long& b, int& q) {
  b = ptr[c];
  q = (int)(ptr[c + 1] - b);
}

// _soc_tail_norm (soc_lorentz_kernels.jl:22-34)
__device__ __forceinline__ double tail_norm(const double* v, int q) {
  double scale = 0.0;
  for (int i = 1; i < q; ++i) scale = fmax(scale, fabs(v[i]));
  if (scale == 0.0) return scale;
  double ss = 0.0;
  for (int i = 1; i < q; ++i) {
    const double t = v[i] / scale;
    ss += t * t;
  }
  return scale * sqrt(ss);
}

// tail norm of s + t*ds without materialising the trial point
__device__ double trial_tail_norm(const double* s, const double* ds, double t, int q) {
  double scale = 0.0;
  for (int i = 1; i < q; ++i) scale = fmax(scale, fabs(s[i] + t * ds[i]));
  if (scale == 0.0) return scale;
  double ss = 0.0;
  for (int i = 1; i < q; ++i) {
    const double v = (s[i] + t * ds[i]) / scale;
    ss += v * v;
  }
  return scale * sqrt(ss);
}

// _soc_is_interior (soc.jl:122-123)
__device__ bool is_interior(const double* v, int q) {
  return v[0] > 0.0 && (v[0] - tail_norm(v, q)) > 0.0;
}

__device__ bool trial_is_interior(const double* s, const double* ds, double t, int q) {
  const double h = s[0] + t * ds[0];
  return h > 0.0 && (h - trial_tail_norm(s, ds, t, q)) > 0.0;
}

// _soc_determinant (soc.jl:114-120)
__device__ double determinant(const double* v, int q) {
  double d = v[0] * v[0];
  for (int i = 1; i < q; ++i) d -= v[i] * v[i];
  return d;
}

// _soc_sqrt_determinant (soc_lorentz_kernels.jl:42-62)
__device__ __forceinline__ bool sqrt_determinant(const double* v, int q, double& r) {
  r = 0.0;
  double scale = 0.0;
  for (int i = 0; i < q; ++i) {
    if (!isfinite(v[i])) return false;
    scale = fmax(scale, fabs(v[i]));
  }
  if (scale == 0.0) return false;
  const double head = v[0];
  const double tail = tail_norm(v, q);
  const double mg = head - tail;
  if (!(mg > 0.0)) return false;
  const double residual = (mg / scale) * (head / scale + tail / scale);
  if (!(residual > 0.0)) return false;
  const double res = scale * sqrt(residual);
  r = res;
  return isfinite(res) && res > 0.0;
}

// Julia's max(a, b) for the finite values these kernels compare
__device__ __forceinline__ double jmax(double a, double b) { return a > b ? a : b; }

// _soc_nt_scaling! (soc_lorentz_kernels.jl:72-184).  The a, b, raw vectors
// of the reference are recomputed from s and z where needed (same values).
// Returns false without writing w / lambda when any check fails.
__device__ __forceinline__ bool nt_scaling(const double* s, const double* z, int q, double* w, double* lam,
                                           double& eta_out, double& eta2_out) {
  eta_out = 0.0;
  eta2_out = 0.0;
  double ss, zs;
  const bool sok = sqrt_determinant(s, q, ss);
  const bool zok = sqrt_determinant(z, q, zs);
  if (!(sok && zok)) return false;
  const double eta = sqrt(ss) / sqrt(zs);
  const double eta2 = eta * eta;
  if (!(isfinite(eta2) && eta2 > 0.0)) return false;
  // a = s / ss, b = z / zs, raw = (a0 + b0, a_tail - b_tail)
  double ms = 1.0;
  for (int i = 0; i < q; ++i) {
    const double a = s[i] / ss, b = z[i] / zs;
    const double raw = i == 0 ? a + b : a - b;
    if (!(isfinite(a) && isfinite(b) && isfinite(raw))) return false;
    ms = jmax(ms, jmax(fabs(a), fabs(b)));
  }
  const double ims = 1.0 / ms;
  double cross = ((s[0] / ss) * ims) * ((z[0] / zs) * ims);
  for (int i = 1; i < q; ++i) cross += ((s[i] / ss) * ims) * ((z[i] / zs) * ims);
  const double radicand = ims * ims + cross;
  if (!(isfinite(radicand) && radicand > 0.0)) return false;
  const double wscale = sqrt(2.0) * ms * sqrt(radicand);
  if (!(isfinite(wscale) && wscale > 0.0)) return false;
  double nts = 1.0;
  for (int i = 1; i < q; ++i) nts = jmax(nts, fabs(((s[i] / ss) - (z[i] / zs)) / wscale));
  const double its = 1.0 / nts;
  double acc = its * its;
  for (int i = 1; i < q; ++i) {
    const double t = (((s[i] / ss) - (z[i] / zs)) / wscale) / nts;
    acc += t * t;
  }
  const double nhead = nts * sqrt(acc);
  if (!(isfinite(nhead) && nhead > 0.0)) return false;
  const double gamma = wscale / 2.0;
  const double a0 = s[0] / ss, b0 = z[0] / zs;
  double ls = jmax(jmax(jmax(1.0, fabs(a0)), fabs(b0)), fabs(gamma));
  for (int i = 1; i < q; ++i) ls = jmax(ls, jmax(fabs(s[i] / ss), fabs(z[i] / zs)));
  const double ils = 1.0 / ls;
  const double aah = a0 * ils, bbh = b0 * ils, gg = gamma * ils;
  const double den = aah + bbh + 2.0 * gg;
  if (!(isfinite(den) && den > 0.0)) return false;
  const double geo = sqrt(ss) * sqrt(zs);
  const double lhead = geo * gamma;
  if (!isfinite(lhead)) return false;
  for (int i = 1; i < q; ++i) {
    const double aa = (s[i] / ss) * ils, bb = (z[i] / zs) * ils;
    const double scaled = (gg + bbh) * aa + (gg + aah) * bb;
    const double v = geo * (scaled / den) * ls;
    if (!isfinite(v)) return false;
  }
  // every check passed: commit
  w[0] = nhead;
  lam[0] = lhead;
  for (int i = 1; i < q; ++i) {
    const double aa = (s[i] / ss) * ils, bb = (z[i] / zs) * ils;
    const double scaled = (gg + bbh) * aa + (gg + aah) * bb;
    lam[i] = geo * (scaled / den) * ls;
    w[i] = ((s[i] / ss) - (z[i] / zs)) / wscale;
  }
  eta_out = eta;
  eta2_out = eta2;
  return true;
}

__global__ void nt_scaling_kernel(Cone cn, const double* __restrict__ S, const double* __restrict__ Z, double* W,
                                  double* L, double* eta, double* eta2, int* ok, int* nfail) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  double e = 0.0, e2 = 0.0;
  const bool good = nt_scaling(S + b, Z + b, q, W + b, L + b, e, e2);
  eta[c] = e;
  eta2[c] = e2;
  ok[c] = good ? 1 : 0;
  if (!good) atomicAdd(nfail, 1);
}

// op 0: W (soc_lorentz_kernels.jl:187-205), 1: W^-1 (:208-226), 2: (W'W)^-1 (:229-248).
// dst may alias src (the source cone is read into the output cone in order:
// every output entry depends only on src[i] and the two reductions).
__device__ __forceinline__ void apply_cone(int op, const double* w, double e, double e2, const double* x, int q,
                                           double* y) {
  const double head = x[0];
  if (op == 2) {
    double ld = w[0] * head;
    for (int i = 1; i < q; ++i) ld -= w[i] * x[i];
    y[0] = (2.0 * w[0] * ld - head) / e2;
    for (int i = 1; i < q; ++i) y[i] = (-2.0 * w[i] * ld + x[i]) / e2;
    return;
  }
  double dt = 0.0;
  for (int i = 1; i < q; ++i) dt += w[i] * x[i];
  if (op == 0) {
    const double corr = head + dt / (1.0 + w[0]);
    y[0] = e * (w[0] * head + dt);
    for (int i = 1; i < q; ++i) y[i] = e * (x[i] + corr * w[i]);
  } else {
    const double corr = -head + dt / (1.0 + w[0]);
    y[0] = (w[0] * head - dt) / e;
    for (int i = 1; i < q; ++i) y[i] = (x[i] + corr * w[i]) / e;
  }
}

__global__ void apply_kernel(Cone cn, int op, const double* __restrict__ W, const double* eta,
                             const double* eta2, const double* src, double* dst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  // the source cone is read in order before each output entry is written, so dst may alias src
  apply_cone(op, W + b, op == 2 ? 0.0 : eta[c], op == 2 ? eta2[c] : 0.0, src + b, q, dst + b);
}

// ---- every cone of the same dimension Q (the common case: SOCP constraints of
// one shape): no offset table, the cone lives in registers, 16-byte accesses
template <int Q>
__device__ __forceinline__ void load_cone(const double* __restrict__ p, double* v) {
  if constexpr (Q % 2 == 0) {
#pragma unroll
    for (int i = 0; i < Q; i += 2) {
      const double2 t = *reinterpret_cast<const double2*>(p + i);
      v[i] = t.x;
      v[i + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) v[i] = p[i];
  }
}

template <int Q>
__device__ __forceinline__ void store_cone(double* p, const double* v) {
  if constexpr (Q % 2 == 0) {
#pragma unroll
    for (int i = 0; i < Q; i += 2) *reinterpret_cast<double2*>(p + i) = make_double2(v[i], v[i + 1]);
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) p[i] = v[i];
  }
}

template <int Q>
__global__ void nt_scaling_fixed_kernel(int ncones, const double* __restrict__ S, const double* __restrict__ Z,
                                        double* W, double* L, double* eta, double* eta2, int* ok, int* nfail) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncones) return;
  const long long b = (long long)c * Q;
  double s[Q], z[Q], w[Q], lam[Q];
  load_cone<Q>(S + b, s);
  load_cone<Q>(Z + b, z);
  double e = 0.0, e2 = 0.0;
  const bool good = nt_scaling(s, z, Q, w, lam, e, e2);
  if (good) {
    store_cone<Q>(W + b, w);
    store_cone<Q>(L + b, lam);
  } else {
    atomicAdd(nfail, 1);
  }
  eta[c] = e;
  eta2[c] = e2;
  ok[c] = good ? 1 : 0;
}

template <int Q>
__global__ void apply_fixed_kernel(int ncones, int op, const double* __restrict__ W, const double* eta,
                                   const double* eta2, const double* src, double* dst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncones) return;
  const long long b = (long long)c * Q;
  double w[Q], x[Q], y[Q];
  load_cone<Q>(W + b, w);
  load_cone<Q>(src + b, x);
  apply_cone(op, w, op == 2 ? 0.0 : eta[c], op == 2 ? eta2[c] : 0.0, x, Q, y);
  store_cone<Q>(dst + b, y);
}

// _soc_jordan! (soc.jl:94-112); dst may alias either operand
__global__ void jordan_kernel(Cone cn, const double* left, const double* right, double* dst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  const double* l = left + b;
  const double* r = right + b;
  double* d = dst + b;
  const double lh = l[0], rh = r[0];
  double v = lh * rh;
  for (int i = 1; i < q; ++i) v += l[i] * r[i];
  for (int i = 1; i < q; ++i) d[i] = lh * r[i] + rh * l[i];
  d[0] = v;
}

// _soc_jordan_solve! (soc_lorentz_kernels.jl:464-485); cones whose divisor is
// not interior are counted and left untouched (the reference throws)
__global__ void jordan_solve_kernel(Cone cn, const double* left, const double* right, double* dst, int* nbad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  const double* l = left + b;
  const double* r = right + b;
  double* d = dst + b;
  if (!is_interior(l, q)) {
    atomicAdd(nbad, 1);
    return;
  }
  const double lh = l[0], rh = r[0];
  const double det = determinant(l, q);
  double td = 0.0;
  for (int i = 1; i < q; ++i) td += l[i] * r[i];
  const double rhead = (lh * rh - td) / det;
  const double ts = (td / lh - rh) / det;
  for (int i = 1; i < q; ++i) d[i] = ts * l[i] + r[i] / lh;
  d[0] = rhead;
}

// _soc_inverse! (soc.jl:126-138)
__global__ void inverse_kernel(Cone cn, const double* src, double* dst, int* nbad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  const double* x = src + b;
  double* d = dst + b;
  if (!is_interior(x, q)) {
    atomicAdd(nbad, 1);
    return;
  }
  const double det = determinant(x, q);
  const double head = x[0];
  for (int i = 1; i < q; ++i) d[i] = -x[i] / det;
  d[0] = head / det;
}

// _soc_fraction_to_boundary (soc.jl:148-214)
__device__ double fraction_to_boundary(const double* s, const double* ds, int q) {
  double scale = 0.0;
  for (int i = 0; i < q; ++i) scale = jmax(scale, jmax(fabs(s[i]), fabs(ds[i])));
  if (scale == 0.0) return 1.0;
  const double s0 = s[0] / scale, d0 = ds[0] / scale;
  double c0 = s0 * s0;
  double c1 = 2.0 * s0 * d0;
  double c2 = d0 * d0;
  const double ff = s0 + d0;
  double fdet = ff * ff;
  for (int i = 1; i < q; ++i) {
    const double si = s[i] / scale, di = ds[i] / scale;
    c0 -= si * si;
    c1 -= 2.0 * si * di;
    c2 -= di * di;
    const double fv = si + di;
    fdet -= fv * fv;
  }
  if (s0 == 0.0 && d0 < 0.0) return 0.0;
  if (c0 == 0.0 && (c1 < 0.0 || (c1 == 0.0 && c2 < 0.0))) return 0.0;
  if (ff >= 0.0 && fdet >= 0.0) return 1.0;
  double root = 1.0;
  if (c2 == 0.0) {
    if (c1 < 0.0) root = fmin(root, -c0 / c1);
  } else {
    const double disc = jmax(c1 * c1 - 4.0 * c2 * c0, 0.0);
    const double sq = sqrt(disc);
    const double qq = c1 >= 0.0 ? -(c1 + sq) / 2.0 : -(c1 - sq) / 2.0;
    if (qq != 0.0) {
      const double first = qq / c2;
      const double second = c0 / qq;
      if (0.0 < first && first <= root) root = first;
      if (0.0 < second && second <= root) root = second;
    } else {
      const double rep = -c1 / (2.0 * c2);
      if (0.0 < rep && rep <= root) root = rep;
    }
  }
  if (d0 < 0.0) root = fmin(root, -s0 / d0);
  // clamp(root, 0, 1) keeps NaN
  return root > 1.0 ? 1.0 : (root < 0.0 ? 0.0 : root);
}

__global__ void ftb_kernel(Cone cn, const double* S, const double* dS, double* bound) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  bound[c] = fraction_to_boundary(S + b, dS + b, q);
}

// min over n values in a fixed order (one block of 1024 threads, strided
// partial minima then a tree); NaN propagates
__global__ void __launch_bounds__(1024) min_reduce_kernel(const double* v, int n, double* out) {
  __shared__ double red[1024];
  double m = INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double x = v[i];
    m = (x != x || m != m) ? NAN : fmin(m, x);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double a = red[threadIdx.x], o = red[threadIdx.x + s];
      red[threadIdx.x] = (a != a || o != o) ? NAN : fmin(a, o);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// per cone: margin v0 - ||v_tail|| (soc_lorentz_kernels.jl:36) and interior flag
__global__ void margin_kernel(Cone cn, const double* V, double* margin, int* ninterior) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  const double* v = V + b;
  const double mg = v[0] - tail_norm(v, q);
  margin[c] = mg;
  if (v[0] > 0.0 && mg > 0.0) atomicAdd(ninterior, 1);
}

// _native_soc_strict_step (soc_native.jl:2705-2745): flags[j] stays 1 while
// every cone's trial point state + steps[j]*direction is interior
__global__ void strict_step_kernel(Cone cn, const double* S, const double* dS, const double* steps, int nsteps,
                                   int* flags) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cn.ncones) return;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  for (int j = 0; j < nsteps; ++j)
    if (!trial_is_interior(S + b, dS + b, steps[j], q)) atomicAnd(flags + j, 0);
}

// ---- fixed-trace Q3 kernels (soc_lorentz_kernels.jl:251-461) -----------------
__device__ __forceinline__ void sym2_inverse(double x0, double x1, double x2, double& i11, double& i12,
                                             double& i22) {
  const double d = x0 * x0 - x1 * x1 - x2 * x2;
  i11 = (x0 - x1) / d;
  i12 = -x2 / d;
  i22 = (x0 + x1) / d;
}

// _soc_fixed_trace_hkm_metric! (soc_lorentz_kernels.jl:353-370)
__global__ void hkm_metric_kernel(int n, const double* Xp, const double* Zd, double* D, int* nfail) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double* x = Xp + 3LL * c;
  const double* z = Zd + 3LL * c;
  const double x0 = x[0], x1 = x[1], x2 = x[2];
  const double z0 = z[0], z1 = z[1], z2 = z[2];
  const double tn = sqrt(x1 * x1 + x2 * x2);
  const double d = x0 * x0 - x1 * x1 - x2 * x2;
  if (!(x0 > tn) || !(d > 0.0)) {
    atomicAdd(nfail, 1);
    return;
  }
  double* o = D + 3LL * c;
  o[0] = (x0 * z0 - x1 * z1 + x2 * z2) / d;
  o[1] = -(x1 * z2 + x2 * z1) / d;
  o[2] = (x0 * z0 + x1 * z1 - x2 * z2) / d;
}

// R = target I - X Y (- dXa dYa) in the 2x2 image, Y = Z / 2
__device__ __forceinline__ void hkm_target(double x11, double x12, double x22, double y11, double y12, double y22,
                                           double target, bool aff, const double* ap, const double* ad, double& r11,
                                           double& r12, double& r21, double& r22) {
  r11 = target - (x11 * y11 + x12 * y12);
  r12 = -(x11 * y12 + x12 * y22);
  r21 = -(x12 * y11 + x22 * y12);
  r22 = target - (x12 * y12 + x22 * y22);
  if (aff) {
    const double dx11 = ap[0] + ap[1], dx12 = ap[2], dx22 = ap[0] - ap[1];
    const double dy11 = (ad[0] + ad[1]) / 2, dy12 = ad[2] / 2, dy22 = (ad[0] - ad[1]) / 2;
    r11 -= dx11 * dy11 + dx12 * dy12;
    r12 -= dx11 * dy12 + dx12 * dy22;
    r21 -= dx12 * dy11 + dx22 * dy12;
    r22 -= dx12 * dy12 + dx22 * dy22;
  }
}

// mode 0: _soc_fixed_trace_hkm_rhs_coordinates (soc_lorentz_kernels.jl:372-418), V = primal residual
// mode 1: _soc_fixed_trace_hkm_recovery!        (soc_lorentz_kernels.jl:420-461), V = primal direction
__global__ void hkm_contract_kernel(int n, int mode, const double* Xp, const double* Zd, const double* V,
                                    const double* AP, const double* AD, double target, int aff, double* D) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const long long o3 = 3LL * c;
  const double x0 = Xp[o3], x1 = Xp[o3 + 1], x2 = Xp[o3 + 2];
  const double z0 = Zd[o3], z1 = Zd[o3 + 1], z2 = Zd[o3 + 2];
  const double v0 = V[o3], v1 = V[o3 + 1], v2 = V[o3 + 2];
  const double x11 = x0 + x1, x12 = x2, x22 = x0 - x1;
  const double y11 = (z0 + z1) / 2, y12 = z2 / 2, y22 = (z0 - z1) / 2;
  const double v11 = v0 + v1, v12 = v2, v22 = v0 - v1;
  double r11, r12, r21, r22;
  hkm_target(x11, x12, x22, y11, y12, y22, target, aff != 0, aff ? AP + o3 : nullptr, aff ? AD + o3 : nullptr, r11,
             r12, r21, r22);
  double w11, w12, w21, w22;
  if (mode == 0) {
    w11 = v11 * y11 + v12 * y12 - r11;
    w12 = v11 * y12 + v12 * y22 - r12;
    w21 = v12 * y11 + v22 * y12 - r21;
    w22 = v12 * y12 + v22 * y22 - r22;
  } else {
    w11 = r11 - (v11 * y11 + v12 * y12);
    w12 = r12 - (v11 * y12 + v12 * y22);
    w21 = r21 - (v12 * y11 + v22 * y12);
    w22 = r22 - (v12 * y12 + v22 * y22);
  }
  double i11, i12, i22;
  sym2_inverse(x0, x1, x2, i11, i12, i22);
  const double q11 = i11 * w11 + i12 * w21;
  const double q12 = i11 * w12 + i12 * w22;
  const double q21 = i12 * w11 + i22 * w21;
  const double q22 = i12 * w12 + i22 * w22;
  D[o3] = q11 + q22;
  D[o3 + 1] = q11 - q22;
  D[o3 + 2] = q12 + q21;
}

__global__ void zero_ints_kernel(int* p, int n, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// ---- metric assembly (soc_native.jl:1119-1218) --------------------------------
// Atil = W^-1 A for every cone c and column j of the coefficient slab A
// ((sum q) x n, column-major, cone c on rows [ptr[c], ptr[c+1])).  The Gram
// product Atil' Atil then equals sum_c A_c' (W'W)^-1 A_c: W^-1 is symmetric,
// so (W'W)^-1 = W^-1 W^-1 = W^-T W^-1.  Cones vary fastest within a column so
// a wavefront reads consecutive coordinates.
__global__ void winv_columns_kernel(Cone cn, const double* __restrict__ W, const double* __restrict__ eta,
                                    const double* __restrict__ A, long long lda, long long ncols,
                                    double* __restrict__ At) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)cn.ncones * ncols) return;
  const int c = (int)(idx % cn.ncones);
  const long long j = idx / cn.ncones;
  long long b;
  int q;
  cone_range(cn.ptr, c, b, q);
  apply_cone(1, W + b, eta[c], 0.0, A + b + j * lda, q, At + b + j * lda);
}

// strict upper triangle of the n x n matrix := its lower triangle
__global__ void mirror_lower_kernel(double* H, int n) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r < c) H[r + (long long)c * n] = H[c + (long long)r * n];
}

__global__ void count_not_ok_kernel(const int* ok, int n, int* count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n && !ok[c]) atomicAdd(count, 1);
}

}  // namespace
}  // namespace sdpx

using namespace sdpx;

struct sdpx_soc {
  int device = 0;
  int ncones = 0;
  long long total = 0;
  int maxq = 0;
  bool all_q3 = false;
  int uniform_q = 0;  // every cone has this dimension (2..kFixedQMax), else 0
  hipStream_t stream = nullptr;
  long long* ptr = nullptr;
  double* vec[kSocSlots] = {};
  double *eta = nullptr, *eta2 = nullptr, *bound = nullptr, *steps = nullptr, *dmin = nullptr;
  int* ok = nullptr;
  // metric assembly: coefficient slab A ((sum q) x nvars), W^-1 A, the Hessian
  // (nvars x nvars) and the Gram kernel's split-K partials and work list
  int nvars = 0, metric_split = 1;
  double *A = nullptr, *At = nullptr, *H = nullptr, *metric_ws = nullptr;
  int2* metric_items = nullptr;
  int* counters = nullptr;  // [0] failures / bad cones / interior count, [1..25] strict-step flags
  int* hcounters = nullptr;
  double* hscalar = nullptr;
  long long bytes = 0;
  std::vector<void*> allocs;
  hipEvent_t ev[2] = {};
  bool prof = false;
  float last_ms = 0.f;
};

namespace {

int soc_alloc(sdpx_soc_t* h, void** p, size_t bytes) {
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

int soc_check(sdpx_soc_t* h, int slot = 0) {
  if (!h) {
    set_last_error("null SOC handle");
    return kInvalidArgument;
  }
  if (slot < 0 || slot >= kSocSlots) {
    set_last_error("SOC slot out of range");
    return kInvalidArgument;
  }
  SDPX_HIP(hipSetDevice(h->device));
  return kOk;
}

inline dim3 soc_grid(int n) { return dim3((unsigned)std::max(1, ceil_div(n, 256))); }

// Calls f(std::integral_constant<int, q>) for a supported uniform cone
// dimension q; returns false (nothing launched) otherwise.
template <class F>
bool launch_fixed_q(int q, F&& f) {
  switch (q) {
    case 2: f(std::integral_constant<int, 2>{}); return true;
    case 3: f(std::integral_constant<int, 3>{}); return true;
    case 4: f(std::integral_constant<int, 4>{}); return true;
    case 5: f(std::integral_constant<int, 5>{}); return true;
    case 6: f(std::integral_constant<int, 6>{}); return true;
    case 8: f(std::integral_constant<int, 8>{}); return true;
    default: return false;
  }
}

int begin(sdpx_soc_t* h) {
  if (h->prof) SDPX_HIP(hipEventRecord(h->ev[0], h->stream));
  return kOk;
}

int end(sdpx_soc_t* h) {
  SDPX_HIP(hipGetLastError());
  if (h->prof) {
    SDPX_HIP(hipEventRecord(h->ev[1], h->stream));
    SDPX_HIP(hipEventSynchronize(h->ev[1]));
    SDPX_HIP(hipEventElapsedTime(&h->last_ms, h->ev[0], h->ev[1]));
  }
  return kOk;
}

int reset_counters(sdpx_soc_t* h, int n, int value) {
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, h->stream, h->counters, n, value);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int read_counter(sdpx_soc_t* h, int* out, int n = 1) {
  SDPX_HIP(hipMemcpyAsync(h->hcounters, h->counters, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  for (int i = 0; i < n; ++i) out[i] = h->hcounters[i];
  return kOk;
}

}  // namespace

// Scratch of sdpx_soc_fraction_to_boundary_noalloc: the per-cone bounds that
// the fixed-order minimum reads.  Each device holds its own copy from module
// load until the process ends; every call on a device shares it.
static __device__ __align__(256) double scratch[SDPX_SOC_NOALLOC_MAX_CONES];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device (resolved once per device;
// the array's host-side name is not a device address).
int noalloc_scratch(int device, double** out) {
  if (device < 0 || device >= kNoAllocMaxDevices) {
    set_last_error("sdpx_soc_fraction_to_boundary_noalloc: device index out of range");
    return kInvalidArgument;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return kHipError;
    }
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return kOk;
}

}  // namespace

extern "C" {

int sdpx_soc_create(int device, int ncones, const int* dims, sdpx_soc_t** out) {
  if (!out || ncones <= 0 || !dims) {
    set_last_error("sdpx_soc_create: invalid arguments");
    return kInvalidArgument;
  }
  *out = nullptr;
  std::vector<long long> ptr((size_t)ncones + 1, 0);
  int maxq = 0;
  bool all3 = true;
  for (int c = 0; c < ncones; ++c) {
    if (dims[c] < 1) {
      set_last_error("sdpx_soc_create: cone dimensions must be positive");
      return kInvalidArgument;
    }
    ptr[c + 1] = ptr[c] + dims[c];
    maxq = std::max(maxq, dims[c]);
    all3 = all3 && dims[c] == 3;
  }
  if (hipSetDevice(device) != hipSuccess) {
    set_last_error("sdpx_soc_create: hipSetDevice failed");
    return kHipError;
  }
  auto* h = new sdpx_soc();
  h->device = device;
  h->ncones = ncones;
  h->total = ptr[ncones];
  h->maxq = maxq;
  h->all_q3 = all3;
  {
    bool uniform = true;
    for (int c = 1; c < ncones && uniform; ++c) uniform = dims[c] == dims[0];
    h->uniform_q = (uniform && dims[0] >= 2 && dims[0] <= kFixedQMax) ? dims[0] : 0;
  }
  auto fail = [&](int s) {
    sdpx_soc_destroy(h);
    return s;
  };
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    set_last_error("hipStreamCreate failed");
    return fail(kHipError);
  }
  for (auto& e : h->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(kHipError);
  int s;
  if ((s = soc_alloc(h, reinterpret_cast<void**>(&h->ptr), sizeof(long long) * ptr.size()))) return fail(s);
  for (int i = 0; i < kSocSlots; ++i)
    if ((s = soc_alloc(h, reinterpret_cast<void**>(&h->vec[i]), sizeof(double) * (size_t)h->total)))
      return fail(s);
  double** scal[] = {&h->eta, &h->eta2, &h->bound};
  for (double** p : scal)
    if ((s = soc_alloc(h, reinterpret_cast<void**>(p), sizeof(double) * (size_t)ncones))) return fail(s);
  if ((s = soc_alloc(h, reinterpret_cast<void**>(&h->steps), sizeof(double) * kStrictCandidates))) return fail(s);
  if ((s = soc_alloc(h, reinterpret_cast<void**>(&h->dmin), sizeof(double)))) return fail(s);
  if ((s = soc_alloc(h, reinterpret_cast<void**>(&h->ok), sizeof(int) * (size_t)ncones))) return fail(s);
  if ((s = soc_alloc(h, reinterpret_cast<void**>(&h->counters), sizeof(int) * 64))) return fail(s);
  if (hipHostMalloc(reinterpret_cast<void**>(&h->hcounters), sizeof(int) * 64) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&h->hscalar), sizeof(double) * 64) != hipSuccess) {
    set_last_error("hipHostMalloc failed");
    return fail(kOutOfMemory);
  }
  if (hipMemcpy(h->ptr, ptr.data(), sizeof(long long) * ptr.size(), hipMemcpyHostToDevice) != hipSuccess) {
    set_last_error("hipMemcpy(cone offsets) failed");
    return fail(kHipError);
  }
  for (int i = 0; i < kSocSlots; ++i)
    if (hipMemset(h->vec[i], 0, sizeof(double) * (size_t)h->total) != hipSuccess) return fail(kHipError);
  // no NT scaling yet: every cone reads as not scaled
  if (hipMemset(h->ok, 0, sizeof(int) * (size_t)ncones) != hipSuccess ||
      hipMemset(h->eta, 0, sizeof(double) * (size_t)ncones) != hipSuccess ||
      hipMemset(h->eta2, 0, sizeof(double) * (size_t)ncones) != hipSuccess)
    return fail(kHipError);
  *out = h;
  return kOk;
}

int sdpx_soc_destroy(sdpx_soc_t* h) {
  if (!h) return kOk;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (void* p : h->allocs) hipFree(p);
  if (h->hcounters) hipHostFree(h->hcounters);
  if (h->hscalar) hipHostFree(h->hscalar);
  for (auto& e : h->ev)
    if (e) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return kOk;
}

int sdpx_soc_memory(sdpx_soc_t* h, long long* bytes) {
  SDPX_CHECK(soc_check(h));
  if (bytes) *bytes = h->bytes;
  return kOk;
}

int sdpx_soc_set_vector(sdpx_soc_t* h, int slot, const double* v) {
  SDPX_CHECK(soc_check(h, slot));
  SDPX_HIP(hipMemcpyAsync(h->vec[slot], v, sizeof(double) * (size_t)h->total, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_soc_get_vector(sdpx_soc_t* h, int slot, double* v) {
  SDPX_CHECK(soc_check(h, slot));
  SDPX_HIP(hipMemcpyAsync(v, h->vec[slot], sizeof(double) * (size_t)h->total, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_soc_copy_vector(sdpx_soc_t* h, int src, int dst) {
  SDPX_CHECK(soc_check(h, src));
  SDPX_CHECK(soc_check(h, dst));
  if (src == dst) return kOk;
  SDPX_HIP(hipMemcpyAsync(h->vec[dst], h->vec[src], sizeof(double) * (size_t)h->total, hipMemcpyDeviceToDevice,
                          h->stream));
  return kOk;
}

int sdpx_soc_get_scaling(sdpx_soc_t* h, double* eta, double* eta2, int* ok) {
  SDPX_CHECK(soc_check(h));
  const size_t n = (size_t)h->ncones;
  if (eta) SDPX_HIP(hipMemcpyAsync(eta, h->eta, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  if (eta2) SDPX_HIP(hipMemcpyAsync(eta2, h->eta2, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  if (ok) SDPX_HIP(hipMemcpyAsync(ok, h->ok, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_soc_nt_scaling(sdpx_soc_t* h, int s, int z, int w, int lambda, int* nfailed) {
  SDPX_CHECK(soc_check(h, s));
  SDPX_CHECK(soc_check(h, z));
  SDPX_CHECK(soc_check(h, w));
  SDPX_CHECK(soc_check(h, lambda));
  if (w == lambda || w == s || w == z || lambda == s || lambda == z) {
    set_last_error("sdpx_soc_nt_scaling: outputs must not alias inputs or each other");
    return kInvalidArgument;
  }
  SDPX_CHECK(reset_counters(h, 1, 0));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  if (!launch_fixed_q(h->uniform_q, [&](auto qc) {
        constexpr int Q = decltype(qc)::value;
        hipLaunchKernelGGL(nt_scaling_fixed_kernel<Q>, soc_grid(h->ncones), dim3(256), 0, h->stream, h->ncones,
                           h->vec[s], h->vec[z], h->vec[w], h->vec[lambda], h->eta, h->eta2, h->ok, h->counters);
      }))
    hipLaunchKernelGGL(nt_scaling_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, h->vec[s], h->vec[z],
                       h->vec[w], h->vec[lambda], h->eta, h->eta2, h->ok, h->counters);
  SDPX_CHECK(end(h));
  int nf = 0;
  SDPX_CHECK(read_counter(h, &nf));
  if (nfailed) *nfailed = nf;
  return kOk;
}

int sdpx_soc_apply(sdpx_soc_t* h, int op, int w, int src, int dst) {
  SDPX_CHECK(soc_check(h, w));
  SDPX_CHECK(soc_check(h, src));
  SDPX_CHECK(soc_check(h, dst));
  if (op < 0 || op > 2 || dst == w) {
    set_last_error("sdpx_soc_apply: op must be 0..2 and the destination must not alias the scaling point");
    return kInvalidArgument;
  }
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  if (!launch_fixed_q(h->uniform_q, [&](auto qc) {
        constexpr int Q = decltype(qc)::value;
        hipLaunchKernelGGL(apply_fixed_kernel<Q>, soc_grid(h->ncones), dim3(256), 0, h->stream, h->ncones, op,
                           h->vec[w], h->eta, h->eta2, h->vec[src], h->vec[dst]);
      }))
    hipLaunchKernelGGL(apply_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, op, h->vec[w], h->eta,
                       h->eta2, h->vec[src], h->vec[dst]);
  return end(h);
}

int sdpx_soc_jordan(sdpx_soc_t* h, int left, int right, int dst) {
  SDPX_CHECK(soc_check(h, left));
  SDPX_CHECK(soc_check(h, right));
  SDPX_CHECK(soc_check(h, dst));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  hipLaunchKernelGGL(jordan_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, h->vec[left], h->vec[right],
                     h->vec[dst]);
  return end(h);
}

int sdpx_soc_jordan_solve(sdpx_soc_t* h, int left, int right, int dst, int* nbad) {
  SDPX_CHECK(soc_check(h, left));
  SDPX_CHECK(soc_check(h, right));
  SDPX_CHECK(soc_check(h, dst));
  if (dst == left) {
    set_last_error("sdpx_soc_jordan_solve: the destination must not alias the divisor");
    return kInvalidArgument;
  }
  SDPX_CHECK(reset_counters(h, 1, 0));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  hipLaunchKernelGGL(jordan_solve_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, h->vec[left],
                     h->vec[right], h->vec[dst], h->counters);
  SDPX_CHECK(end(h));
  int nb = 0;
  SDPX_CHECK(read_counter(h, &nb));
  if (nbad) *nbad = nb;
  return nb ? kNotPositiveDefinite : kOk;
}

int sdpx_soc_inverse(sdpx_soc_t* h, int src, int dst, int* nbad) {
  SDPX_CHECK(soc_check(h, src));
  SDPX_CHECK(soc_check(h, dst));
  SDPX_CHECK(reset_counters(h, 1, 0));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  hipLaunchKernelGGL(inverse_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, h->vec[src], h->vec[dst],
                     h->counters);
  SDPX_CHECK(end(h));
  int nb = 0;
  SDPX_CHECK(read_counter(h, &nb));
  if (nbad) *nbad = nb;
  return nb ? kNotPositiveDefinite : kOk;
}

int sdpx_soc_fraction_to_boundary(sdpx_soc_t* h, int s, int ds, double* min_bound, double* per_cone) {
  SDPX_CHECK(soc_check(h, s));
  SDPX_CHECK(soc_check(h, ds));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  hipLaunchKernelGGL(ftb_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, h->vec[s], h->vec[ds], h->bound);
  SDPX_CHECK(end(h));
  hipLaunchKernelGGL(min_reduce_kernel, dim3(1), dim3(1024), 0, h->stream, h->bound, h->ncones, h->dmin);
  SDPX_HIP(hipGetLastError());
  SDPX_HIP(hipMemcpyAsync(h->hscalar, h->dmin, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (per_cone)
    SDPX_HIP(hipMemcpyAsync(per_cone, h->bound, sizeof(double) * (size_t)h->ncones, hipMemcpyDeviceToHost,
                            h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  if (min_bound) *min_bound = h->hscalar[0];
  return kOk;
}

int sdpx_soc_margins(sdpx_soc_t* h, int slot, double* margins, int* ninterior) {
  SDPX_CHECK(soc_check(h, slot));
  SDPX_CHECK(reset_counters(h, 1, 0));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  hipLaunchKernelGGL(margin_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, h->vec[slot], h->bound,
                     h->counters);
  SDPX_CHECK(end(h));
  if (margins)
    SDPX_HIP(hipMemcpyAsync(margins, h->bound, sizeof(double) * (size_t)h->ncones, hipMemcpyDeviceToHost,
                            h->stream));
  int ni = 0;
  SDPX_CHECK(read_counter(h, &ni));
  if (ninterior) *ninterior = ni;
  return kOk;
}

int sdpx_soc_strict_step(sdpx_soc_t* h, int state, int direction, double bound, int allow_full_step,
                         double* step) {
  SDPX_CHECK(soc_check(h, state));
  SDPX_CHECK(soc_check(h, direction));
  // candidate sequence of the reference, same arithmetic
  const double safety = 99.0 / 100.0;
  double cand = bound > 0.0 ? bound : 0.0;
  cand = cand < 1.0 ? cand : 1.0;
  std::vector<double> steps;
  const bool full = allow_full_step && cand == 1.0;
  if (full) steps.push_back(1.0);
  double t = safety * cand;
  for (int i = 0; i < 24; ++i) {
    steps.push_back(t);
    t *= 1.0 / 2.0;
  }
  const int ns = (int)steps.size();
  SDPX_HIP(hipMemcpyAsync(h->steps, steps.data(), sizeof(double) * ns, hipMemcpyHostToDevice, h->stream));
  SDPX_CHECK(reset_counters(h, ns, 1));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  hipLaunchKernelGGL(strict_step_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, cn, h->vec[state],
                     h->vec[direction], h->steps, ns, h->counters);
  SDPX_CHECK(end(h));
  int flags[kStrictCandidates];
  SDPX_CHECK(read_counter(h, flags, ns));
  double result = 0.0;
  for (int j = 0; j < ns; ++j)
    if (flags[j]) {
      result = steps[j];
      break;
    }
  if (step) *step = result;
  return kOk;
}

int sdpx_soc_hkm_metric(sdpx_soc_t* h, int primal, int dual, int dst, int* nfailed) {
  SDPX_CHECK(soc_check(h, primal));
  SDPX_CHECK(soc_check(h, dual));
  SDPX_CHECK(soc_check(h, dst));
  if (!h->all_q3) {
    set_last_error("sdpx_soc_hkm_metric: fixed-trace kernels need every cone of dimension 3");
    return kInvalidArgument;
  }
  SDPX_CHECK(reset_counters(h, 1, 0));
  SDPX_CHECK(begin(h));
  hipLaunchKernelGGL(hkm_metric_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, h->ncones, h->vec[primal],
                     h->vec[dual], h->vec[dst], h->counters);
  SDPX_CHECK(end(h));
  int nf = 0;
  SDPX_CHECK(read_counter(h, &nf));
  if (nfailed) *nfailed = nf;
  return kOk;
}

int sdpx_soc_hkm_contract(sdpx_soc_t* h, int mode, int primal, int dual, int v, int affine_primal,
                          int affine_dual, double target, int include_affine, int dst) {
  SDPX_CHECK(soc_check(h, primal));
  SDPX_CHECK(soc_check(h, dual));
  SDPX_CHECK(soc_check(h, v));
  SDPX_CHECK(soc_check(h, dst));
  if (include_affine) {
    SDPX_CHECK(soc_check(h, affine_primal));
    SDPX_CHECK(soc_check(h, affine_dual));
  }
  if (!h->all_q3 || (mode != 0 && mode != 1)) {
    set_last_error("sdpx_soc_hkm_contract: mode 0/1 and every cone of dimension 3 required");
    return kInvalidArgument;
  }
  SDPX_CHECK(begin(h));
  hipLaunchKernelGGL(hkm_contract_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, h->ncones, mode,
                     h->vec[primal], h->vec[dual], h->vec[v], include_affine ? h->vec[affine_primal] : nullptr,
                     include_affine ? h->vec[affine_dual] : nullptr, target, include_affine, h->vec[dst]);
  return end(h);
}

int sdpx_soc_set_profiling(sdpx_soc_t* h, int enabled) {
  SDPX_CHECK(soc_check(h));
  h->prof = enabled != 0;
  return kOk;
}

int sdpx_soc_last_kernel_ms(sdpx_soc_t* h, double* ms) {
  SDPX_CHECK(soc_check(h));
  if (ms) *ms = (double)h->last_ms;
  return kOk;
}

int sdpx_soc_synchronize(sdpx_soc_t* h) {
  SDPX_CHECK(soc_check(h));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_soc_fraction_to_boundary_noalloc(int device, int ncones, const long long* offsets, const double* s,
                                          const double* ds, double* min_bound, double* per_cone, void* stream) {
  if (ncones <= 0 || !offsets || !s || !ds || !min_bound) {
    set_last_error("sdpx_soc_fraction_to_boundary_noalloc: invalid arguments");
    return kInvalidArgument;
  }
  if (!per_cone && ncones > SDPX_SOC_NOALLOC_MAX_CONES) {
    set_last_error("sdpx_soc_fraction_to_boundary_noalloc: " + std::to_string(ncones) +
                   " cones exceed the scratch capacity of " + std::to_string(SDPX_SOC_NOALLOC_MAX_CONES) +
                   " (pass per_cone to use caller memory instead)");
    return kInvalidArgument;
  }
  SDPX_HIP(hipSetDevice(device));
  double* bound = per_cone;
  if (!bound) SDPX_CHECK(noalloc_scratch(device, &bound));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Cone cn{offsets, ncones};
  hipLaunchKernelGGL(ftb_kernel, soc_grid(ncones), dim3(256), 0, st, cn, s, ds, bound);
  SDPX_HIP(hipGetLastError());
  hipLaunchKernelGGL(min_reduce_kernel, dim3(1), dim3(1024), 0, st, bound, ncones, min_bound);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int sdpx_soc_set_constraints(sdpx_soc_t* h, int n, const double* A) {
  SDPX_CHECK(soc_check(h));
  if (n < 1 || !A || h->nvars != 0) {
    set_last_error("sdpx_soc_set_constraints: n >= 1 and A required, once per handle");
    return kInvalidArgument;
  }
  const long long N = h->total;
  const int nt = ceil_div(n, 128);
  const long long tiles = (long long)nt * (nt + 1) / 2;
  // split-K so that the Gram launch fills two rounds of resident groups, each K range >= 256
  int split = (int)std::max<long long>(1, std::min<long long>(16, (1024 + tiles - 1) / tiles));
  split = (int)std::max<long long>(1, std::min<long long>(split, N / 256));
  while (split > 1 && (long long)split * n * n > (1LL << 28)) --split;  // partials within 2 GiB
  SDPX_CHECK(soc_alloc(h, (void**)&h->A, sizeof(double) * (size_t)(N * n)));
  SDPX_CHECK(soc_alloc(h, (void**)&h->At, sizeof(double) * (size_t)(N * n)));
  SDPX_CHECK(soc_alloc(h, (void**)&h->H, sizeof(double) * (size_t)n * n));
  if (split > 1) SDPX_CHECK(soc_alloc(h, (void**)&h->metric_ws, sizeof(double) * (size_t)split * n * n));
  std::vector<int2> items((size_t)tiles * split);
  gram_item_order(n, 8, split, items.data());
  SDPX_CHECK(soc_alloc(h, (void**)&h->metric_items, sizeof(int2) * items.size()));
  SDPX_HIP(hipMemcpyAsync(h->A, A, sizeof(double) * (size_t)(N * n), hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->metric_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice,
                          h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  h->nvars = n;
  h->metric_split = split;
  return kOk;
}

int sdpx_soc_metric(sdpx_soc_t* h, int w, const double* H0, double* H, int* nfailed) {
  SDPX_CHECK(soc_check(h, w));
  if (h->nvars == 0) {
    set_last_error("sdpx_soc_metric: no constraints (sdpx_soc_set_constraints)");
    return kInvalidArgument;
  }
  const int n = h->nvars;
  const long long N = h->total;
  // cones whose last NT scaling failed have no metric: report them, assemble nothing
  SDPX_CHECK(reset_counters(h, 1, 0));
  hipLaunchKernelGGL(count_not_ok_kernel, soc_grid(h->ncones), dim3(256), 0, h->stream, h->ok, h->ncones,
                     h->counters);
  int bad = 0;
  SDPX_CHECK(read_counter(h, &bad));
  if (nfailed) *nfailed = bad;
  if (bad) return 1;
  if (H0) SDPX_HIP(hipMemcpyAsync(h->H, H0, sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice, h->stream));
  SDPX_CHECK(begin(h));
  Cone cn{h->ptr, h->ncones};
  const long long work = (long long)h->ncones * n;
  hipLaunchKernelGGL(winv_columns_kernel, dim3((unsigned)ceil_div(work, 256)), dim3(256), 0, h->stream, cn,
                     h->vec[w], h->eta, h->A, N, (long long)n, h->At);
  SDPX_HIP(hipGetLastError());
  GemmArgs g{};
  g.M = g.N = n;
  g.K = (int)N;
  g.alpha = 1.0;
  g.beta = H0 ? 1.0 : 0.0;
  g.A = g.B = h->At;
  g.lda = g.ldb = N;
  g.C = h->H;
  g.ldc = n;
  g.batch = 1;
  g.flags = kGemmTransA | kGemmLowerOut | kGemmSchurGram;
  g.item_order = h->metric_items;
  SDPX_CHECK(gemm_f64(g, h->stream, h->metric_split, h->metric_ws, nullptr));
  hipLaunchKernelGGL(mirror_lower_kernel, dim3((unsigned)ceil_div((long long)n * n, 256)), dim3(256), 0, h->stream,
                     h->H, n);
  SDPX_CHECK(end(h));
  if (H) SDPX_HIP(hipMemcpyAsync(H, h->H, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

}  // extern "C"
