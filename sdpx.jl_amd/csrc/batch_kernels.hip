// Kernels of the batched route (see batch_kernels.hpp): one work-group of 256
// threads per listed member; k x k blocks (k <= 64) are staged in LDS,
// constraint panels are streamed from HBM with coalesced loads, the Schur
// panel products run on FP64 MFMA.
#include "batch_kernels.hpp"

namespace sdpx {
namespace batch {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---- block-level helpers ---------------------------------------------------------
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || b != b) ? NAN : fmax(a, b); }

// deterministic reductions over the 256 threads (red: >= 256 doubles of LDS)
__device__ double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = red[t] + red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ double block_max(double v, double* red) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = nan_max(red[t], red[t + s]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ double block_min(double v, double* red) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = fmin(red[t], red[t + s]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// s(r, c) = s[r + c * LP]
__device__ __forceinline__ void lds_load(double* s, const double* g, int k) {
  for (int e = threadIdx.x; e < k * k; e += NT) s[(e % k) + (e / k) * LP] = g[e];
}
__device__ __forceinline__ void lds_store(double* g, const double* s, int k) {
  for (int e = threadIdx.x; e < k * k; e += NT) g[e] = s[(e % k) + (e / k) * LP];
}
__device__ __forceinline__ void lds_store_lower(double* g, const double* s, int k) {
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    g[e] = r >= c ? s[r + c * LP] : 0.0;
  }
}

// C = alpha op(A) op(B) + beta C   (C must not alias A or B)
__device__ void lds_mm(double* C, const double* A, bool ta, const double* B, bool tb, int k, double alpha,
                       double beta) {
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    double acc = 0.0;
    for (int q = 0; q < k; ++q) {
      const double av = ta ? A[q + r * LP] : A[r + q * LP];
      const double bv = tb ? B[c + q * LP] : B[q + c * LP];
      acc += av * bv;
    }
    double* cp = C + r + c * LP;
    *cp = beta == 0.0 ? alpha * acc : alpha * acc + beta * (*cp);
  }
}

// In-place lower Cholesky of s (k x k): the same unscaled right-looking
// elimination and final scaling as the diagonal-block kernel of chol_f64.hip.
// Returns 0 or the 1-based column of the first non-positive / NaN pivot.
// The strict upper triangle of s is left as it was.
__device__ int lds_chol(double* s, int k, double* red) {
  const int tid = threadIdx.x;
  const int row = tid % KMAX, part = tid / KMAX;
  __syncthreads();
  for (int c = 0; c < k; ++c) {
    const double p = s[c + c * LP];
    if (!(p > 0.0)) {
      __syncthreads();
      return c + 1;
    }
    if (row > c && row < k) {
      const double f = s[row + c * LP] / p;
      for (int c2 = c + 1 + part; c2 <= row; c2 += 4) s[row + c2 * LP] -= f * s[c2 + c * LP];
    }
    __syncthreads();
  }
  if (tid < k) red[tid] = sqrt(s[tid + tid * LP]);
  __syncthreads();
  for (int e = tid; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    if (r > c) s[r + c * LP] = s[r + c * LP] / red[c];
  }
  __syncthreads();
  if (tid < k) s[tid + tid * LP] = red[tid];
  __syncthreads();
  return 0;
}

// Li = L^-1 (lower, strict upper zero) by forward substitution on the identity
// (the same summation as the diagonal-block inverse of chol_f64.hip)
__device__ void lds_trtri(const double* L, double* Li, int k, double* red4) {
  const int tid = threadIdx.x;
  const int c = tid % KMAX, part = tid / KMAX;
  for (int e = tid; e < k * k; e += NT) Li[(e % k) + (e / k) * LP] = 0.0;
  __syncthreads();
  for (int r = 0; r < k; ++r) {
    double partial = 0.0;
    if (c < r)
      for (int q = c + part; q < r; q += 4) partial += L[r + q * LP] * Li[q + c * LP];
    red4[part * KMAX + c] = partial;
    __syncthreads();
    if (part == 0 && c <= r)
      Li[r + c * LP] = (c == r) ? 1.0 / L[r + r * LP]
                                : -((red4[c] + red4[KMAX + c]) + (red4[2 * KMAX + c] + red4[3 * KMAX + c])) /
                                      L[r + r * LP];
    __syncthreads();
  }
}

// |diag| extrema of a factor in LDS (thread 0 result broadcast through red)
__device__ void lds_diag_extrema(const double* s, int k, double* red, double& mn, double& mx) {
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = __builtin_inf(), b = 0.0;
    for (int i = 0; i < k; ++i) {
      const double v = fabs(s[i + i * LP]);
      a = fmin(a, v);
      b = nan_max(b, v);
    }
    red[0] = a;
    red[1] = b;
  }
  __syncthreads();
  mn = red[0];
  mx = red[1];
  __syncthreads();
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---- HBM-bound constraint-panel passes ----------------------------------------------
// The panel passes split every member's m columns into chunks of kColChunk
// (grid.y = chunk; work-groups past a member's last chunk exit at once), so a
// large member streams through many CUs.  Av x accumulates per chunk into
// gpart (fixed chunk order in the combining kernel: deterministic).
__global__ __launch_bounds__(NT) void gemv_part_kernel(Arrays a, const int* list, int mode) {
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int i0 = blockIdx.y * kColChunk;
  if (i0 >= M.m) return;
  const int i1 = min(M.m, i0 + kColChunk);
  const long long k2 = (long long)M.k * M.k;
  const double* Av = a.Av + M.oAv;
  const double* vec = (mode == 1 || mode == 3 ? a.dx : a.x) + M.oM;
  double* part = a.gpart + M.oG + (long long)blockIdx.y * k2;
  for (long long e = threadIdx.x; e < k2; e += NT) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = i0;
    for (; i + 4 <= i1; i += 4) {
      s0 += Av[(long long)i * k2 + e] * vec[i];
      s1 += Av[(long long)(i + 1) * k2 + e] * vec[i + 1];
      s2 += Av[(long long)(i + 2) * k2 + e] * vec[i + 2];
      s3 += Av[(long long)(i + 3) * k2 + e] * vec[i + 3];
    }
    for (; i < i1; ++i) s0 += Av[(long long)i * k2 + e] * vec[i];
    part[e] = (s0 + s1) + (s2 + s3);
  }
}

// Sparse form of gemv_part: (sum_i vec_i A_i)_e gathered per position e from
// the position-ordered copy (vec in LDS).  A group of 16 lanes owns one
// position at a time: its entries are read 16 at a time (coalesced), each lane
// keeps a running sum and the group reduces them in a fixed tree, so the
// result is deterministic; it is written as the member's single partial.
// Work-groups along y take kSpPositions positions each, so large members spread
// over several CUs (the member's partial vector is written position by position).
constexpr int kSpPositions = 1024;
constexpr int kSpConstraints = 128;  // constraints per work-group of sp_gemvt_kernel

__global__ __launch_bounds__(NT) void sp_gemv_kernel(Arrays a, const int* list, int mode) {
  __shared__ double xs[MMAX];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k2 = M.k * M.k;
  const int e0 = blockIdx.y * kSpPositions, e1 = min(k2, e0 + kSpPositions);
  if (e0 >= k2) return;  // whole work-group
  const double* vec = (mode == 1 || mode == 3 ? a.dx : a.x) + M.oM;
  for (int i = threadIdx.x; i < M.m; i += NT) xs[i] = vec[i];
  __syncthreads();
  const long long* pp = a.sp_pptr + M.oK;
  double* part = a.gpart + M.oG;
  const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
  for (int e = e0 + g; e < e1; e += NT / 16) {
    double acc = 0.0;
    const long long q1 = pp[e + 1];
    for (long long q = pp[e] + gl; q < q1; q += 16) acc += a.sp_pval[q] * xs[a.sp_pcon[q]];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 16);
    if (gl == 0) part[e] = acc;
  }
}

// Sparse form of gemvt: <A_i, W> gathered from an LDS copy of W, one
// wavefront per constraint (four per member), reduced across the wave.
__global__ __launch_bounds__(NT) void sp_gemvt_kernel(Arrays a, const int* list, int mode) {
  __shared__ double W[KMAX * KMAX];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k2 = M.k * M.k;
  const int i0 = blockIdx.y * kSpConstraints, i1 = min(M.m, i0 + kSpConstraints);
  if (i0 >= M.m) return;  // whole work-group
  const double* src = (mode == 1 || mode == 3 ? a.W2 : a.Y) + M.oK;
  for (int e = threadIdx.x; e < k2; e += NT) W[e] = src[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long* ptr = a.sp_ptr + M.oM;
  for (int i = i0 + wave; i < i1; i += NT / 64) {
    double s = 0.0;
    for (long long e = ptr[i] + lane; e < ptr[i + 1]; e += 64) s += a.sp_val[e] * W[a.sp_idx[e]];
    s = wave_sum(s);
    if (lane == 0) {
      const long long o = M.oM + i;
      if (mode == 1)
        a.r[o] = -(a.d[o] + s);
      else if (mode == 3)
        a.rho[o] = a.r[o] - s;
      else
        (mode == 0 ? a.d : a.v)[o] = a.c[o] - s;
    }
  }
}

__global__ __launch_bounds__(NT) void gemv_kernel(Arrays a, const int* list, int mode) {
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const long long k2 = (long long)M.k * M.k;
  const int nch = a.sp_ptr ? 1 : (M.m + kColChunk - 1) / kColChunk;  // the sparse form leaves one partial
  const double* part = a.gpart + M.oG;
  double amax = 0.0;
  for (long long e = threadIdx.x; e < k2; e += NT) {
    double acc = 0.0;
    for (int c = 0; c < nch; ++c) acc += part[(long long)c * k2 + e];
    double out;
    if (mode == 1 || mode == 3) {
      out = mode == 1 ? acc + a.P[M.oK + e] : acc;
      a.dX[M.oK + e] = out;
    } else {
      out = acc - a.X[M.oK + e] - a.C[M.oK + e];
      (mode == 0 ? a.P : a.W1)[M.oK + e] = out;
    }
    amax = nan_max(amax, fabs(out));
  }
  if (mode == 0 || mode == 2) {
    amax = block_max(amax, red);
    if (threadIdx.x == 0) a.sc[(long long)p * NSLOT + (mode == 0 ? S_PRES : S_PSOL)] = amax;
  }
}

__global__ __launch_bounds__(NT) void gemvt_kernel(Arrays a, const int* list, int mode) {
  __shared__ double W[KMAX * KMAX];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int i0 = blockIdx.y * kColChunk;
  if (i0 >= M.m) return;
  const int i1 = min(M.m, i0 + kColChunk);
  const int k2 = M.k * M.k;
  const double* src = (mode == 1 || mode == 3 ? a.W2 : a.Y) + M.oK;
  for (int e = threadIdx.x; e < k2; e += NT) W[e] = src[e];
  __syncthreads();
  const double* Av = a.Av + M.oAv;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // two columns per pass for more loads in flight
  for (int i = i0 + 2 * wave; i < i1; i += 2 * (NT / 64)) {
    const bool two = i + 1 < i1;
    const double* c0 = Av + (long long)i * k2;
    const double* c1 = c0 + k2;
    double s0 = 0.0, s1 = 0.0;
    for (int e = lane; e < k2; e += 64) {
      const double we = W[e];
      s0 += c0[e] * we;
      if (two) s1 += c1[e] * we;
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (lane == 0) {
      for (int h = 0; h < (two ? 2 : 1); ++h) {
        const long long o = M.oM + i + h;
        const double s = h == 0 ? s0 : s1;
        if (mode == 1)
          a.r[o] = -(a.d[o] + s);
        else if (mode == 3)
          a.rho[o] = a.r[o] - s;
        else
          (mode == 0 ? a.d : a.v)[o] = a.c[o] - s;
      }
    }
  }
}

// ||d||_inf (mode 0, S_DRES) or ||v||_inf (mode 2, S_DSOL) after gemvt; mode 3: ||rho||_inf
// (S_RHO) and ||r||_inf (S_RNORM)
__global__ __launch_bounds__(NT) void vec_norm_kernel(Arrays a, const int* list, int mode) {
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const double* v = (mode == 0 ? a.d : mode == 3 ? a.rho : a.v) + M.oM;
  double amax = 0.0;
  for (int i = threadIdx.x; i < M.m; i += NT) amax = nan_max(amax, fabs(v[i]));
  amax = block_max(amax, red);
  if (mode == 3) {
    double rmax = 0.0;
    for (int i = threadIdx.x; i < M.m; i += NT) rmax = nan_max(rmax, fabs(a.r[M.oM + i]));
    rmax = block_max(rmax, red);
    if (threadIdx.x == 0) {
      a.sc[(long long)p * NSLOT + S_RHO] = amax;
      a.sc[(long long)p * NSLOT + S_RNORM] = rmax;
    }
    return;
  }
  if (threadIdx.x == 0) a.sc[(long long)p * NSLOT + (mode == 0 ? S_DRES : S_DSOL)] = amax;
}

// ---- k x k block phases (LDS) ----------------------------------------------------------
__global__ __launch_bounds__(NT) void block_factor_kernel(Arrays a, const int* list) {
  __shared__ double sX[KMAX * LP], sY[KMAX * LP], sA[KMAX * LP], sB[KMAX * LP];
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  lds_load(sX, a.X + o, k);
  lds_load(sY, a.Y + o, k);
  __syncthreads();
  // R = -X Y (+ mu I)
  lds_mm(sA, sX, false, sY, false, k, -1.0, 0.0);
  __syncthreads();
  const bool affine = a.par[(long long)p * NPARAM + P_AFF] != 0.0;
  if (!affine && threadIdx.x < k) sA[threadIdx.x * (LP + 1)] += a.par[(long long)p * NPARAM + P_MU];
  __syncthreads();
  lds_store(a.R + o, sA, k);
  __syncthreads();
  // L_X, L_Y
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    sA[r + c * LP] = sX[r + c * LP];
    sB[r + c * LP] = sY[r + c * LP];
  }
  const int ix = lds_chol(sA, k, red);
  const int iy = lds_chol(sB, k, red);
  double mn, mx;
  if (ix == 0) {
    lds_store_lower(a.LX + o, sA, k);
    lds_diag_extrema(sA, k, red, mn, mx);
    if (threadIdx.x == 0) {
      a.sc[(long long)p * NSLOT + S_LXMIN] = mn;
      a.sc[(long long)p * NSLOT + S_LXMAX] = mx;
    }
  }
  if (iy == 0) {
    lds_store_lower(a.LY + o, sB, k);
    lds_diag_extrema(sB, k, red, mn, mx);
    if (threadIdx.x == 0) {
      a.sc[(long long)p * NSLOT + S_LYMIN] = mn;
      a.sc[(long long)p * NSLOT + S_LYMAX] = mx;
    }
  }
  if (threadIdx.x == 0) {
    a.si[(long long)p * NISLOT + I_INFOX] = ix;
    a.si[(long long)p * NISLOT + I_INFOY] = iy;
  }
  if (ix != 0 || iy != 0) return;  // uniform
  // L_X^-1 (into sX) and X^-1 = L_X^-T L_X^-1 (into sY)
  __syncthreads();
  lds_trtri(sA, sX, k, red);
  lds_store(a.LXinv + o, sX, k);
  lds_mm(sY, sX, true, sX, false, k, 1.0, 0.0);
  __syncthreads();
  lds_store(a.Xinv + o, sY, k);
}

// W2 = X^-1 (P Y - R)
__global__ __launch_bounds__(NT) void rhs_block_kernel(Arrays a, const int* list) {
  __shared__ double sP[KMAX * LP], sY[KMAX * LP], sA[KMAX * LP];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  lds_load(sP, a.P + o, k);
  lds_load(sY, a.Y + o, k);
  lds_load(sA, a.R + o, k);
  __syncthreads();
  // sA = P Y - R
  lds_mm(sA, sP, false, sY, false, k, 1.0, -1.0);
  __syncthreads();
  lds_load(sP, a.Xinv + o, k);
  __syncthreads();
  lds_mm(sY, sP, false, sA, false, k, 1.0, 0.0);
  __syncthreads();
  lds_store(a.W2 + o, sY, k);
}

// W2 = X^-1 dX Y, dX = sum_j dx_j A_j (gemv mode 3): the k x k part of the matrix-free FP64 Schur
// operator (S dx)_i = <A_i, X^-1 dX Y> (HKM: S_ij = tr(A_i X^-1 A_j Y))
__global__ __launch_bounds__(NT) void exact_block_kernel(Arrays a, const int* list) {
  __shared__ double sD[KMAX * LP], sY[KMAX * LP], sA[KMAX * LP];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  lds_load(sD, a.dX + o, k);
  lds_load(sY, a.Y + o, k);
  __syncthreads();
  lds_mm(sA, sD, false, sY, false, k, 1.0, 0.0);  // dX Y
  __syncthreads();
  lds_load(sD, a.Xinv + o, k);
  __syncthreads();
  lds_mm(sY, sD, false, sA, false, k, 1.0, 0.0);
  __syncthreads();
  lds_store(a.W2 + o, sY, k);
}

// dY = sym(X^-1 (R - dX Y))
__global__ __launch_bounds__(NT) void dir_block_kernel(Arrays a, const int* list) {
  __shared__ double sD[KMAX * LP], sY[KMAX * LP], sA[KMAX * LP];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  lds_load(sD, a.dX + o, k);
  lds_load(sY, a.Y + o, k);
  lds_load(sA, a.R + o, k);
  __syncthreads();
  lds_mm(sA, sD, false, sY, false, k, -1.0, 1.0);  // R - dX Y
  __syncthreads();
  lds_load(sD, a.Xinv + o, k);
  __syncthreads();
  lds_mm(sY, sD, false, sA, false, k, 1.0, 0.0);
  __syncthreads();
  double* dY = a.dY + o;
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    dY[e] = (sY[r + c * LP] + sY[c + r * LP]) / 2.0;
  }
}

__global__ __launch_bounds__(NT) void comp_kernel(Arrays a, const int* list) {
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k2 = M.k * M.k;
  const long long o = M.oK;
  double ta = a.par[(long long)p * NPARAM + P_A], tb = a.par[(long long)p * NPARAM + P_B];
  // a negative step means: the affine step the fraction-to-boundary pass just left in S_TX / S_TY
  if (ta < 0.0) ta = a.sc[(long long)p * NSLOT + S_TX];
  if (tb < 0.0) tb = a.sc[(long long)p * NSLOT + S_TY];
  double s0 = 0.0, s1 = 0.0;
  for (int e = threadIdx.x; e < k2; e += NT) {
    const double x = a.X[o + e], y = a.Y[o + e];
    s0 += x * y;
    s1 += (x + ta * a.dX[o + e]) * (y + tb * a.dY[o + e]);
  }
  s0 = block_sum(s0, red);
  s1 = block_sum(s1, red);
  if (threadIdx.x == 0) {
    a.sc[(long long)p * NSLOT + S_COMP] = s0;
    a.sc[(long long)p * NSLOT + S_COMPT] = s1;
  }
}

// R = -X Y - dX dY + target I (threaded.jl:839-854, 943-1003)
__global__ __launch_bounds__(NT) void corrector_block_kernel(Arrays a, const int* list) {
  __shared__ double sX[KMAX * LP], sY[KMAX * LP], sA[KMAX * LP];
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k, k2 = k * k;
  const long long o = M.oK;
  const double* par = a.par + (long long)p * NPARAM;
  const int mode = (int)par[P_MODE];
  double dotv = 0.0;
  for (int e = threadIdx.x; e < k2; e += NT) {
    const double x = a.X[o + e], y = a.Y[o + e];
    dotv += mode == 0 ? (x + a.dX[o + e]) * (y + a.dY[o + e]) : x * y;
  }
  dotv = block_sum(dotv, red);
  double target;
  if (mode == 0) {
    const double rl = dotv / par[P_MU] / (double)k;
    target = fmax(rl < 1.0 ? rl * rl : rl, par[P_BETA]) * par[P_MU];
  } else if (mode == 1) {
    target = par[P_SIGMA] * dotv / (double)k;
  } else {
    target = par[P_SIGMA] * par[P_MUG];
  }
  lds_load(sX, a.X + o, k);
  lds_load(sY, a.Y + o, k);
  __syncthreads();
  lds_mm(sA, sX, false, sY, false, k, -1.0, 0.0);
  __syncthreads();
  lds_load(sX, a.dX + o, k);
  lds_load(sY, a.dY + o, k);
  __syncthreads();
  lds_mm(sA, sX, false, sY, false, k, -1.0, 1.0);
  __syncthreads();
  if (threadIdx.x < k) sA[threadIdx.x * (LP + 1)] += target;
  __syncthreads();
  lds_store(a.R + o, sA, k);
}

// ---- step length / PSD checks (kernels/generic.jl:292-420, step.jl:1025-1098) ---------
// trial = base + t * dir (lower triangle) factored in sT; returns info
// (base and dir are read from global memory, leading dimension k: they stay in L2)
__device__ int trial_chol(const double* base, const double* dir, double t, double* sT, int k, double* red) {
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    if (r >= c) sT[r + c * LP] = base[e] + t * dir[e];
  }
  return lds_chol(sT, k, red);
}

// trial_has_cholesky_margin! (generic.jl:318-340); ratio 0: plain positive definiteness
__device__ bool trial_ok(const double* base, const double* dir, double t, double ratio, double* sT, int k,
                         double* red) {
  if (trial_chol(base, dir, t, sT, k, red) != 0) return false;
  if (!(ratio > 0.0)) return true;
  double mn, mx;
  lds_diag_extrema(sT, k, red, mn, mx);
  return isfinite(mn) && isfinite(mx) && mx > 0.0 && mn > ratio * mx;
}

// fraction_to_boundary_bound! for k >= 3: full step, else 64-step bisection
__device__ double ftb_side(const double* base, const double* dir, double* sT, int k, double* red) {
  if (trial_chol(base, dir, 1.0, sT, k, red) == 0) return 1.0;
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 64; ++it) {
    const double mid = (lo + hi) / 2.0;
    if (trial_chol(base, dir, mid, sT, k, red) == 0)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

#pragma clang fp contract(off)
// closed forms for k <= 2 (generic.jl:356-401), as in the dense route
__device__ bool small_isposdef(const double* X, const double* dX, int k, double t) {
  if (k == 1) return X[0] + t * dX[0] > 0.0;
  const double a0 = X[0] + t * dX[0];
  if (!(a0 > 0.0)) return false;
  const double b0 = X[1] + t * dX[1];
  const double d0 = X[LP + 1] + t * dX[LP + 1];
  return a0 * d0 > b0 * b0;
}

__device__ double small_ftb(const double* X, const double* dX, int k) {
  if (k == 1) {
    const double d = dX[0];
    return d < 0.0 ? fmin(1.0, -X[0] / d) : 1.0;
  }
  const double a0 = X[0], b0 = X[LP], d0 = X[LP + 1];
  const double da = dX[0], db = (dX[LP] + dX[1]) / 2.0, dd = dX[LP + 1];
  const double a1 = a0 + da, b1 = b0 + db, d1 = d0 + dd;
  if (a1 > 0.0 && a1 * d1 > b1 * b1) return 1.0;
  const double c0 = a0 * d0 - b0 * b0;
  const double c1 = da * d0 + a0 * dd - 2.0 * b0 * db;
  const double c2 = da * dd - db * db;
  double root = 1.0;
  bool found = false;
  if (c2 == 0.0) {
    if (c1 < 0.0) {
      const double cand = -c0 / c1;
      if (0.0 < cand && cand <= 1.0) {
        root = cand;
        found = true;
      }
    }
  } else {
    const double disc = c1 * c1 - 4.0 * c2 * c0;
    if (disc >= 0.0) {
      const double sd = sqrt(disc);
      const double den = 2.0 * c2;
      const double r1 = (-c1 - sd) / den;
      const double r2 = (-c1 + sd) / den;
      if (0.0 < r1 && r1 <= root) {
        root = r1;
        found = true;
      }
      if (0.0 < r2 && r2 <= root) {
        root = r2;
        found = true;
      }
    }
  }
  if (found) return root;
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 64; ++it) {
    const double mid = (lo + hi) / 2.0;
    if (small_isposdef(X, dX, k, mid))
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ bool small_margin(const double* X, const double* dX, int k, double t, double ratio) {
  if (!(ratio > 0.0)) return small_isposdef(X, dX, k, t);
  double mn, mx;
  const double d11 = X[0] + t * dX[0];
  if (!(d11 > 0.0)) return false;
  if (k == 1) {
    mn = mx = fabs(sqrt(d11));
  } else {
    const double l11 = sqrt(d11);
    const double l21 = (X[1] + t * dX[1]) / l11;
    const double d22 = (X[LP + 1] + t * dX[LP + 1]) - l21 * l21;
    if (!(d22 > 0.0)) return false;
    const double l22 = sqrt(d22);
    mn = fmin(fabs(l11), fabs(l22));
    mx = fmax(fabs(l11), fabs(l22));
  }
  return isfinite(mn) && isfinite(mx) && mx > 0.0 && mn > ratio * mx;
}
#pragma clang fp contract(on)

__global__ __launch_bounds__(NT) void ftb_kernel(Arrays a, const int* list) {
  __shared__ double sB[2 * LP], sD[2 * LP], sT[KMAX * LP];  // sB, sD: blocks of side <= 2
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  const double* par = a.par + (long long)p * NPARAM;
  double b[2];
  for (int side = 0; side < 2; ++side) {
    const double* base = (side == 0 ? a.X : a.Y) + o;
    const double* dir = (side == 0 ? a.dX : a.dY) + o;
    if (k <= 2) {
      __syncthreads();
      lds_load(sB, base, k);
      lds_load(sD, dir, k);
      __syncthreads();
      b[side] = small_ftb(sB, sD, k);
    } else {
      b[side] = ftb_side(base, dir, sT, k, red);
    }
  }
  if (threadIdx.x == 0) {
    a.sc[(long long)p * NSLOT + S_TX] = b[0] < 1.0 ? par[P_PF] * b[0] : 1.0;
    a.sc[(long long)p * NSLOT + S_TY] = b[1] < 1.0 ? par[P_DF] * b[1] : 1.0;
  }
}

// line_search! (step.jl:1025-1075): t = 1, gamma, gamma^2, ... (repeated
// multiplication); a side stops at its first passing trial or at the first
// candidate below min_step
__global__ __launch_bounds__(NT) void backtrack_kernel(Arrays a, const int* list) {
  __shared__ double sB[2 * LP], sD[2 * LP], sT[KMAX * LP];  // sB, sD: blocks of side <= 2
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  const double* par = a.par + (long long)p * NPARAM;
  const double gamma = par[P_GAMMA], min_step = par[P_MINSTEP], ratio = par[P_RATIO];
  double res[2];
  for (int side = 0; side < 2; ++side) {
    const double* base = (side == 0 ? a.X : a.Y) + o;
    const double* dir = (side == 0 ? a.dX : a.dY) + o;
    if (k <= 2) {
      __syncthreads();
      lds_load(sB, base, k);
      lds_load(sD, dir, k);
      __syncthreads();
    }
    double t = 1.0;
    while (true) {
      const bool pass = k <= 2 ? small_margin(sB, sD, k, t, ratio) : trial_ok(base, dir, t, ratio, sT, k, red);
      if (pass || t < min_step || !(t > 0.0)) break;  // t reaches 0 after ~1100 steps: the loop always ends
      t *= gamma;
    }
    res[side] = t;
  }
  if (threadIdx.x == 0) {
    a.sc[(long long)p * NSLOT + S_TX] = res[0];
    a.sc[(long long)p * NSLOT + S_TY] = res[1];
  }
}

// ---- wavefront forms of the step-length searches -------------------------
// One wavefront per member (four members per work-group, no block barriers):
// lane r holds row r of the trial matrix base + t dir in registers (KB >= k
// columns, padded to the identity), eliminates on the unscaled columns as
// lds_chol does (same operations and order, so every pass / fail decision is
// the same) with the pivot column published through a wave-private LDS row.
// |diag L| extrema (sqrt of the pivots) serve the margin test.
template <int KB>
__device__ __forceinline__ int wave_trial_chol(const double* base, const double* dir, double t, int k, double* col, int lane,
                               double& dmin, double& dmax) {
  // opaque per trial: the row addresses are recomputed here instead of being
  // hoisted out of the bisection loop, where they would hold 2 KB registers
  const double* bp = base;
  const double* dp = dir;
  int row = lane;
  asm volatile("" : "+v"(bp), "+v"(dp), "+v"(row));
  double v[KB];
#pragma unroll
  for (int c = 0; c < KB; ++c)
    v[c] = (row < k && c <= row) ? bp[row + c * k] + t * dp[row + c * k] : (c == row ? 1.0 : 0.0);
  dmin = __builtin_inf();
  dmax = 0.0;
  int info = 0;
  // no early exit inside the unrolled loop (it keeps v[] in registers): steps
  // past k or past a failed pivot are skipped by a wave-uniform test
#pragma unroll
  for (int c = 0; c < KB; ++c) {
    if (c < k && info == 0) {
      col[lane] = v[c];
      __builtin_amdgcn_wave_barrier();
      const double p = col[c];
      if (!(p > 0.0)) {
        info = c + 1;
      } else {
        const double d = fabs(sqrt(p));
        dmin = fmin(dmin, d);
        dmax = nan_max(dmax, d);
        const double f = v[c] / p;
#pragma unroll
        for (int c2 = c + 1; c2 < KB; ++c2) v[c2] -= f * col[c2];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  return info;
}

template <int KB>
__device__ __forceinline__ bool wave_trial_ok(const double* base, const double* dir, double t, double ratio, int k, double* col,
                              int lane) {
  double mn, mx;
  if (wave_trial_chol<KB>(base, dir, t, k, col, lane, mn, mx) != 0) return false;
  if (!(ratio > 0.0)) return true;
  return isfinite(mn) && isfinite(mx) && mx > 0.0 && mn > ratio * mx;
}

// ---- half-wavefront trials (k <= 32) -----------------------------------------------------------
// A half wavefront (32 lanes, one row each) factors one trial exactly as
// wave_trial_chol does (same operations and order, so every decision is the
// same), publishing its pivot column through its own LDS row; the two halves
// of a wavefront use rows 33 doubles apart, so their reads fall in different
// banks.
constexpr int kPairPitch = 33;

template <int KB>
__device__ __forceinline__ bool pair_trial_ok(const double* base, const double* dir, double t, double ratio, int k,
                                              double* col, int row) {
  const double* bp = base;
  const double* dp = dir;
  int r = row;
  asm volatile("" : "+v"(bp), "+v"(dp), "+v"(r));
  double v[KB];
#pragma unroll
  for (int c = 0; c < KB; ++c)
    v[c] = (r < k && c <= r) ? bp[r + c * k] + t * dp[r + c * k] : (c == r ? 1.0 : 0.0);
  double dmin = __builtin_inf(), dmax = 0.0;
  int info = 0;
#pragma unroll
  for (int c = 0; c < KB; ++c) {
    if (c < k && info == 0) {  // uniform within a half
      col[r] = v[c];
      __builtin_amdgcn_wave_barrier();
      const double p = col[c];
      if (!(p > 0.0)) {
        info = c + 1;
      } else {
        if (ratio > 0.0) {
          const double d = fabs(sqrt(p));
          dmin = fmin(dmin, d);
          dmax = nan_max(dmax, d);
        }
        const double f = v[c] / p;
#pragma unroll
        for (int c2 = c + 1; c2 < KB; ++c2) v[c2] -= f * col[c2];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (info != 0) return false;
  if (!(ratio > 0.0)) return true;
  return isfinite(dmin) && isfinite(dmax) && dmax > 0.0 && dmin > ratio * dmax;
}

// 32 < k <= 64: one work-group per (member, side), four wavefront trials per
// round (node = wave): two bisection levels (the first round: the full step
// and the first two levels) or four backtracking candidates, replayed as in
// step_quad_kernel below.
template <int KB>
__global__ __launch_bounds__(NT) void step_quad_wide_kernel(Arrays a, const int* list, int n, int op) {
  __shared__ double cols[NT / 64][KMAX];
  __shared__ int pass[4];
  static_assert(NT == 256, "four wavefronts: one trial each");
  const int p = list[blockIdx.x >> 1];
  const int side = blockIdx.x & 1;
  const Member M = a.mem[p];
  const int k = M.k;
  if (k > KB || k <= KB - 16) return;  // whole work-group
  const int node = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long o = M.oK;
  const double* par = a.par + (long long)p * NPARAM;
  const double* base = (side == 0 ? a.X : a.Y) + o;
  const double* dir = (side == 0 ? a.dX : a.dY) + o;
  double* col = cols[node];
  double res = 1.0;
  bool active = true;
  if (op == 0) {
    double lo = 0.0, hi = 1.0;
    int levels = 0;
    bool first = true;
    while (active) {
      double t = 0.0;
      if (first) {
        t = node == 0 ? 1.0 : node == 1 ? 0.5 : node == 2 ? 0.25 : 0.75;
      } else if (node < 3) {
        const double mid = (lo + hi) / 2.0;
        t = node == 0 ? mid : node == 1 ? (lo + mid) / 2.0 : (mid + hi) / 2.0;
      }
      const bool ok = wave_trial_ok<KB>(base, dir, t, 0.0, k, col, lane);
      __syncthreads();
      if (lane == 0) pass[node] = ok ? 1 : 0;
      __syncthreads();
      int nb = 0;
      if (first && pass[0]) {
        res = 1.0;
        active = false;
        break;
      }
      if (first) nb = 1;
      for (int lvl = 0; lvl < 2; ++lvl) {
        const double mid = (lo + hi) / 2.0;
        if (levels >= 64 || mid == lo || mid == hi) {
          active = false;
          break;
        }
        ++levels;
        const bool up = pass[nb] != 0;
        if (up)
          lo = mid;
        else
          hi = mid;
        nb = first ? (up ? 3 : 2) : (up ? 2 : 1);
      }
      if (!active || levels >= 64) {
        active = false;
        res = lo;
      }
      first = false;
    }
  } else {
    const double gamma = par[P_GAMMA], min_step = par[P_MINSTEP], ratio = par[P_RATIO];
    double tn = 1.0;
    while (active) {
      double t = tn;
      for (int j = 0; j < node; ++j) t *= gamma;
      const bool ok = wave_trial_ok<KB>(base, dir, t, ratio, k, col, lane);
      __syncthreads();
      if (lane == 0) pass[node] = ok ? 1 : 0;
      __syncthreads();
      double tc = tn;
      for (int j = 0; j < 4; ++j) {
        if (pass[j] || tc < min_step || !(tc > 0.0)) {
          res = tc;
          active = false;
          break;
        }
        tc *= gamma;
      }
      tn = tc;  // t reaches 0 after ~1100 candidates: the loop always ends
    }
  }
  if (threadIdx.x == 0) {
    const double frac = side == 0 ? par[P_PF] : par[P_DF];
    a.sc[(long long)p * NSLOT + (side == 0 ? S_TX : S_TY)] = op == 0 ? (res < 1.0 ? frac * res : 1.0) : res;
  }
}

// 2 < k <= 32: one work-group per member and eight half-wavefront trial slots,
// slots 0-3 for the X side and 4-7 for the Y side.  Each round a side tests
// the midpoints of the next two bisection levels (three nodes; the first
// round tests the full step and the first two levels: four nodes), or the
// next four backtracking candidates, and every thread replays the sequential
// search from the outcomes (the same midpoints by the same arithmetic, the
// same adjacency stop and 64-level cap), so the result is the sequential one.
template <int KB>
__global__ __launch_bounds__(NT) void step_quad_kernel(Arrays a, const int* list, int n, int op) {
  __shared__ double cols[NT / 64][2 * kPairPitch];
  __shared__ int pass[8];
  static_assert(NT == 256, "four wavefronts: eight half-wavefront slots");
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  if (k > KB || k <= KB - 16 || k <= 2) return;  // whole work-group
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, row = lane & 31;
  const int slot = 2 * w + half, side = slot >> 2, node = slot & 3;
  const long long o = M.oK;
  const double* par = a.par + (long long)p * NPARAM;
  const double* base = (side == 0 ? a.X : a.Y) + o;
  const double* dir = (side == 0 ? a.dX : a.dY) + o;
  double* hcol = &cols[w][half * kPairPitch];
  double res[2] = {1.0, 1.0};
  bool active[2] = {true, true};
  if (op == 0) {
    double lo[2] = {0.0, 0.0}, hi[2] = {1.0, 1.0};
    int levels[2] = {0, 0};
    bool first = true;
    while (active[0] || active[1]) {
      // this slot's candidate
      double t = 0.0;
      if (active[side]) {
        if (first) {
          t = node == 0 ? 1.0 : node == 1 ? 0.5 : node == 2 ? 0.25 : 0.75;
        } else if (node < 3) {
          const double mid = (lo[side] + hi[side]) / 2.0;
          t = node == 0 ? mid : node == 1 ? (lo[side] + mid) / 2.0 : (mid + hi[side]) / 2.0;
        }
      }
      const bool ok = pair_trial_ok<KB>(base, dir, t, 0.0, k, hcol, row);
      __syncthreads();  // the previous round's outcomes have been read
      if (row == 0) pass[slot] = ok ? 1 : 0;
      __syncthreads();
      for (int sd = 0; sd < 2; ++sd) {
        if (!active[sd]) continue;
        const int* ps = pass + 4 * sd;
        int nb = 0;  // node of the next level's midpoint
        if (first) {
          if (ps[0]) {
            res[sd] = 1.0;
            active[sd] = false;
            continue;
          }
          nb = 1;
        }
        for (int lvl = 0; lvl < 2 && active[sd]; ++lvl) {
          const double mid = (lo[sd] + hi[sd]) / 2.0;
          if (levels[sd] >= 64 || mid == lo[sd] || mid == hi[sd]) {
            active[sd] = false;
            break;
          }
          ++levels[sd];
          const bool up = ps[nb] != 0;
          if (up)
            lo[sd] = mid;
          else
            hi[sd] = mid;
          // next level's node: (lo + mid) / 2 after a failure, (mid + hi) / 2 after a pass
          nb = first ? (up ? 3 : 2) : (up ? 2 : 1);
        }
        if (!active[sd] || levels[sd] >= 64) {
          active[sd] = false;
          res[sd] = lo[sd];
        }
      }
      first = false;
    }
  } else {
    const double gamma = par[P_GAMMA], min_step = par[P_MINSTEP], ratio = par[P_RATIO];
    double tn[2] = {1.0, 1.0};  // next candidate of each side
    while (active[0] || active[1]) {
      double t = 0.0;
      if (active[side]) {
        t = tn[side];
        for (int j = 0; j < node; ++j) t *= gamma;  // t_j = t_0 gamma^j by repeated multiplication
      }
      const bool ok = pair_trial_ok<KB>(base, dir, t, ratio, k, hcol, row);
      __syncthreads();
      if (row == 0) pass[slot] = ok ? 1 : 0;
      __syncthreads();
      for (int sd = 0; sd < 2; ++sd) {
        if (!active[sd]) continue;
        double tc = tn[sd];
        for (int j = 0; j < 4; ++j) {
          if (pass[4 * sd + j] || tc < min_step || !(tc > 0.0)) {
            res[sd] = tc;
            active[sd] = false;
            break;
          }
          tc *= gamma;
        }
        tn[sd] = tc;  // t reaches 0 after ~1100 candidates: the loop always ends
      }
    }
  }
  if (threadIdx.x == 0) {
    if (op == 0) {
      a.sc[(long long)p * NSLOT + S_TX] = res[0] < 1.0 ? par[P_PF] * res[0] : 1.0;
      a.sc[(long long)p * NSLOT + S_TY] = res[1] < 1.0 ? par[P_DF] * res[1] : 1.0;
    } else {
      a.sc[(long long)p * NSLOT + S_TX] = res[0];
      a.sc[(long long)p * NSLOT + S_TY] = res[1];
    }
  }
}

// op 0: fraction to the boundary (S_TX, S_TY scaled by P_PF / P_DF); op 1: backtracking.
// k <= 2: closed forms on the 2 x 2 (or 1 x 1) block, one thread per member.
__global__ __launch_bounds__(NT) void step_small_kernel(Arrays a, const int* list, int n, int op) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const int p = list[i];
  const Member M = a.mem[p];
  const int k = M.k;
  if (k > 2) return;
  const long long o = M.oK;
  const double* par = a.par + (long long)p * NPARAM;
  double res[2];
  for (int side = 0; side < 2; ++side) {
    const double* base = (side == 0 ? a.X : a.Y) + o;
    const double* dir = (side == 0 ? a.dX : a.dY) + o;
    // LP-strided as in the LDS kernels
    double sb[LP + 2] = {}, sd[LP + 2] = {};
    sb[0] = base[0];
    sd[0] = dir[0];
    if (k == 2) {
      sb[1] = base[1];
      sb[LP] = base[2];
      sb[LP + 1] = base[3];
      sd[1] = dir[1];
      sd[LP] = dir[2];
      sd[LP + 1] = dir[3];
    }
    if (op == 0) {
      res[side] = small_ftb(sb, sd, k);
    } else {
      double t = 1.0;
      while (true) {
        const bool pass = small_margin(sb, sd, k, t, par[P_RATIO]);
        if (pass || t < par[P_MINSTEP] || !(t > 0.0)) break;  // t reaches 0 after ~1100 steps
        t *= par[P_GAMMA];
      }
      res[side] = t;
    }
  }
  if (op == 0) {
    a.sc[(long long)p * NSLOT + S_TX] = res[0] < 1.0 ? par[P_PF] * res[0] : 1.0;
    a.sc[(long long)p * NSLOT + S_TY] = res[1] < 1.0 ? par[P_DF] * res[1] : 1.0;
  } else {
    a.sc[(long long)p * NSLOT + S_TX] = res[0];
    a.sc[(long long)p * NSLOT + S_TY] = res[1];
  }
}

// ---- accepted step ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void update_kernel(Arrays a, const int* list) {
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k2 = M.k * M.k;
  const long long o = M.oK;
  const double tX = a.par[(long long)p * NPARAM + P_A], tY = a.par[(long long)p * NPARAM + P_B];
  double comp = 0.0, xm = 0.0, ym = 0.0, vm = 0.0;
  for (int e = threadIdx.x; e < k2; e += NT) {
    const double x = a.X[o + e] + tX * a.dX[o + e];
    const double y = a.Y[o + e] + tY * a.dY[o + e];
    a.X[o + e] = x;
    a.Y[o + e] = y;
    comp += x * y;
    xm = nan_max(xm, fabs(x));
    ym = nan_max(ym, fabs(y));
  }
  for (int i = threadIdx.x; i < M.m; i += NT) {
    const double v = a.x[M.oM + i] + tX * a.dx[M.oM + i];
    a.x[M.oM + i] = v;
    vm = nan_max(vm, fabs(v));
  }
  comp = block_sum(comp, red);
  xm = block_max(xm, red);
  ym = block_max(ym, red);
  vm = block_max(vm, red);
  if (threadIdx.x == 0) {
    double* sc = a.sc + (long long)p * NSLOT;
    sc[S_COMP] = comp;
    sc[S_XBMAX] = xm;
    sc[S_YBMAX] = ym;
    sc[S_XMAX] = vm;
  }
}

__global__ __launch_bounds__(NT) void accepted_kernel(Arrays a, const int* list) {
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const double tY = a.par[(long long)p * NPARAM + P_B];
  double mx = 0.0;
  for (int i = threadIdx.x; i < M.m; i += NT) {
    const long long o = M.oM + i;
    const double v = (1.0 - tY) * a.d[o] - tY * a.rho[o];
    a.d[o] = v;
    mx = nan_max(mx, fabs(v));
  }
  mx = block_max(mx, red);
  if (threadIdx.x == 0) a.sc[(long long)p * NSLOT + S_DACC] = mx;
}

__global__ __launch_bounds__(NT) void objectives_kernel(Arrays a, const int* list) {
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  double po = 0.0, du = 0.0;
  for (int i = threadIdx.x; i < M.m; i += NT) po += a.c[M.oM + i] * a.x[M.oM + i];
  for (int e = threadIdx.x; e < M.k * M.k; e += NT) du += a.C[M.oK + e] * a.Y[M.oK + e];
  po = block_sum(po, red);
  du = block_sum(du, red);
  if (threadIdx.x == 0) {
    a.sc[(long long)p * NSLOT + S_POBJ] = po;
    a.sc[(long long)p * NSLOT + S_DOBJ] = du;
  }
}

__global__ __launch_bounds__(NT) void scale_kernel(Arrays a, const int* list, int which) {
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const double f = a.par[(long long)p * NPARAM + P_SCALE];
  double* B = (which == 1 ? a.X : a.Y) + M.oK;
  for (int e = threadIdx.x; e < M.k * M.k; e += NT) B[e] = f * B[e];
}

__global__ __launch_bounds__(NT) void best_kernel(Arrays a, const int* list, int dir) {
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const bool ck = dir >= 2;
  double* sX = ck ? a.ckX : a.bX;
  double* sY = ck ? a.ckY : a.bY;
  double* sx = ck ? a.ckx : a.bx;
  const bool save = (dir & 1) == 0;
  for (int e = threadIdx.x; e < M.k * M.k; e += NT) {
    const long long o = M.oK + e;
    if (save) {
      sX[o] = a.X[o];
      sY[o] = a.Y[o];
    } else {
      a.X[o] = sX[o];
      a.Y[o] = sY[o];
    }
  }
  for (int i = threadIdx.x; i < M.m; i += NT) {
    const long long o = M.oM + i;
    if (save)
      sx[o] = a.x[o];
    else
      a.x[o] = sx[o];
  }
}

// ---- adjoint of L = <G, Y> + h'x (the core form of tr(G X*) + h'y*) ------------------
// W = sym(X^-1 G Y) into W1 and W2; d = -h (so the gemvt pass forms r = h - Av' vec(W))
__global__ __launch_bounds__(NT) void adj_w_kernel(Arrays a, const int* list) {
  __shared__ double sA[KMAX * LP], sB[KMAX * LP], sC[KMAX * LP];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  lds_load(sA, a.G + o, k);
  lds_load(sB, a.Y + o, k);
  __syncthreads();
  lds_mm(sC, sA, false, sB, false, k, 1.0, 0.0);  // G Y
  __syncthreads();
  lds_load(sA, a.Xinv + o, k);
  __syncthreads();
  lds_mm(sB, sA, false, sC, false, k, 1.0, 0.0);  // X^-1 G Y
  __syncthreads();
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    const double w = (sB[r + c * LP] + sB[c + r * LP]) / 2.0;
    a.W1[o + e] = w;
    a.W2[o + e] = w;
  }
  for (int i = threadIdx.x; i < M.m; i += NT) a.d[M.oM + i] = -a.v[M.oM + i];
}

// dL/dC_std = -(sym(X^-1 A_lambda Y) + W) into W2 (A_lambda = sum lambda_i A_i is in dX)
__global__ __launch_bounds__(NT) void adj_final_kernel(Arrays a, const int* list) {
  __shared__ double sA[KMAX * LP], sB[KMAX * LP], sC[KMAX * LP];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int k = M.k;
  const long long o = M.oK;
  lds_load(sA, a.dX + o, k);
  lds_load(sB, a.Y + o, k);
  __syncthreads();
  lds_mm(sC, sA, false, sB, false, k, 1.0, 0.0);  // A_lambda Y
  __syncthreads();
  lds_load(sA, a.Xinv + o, k);
  __syncthreads();
  lds_mm(sB, sA, false, sC, false, k, 1.0, 0.0);  // X^-1 A_lambda Y
  __syncthreads();
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int r = e % k, c = e / k;
    a.W2[o + e] = -((sB[r + c * LP] + sB[c + r * LP]) / 2.0 + a.W1[o + e]);
  }
}

__global__ void set_int_kernel(Arrays a, const int* list, int n, int slot, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a.si[(long long)list[i] * NISLOT + slot] = value;
}

// ---- Schur panel Phat_i = L_X^-1 A_i L_Y on FP64 MFMA ------------------------------------
// One work-group per (member, run of `chunk` constraints), one instance per
// padded block size KP = 16 ceil(k/16) (compile-time trip counts and LDS
// sized to the block).  L_X^-1 and L_Y are staged once; each A_i is
// prefetched into registers one constraint ahead and staged in LDS,
// T = L_X^-1 A_i goes to LDS and Phat_i = T L_Y to HBM.  Products run on
// 16 x 16 MFMA tiles; small blocks put several constraints in flight per
// work-group (CP slots of WPC waves: KP = 16 eight slots of one wave, KP = 32
// two slots of four waves, one tile each), larger ones spread the tiles of
// one constraint over the eight waves (tile (ti, tj) to group (ti + tj) % 4,
// ranks 0 and 3 of a group to wave g, 1 and 2 to wave g + 4, which balances
// the triangular K ranges).  The operands are fed swapped so the accumulator
// holds the tile transposed (lane % 16 indexes rows: contiguous column-major
// stores).
constexpr int PT = 512;  // 8 waves: two per SIMD hide MFMA / LDS latency

template <int KP>
struct PanelShape {
  static constexpr int NTL = KP / 16;                         // tiles per side
  static constexpr int CP = KP == 16 ? 8 : KP == 32 ? 2 : 1;  // constraints in flight
  static constexpr int WPC = 8 / CP;                          // waves per constraint
  // pitches: operands read as (16 consecutive columns x 4 rows) fragments use
  // LPA with gcd(LPA, 32) = 2 (a half-wave's 8-byte reads land on 32 distinct
  // bank pairs); (16 consecutive rows x 4 columns) fragments use LPB = 16 mod 32
  // (the two row groups of a half-wave fall in opposite bank halves)
  static constexpr int LPA = KP + 2;
  static constexpr int LPB = KP == 16 ? 16 : KP == 32 ? 48 : KP == 48 ? 48 : 80;
  static constexpr int NPF = (KP * KP + WPC * 64 - 1) / (WPC * 64);  // doubles of A_i per thread
};

// Tiles of wave w of a constraint slot, fixed at compile time so that the
// product loops unroll with constant K ranges (no branches between MFMAs).
struct PanelTiles {
  int n;
  int r0[2], c0[2];
};
template <int KP>
constexpr PanelTiles panel_tiles(int w) {
  PanelTiles ts{0, {0, 0}, {0, 0}};
  constexpr int NTL = KP / 16;
  if (PanelShape<KP>::WPC < 8) {
    if (w < NTL * NTL) {
      ts.n = 1;
      ts.r0[0] = (w % NTL) * 16;
      ts.c0[0] = (w / NTL) * 16;
    }
    return ts;
  }
  for (int t = 0; t < NTL * NTL; ++t) {
    const int ti = t % NTL, tj = t / NTL;
    const int g = (ti + tj) % 4;
    int rank = 0, size = 0;
    for (int t2 = 0; t2 < NTL * NTL; ++t2) {
      const int ti2 = t2 % NTL, tj2 = t2 / NTL;
      if ((ti2 + tj2) % 4 != g) continue;
      ++size;
      if (ti2 < ti) ++rank;
    }
    const int half = size == 4 ? (rank == 1 || rank == 2) : (rank & 1);
    if (g + 4 * half == w && ts.n < 2) {
      ts.r0[ts.n] = ti * 16;
      ts.c0[ts.n] = tj * 16;
      ++ts.n;
    }
  }
  return ts;
}

// Both products of one constraint for wave W of its slot (the barrier between
// them is taken by every wave of the work-group: one instance per wave).
template <int KP, int W>
__device__ __forceinline__ void panel_products(const double* mA, double* mT, const double* sLi, const double* sLy,
                                               double* Pi, float* Pf, int k, int lr, int lq) {
  using S = PanelShape<KP>;
  constexpr int LPA = S::LPA, LPB = S::LPB;
  constexpr PanelTiles ts = panel_tiles<KP>(W);
  // T = L_X^-1 A_i: tile rows r0, K range q < r0 + 16 (L_X^-1 lower)
  {
    double4_t acc[2] = {double4_t{0.0, 0.0, 0.0, 0.0}, double4_t{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int q0 = 0; q0 < KP; q0 += 4) {
#pragma unroll
      for (int u = 0; u < ts.n; ++u) {
        if (q0 >= ts.r0[u] + 16) continue;
        const double fb = mA[(q0 + lq) + (ts.c0[u] + lr) * LPA];   // B'[i'][kq] = A[q0+kq][c0+i']
        const double fl = sLi[(ts.r0[u] + lr) + (q0 + lq) * LPB];  // L'[kq][j'] = Li[r0+j'][q0+kq]
        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb, fl, acc[u], 0, 0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < ts.n; ++u)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) mT[(ts.r0[u] + lr) + (ts.c0[u] + lq + 4 * rr) * LPB] = acc[u][rr];
  }
  __syncthreads();
  // Phat_i = T L_Y: tile columns c0, K range q >= c0 (L_Y lower)
  {
    double4_t acc[2] = {double4_t{0.0, 0.0, 0.0, 0.0}, double4_t{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int q0 = 0; q0 < KP; q0 += 4) {
#pragma unroll
      for (int u = 0; u < ts.n; ++u) {
        if (q0 < ts.c0[u]) continue;
        const double fb = sLy[(q0 + lq) + (ts.c0[u] + lr) * LPA];  // Ly[q0+kq][c0+i']
        const double fl = mT[(ts.r0[u] + lr) + (q0 + lq) * LPB];   // T[r0+j'][q0+kq]
        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb, fl, acc[u], 0, 0, 0);
      }
    }
    if (Pi) {
#pragma unroll
      for (int u = 0; u < ts.n; ++u) {
        const int row = ts.r0[u] + lr;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int col = ts.c0[u] + lq + 4 * rr;
          if (row < k && col < k) Pi[row + col * k] = acc[u][rr];
        }
      }
    } else if (Pf) {
#pragma unroll
      for (int u = 0; u < ts.n; ++u) {
        const int row = ts.r0[u] + lr;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int col = ts.c0[u] + lq + 4 * rr;
          if (row < k && col < k) Pf[row + col * k] = (float)acc[u][rr];
        }
      }
    }
  }
}

template <int KP>
__global__ __launch_bounds__(PT) void panel_kernel(Arrays a, const int2* work, int chunk) {
  using S = PanelShape<KP>;
  constexpr int LPA = S::LPA, LPB = S::LPB, CP = S::CP, WPC = S::WPC, NPF = S::NPF;
  __shared__ double sLi[KP * LPB], sLy[KP * LPA], sA[CP][KP * LPA], sT[CP][KP * LPB];
  const int2 w = work[blockIdx.x];
  const int p = w.x;
  const Member M = a.mem[p];
  const bool skip64 = a.si[(long long)p * NISLOT + I_SKIP] != 0;
  if (skip64 && a.si[(long long)p * NISLOT + I_SKIP32] != 0) return;
  const bool f32 = skip64;  // FP32 Phat for the mixed-precision Gram product
  const int k = M.k;
  const int i0 = w.y, i1 = min(M.m, w.y + chunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int slot = wave / WPC, lw = wave % WPC;  // constraint slot, wave within the slot
  const int st = threadIdx.x - slot * WPC * 64;   // thread within the slot
  // this thread's share of a staged block: global offset (or -1 for padding) and LDS offset
  int goff[NPF], loff[NPF];
#pragma unroll
  for (int j = 0; j < NPF; ++j) {
    const int e = st + WPC * 64 * j;
    const int r = e % KP, c = e / KP;
    const bool live = e < KP * KP;
    loff[j] = live ? r + c * LPA : -1;
    goff[j] = (live && r < k && c < k) ? r + c * k : -1;
  }
  for (int e = threadIdx.x; e < KP * KP; e += PT) {
    const int r = e % KP, c = e / KP;
    const int g = (r < k && c < k) ? r + c * k : -1;
    sLi[r + c * LPB] = g >= 0 ? a.LXinv[M.oK + g] : 0.0;
    sLy[r + c * LPA] = g >= 0 ? a.LY[M.oK + g] : 0.0;
  }
  const long long k2 = (long long)k * k;
  double pf[NPF];
  auto prefetch = [&](int i) {
    const double* Ai = a.Av + M.oAv + (long long)i * k2;
#pragma unroll
    for (int j = 0; j < NPF; ++j) pf[j] = (goff[j] >= 0 && i < i1) ? Ai[goff[j]] : 0.0;
  };
  double* mA = sA[slot];
  double* mT = sT[slot];
  prefetch(i0 + slot);
  // every slot runs the same number of passes (barriers); a slot past i1 idles
  for (int base = i0; base < i1; base += CP) {
    const int i = base + slot;
#pragma unroll
    for (int j = 0; j < NPF; ++j)
      if (loff[j] >= 0) mA[loff[j]] = pf[j];
    prefetch(i + CP);
    __syncthreads();
    double* Pi = (!f32 && i < i1) ? a.Ph + M.oPh + (long long)i * M.ldP : nullptr;
    float* Pf = (f32 && i < i1) ? reinterpret_cast<float*>(a.Ph + M.oPh) + (long long)i * M.ldPf : nullptr;
    if (Pf && st < M.ldPf - k2) Pf[k2 + st] = 0.f;  // padding rows read by the FP32 Gram product
    switch (lw) {
      case 0: panel_products<KP, 0>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
      case 1: panel_products<KP, 1>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
      case 2: panel_products<KP, 2>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
      case 3: panel_products<KP, 3>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
      case 4: panel_products<KP, 4>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
      case 5: panel_products<KP, 5>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
      case 6: panel_products<KP, 6>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
      default: panel_products<KP, 7>(mA, mT, sLi, sLy, Pi, Pf, k, lr, lq); break;
    }
    // no barrier here: the next pass writes mA after every wave has passed the
    // barrier between the products (its product-1 reads are done) and mT after
    // the next pass's first barrier
  }
}

// ---- Schur factor helpers ----------------------------------------------------------------------
__global__ __launch_bounds__(NT) void shift_copy_kernel(Arrays a, const int* list) {
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const double shift = a.par[(long long)p * NPARAM + P_REG];
  const double* S = a.S + M.oS;
  double* Sf = a.Sf + M.oS;
  const long long mm = (long long)M.m * M.m;
  for (long long idx = threadIdx.x; idx < mm; idx += NT) {
    const int r = (int)(idx % M.m), c = (int)(idx / M.m);
    if (r < c) continue;
    double v = S[idx];
    if (r == c && shift != 0.0) v += shift * fmax(fabs(v), 1.0);
    Sf[idx] = v;
  }
  if (threadIdx.x == 0) a.si[(long long)p * NISLOT + I_INFOS] = 0;
}

__global__ __launch_bounds__(NT) void factor_extrema_kernel(Arrays a, const int* list) {
  __shared__ double red[NT];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const double* L = a.Sf + M.oS;
  double mn = __builtin_inf(), mx = 0.0;
  for (int i = threadIdx.x; i < M.m; i += NT) {
    const double v = fabs(L[i + (long long)i * M.m]);
    mn = fmin(mn, v);
    mx = nan_max(mx, v);
  }
  mn = block_min(mn, red);
  mx = block_max(mx, red);
  if (threadIdx.x == 0) {
    a.sc[(long long)p * NSLOT + S_SFMIN] = mn;
    a.sc[(long long)p * NSLOT + S_SFMAX] = mx;
  }
}

// x = Sf^-T Sf^-1 b with the 64 x 64 diagonal-block inverses; x staged in LDS
__global__ __launch_bounds__(NT) void solve_kernel(Arrays a, const int* list, int mode) {
  __shared__ double sx[MMAX];
  __shared__ double part[4][64];
  __shared__ double t[64];
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int m = M.m;
  const double* L = a.Sf + M.oS;
  const double* dinv = a.dinv + M.oD;
  const double* b = (mode == 0 ? a.r : a.rho) + M.oM;
  double* out = (mode == 0 ? a.dx : a.delta) + M.oM;
  const int tid = threadIdx.x, row = tid % 64, q = tid / 64;
  for (int i = tid; i < m; i += NT) sx[i] = b[i];
  __syncthreads();
  // forward: x_J = Dinv_J (x_J - L_{J,<J} x_{<J})
  for (int J = 0; J < M.nblk; ++J) {
    const int r0 = J * 64, jb = min(64, m - r0);
    double s = 0.0;
    if (row < jb)
      for (int c = q; c < r0; c += 4) s += L[(r0 + row) + (long long)c * m] * sx[c];
    part[q][row] = s;
    __syncthreads();
    if (q == 0) t[row] = row < jb ? sx[r0 + row] - ((part[0][row] + part[1][row]) + (part[2][row] + part[3][row])) : 0.0;
    __syncthreads();
    const double* D = dinv + (long long)J * 64 * 64;
    s = 0.0;
    if (row < jb)
      for (int c = q; c <= row; c += 4) s += D[row + c * 64] * t[c];
    part[q][row] = s;
    __syncthreads();
    if (q == 0 && row < jb) sx[r0 + row] = (part[0][row] + part[1][row]) + (part[2][row] + part[3][row]);
    __syncthreads();
  }
  // backward: x_J = Dinv_J' (x_J - L_{>J,J}' x_{>J})
  for (int J = M.nblk - 1; J >= 0; --J) {
    const int r0 = J * 64, jb = min(64, m - r0);
    double s = 0.0;
    if (row < jb)
      for (int c = r0 + 64 + q; c < m; c += 4) s += L[c + (long long)(r0 + row) * m] * sx[c];
    part[q][row] = s;
    __syncthreads();
    if (q == 0) t[row] = row < jb ? sx[r0 + row] - ((part[0][row] + part[1][row]) + (part[2][row] + part[3][row])) : 0.0;
    __syncthreads();
    const double* D = dinv + (long long)J * 64 * 64;
    s = 0.0;
    if (row < jb)
      for (int c = row + q; c < jb; c += 4) s += D[c + row * 64] * t[c];
    part[q][row] = s;
    __syncthreads();
    if (q == 0 && row < jb) sx[r0 + row] = (part[0][row] + part[1][row]) + (part[2][row] + part[3][row]);
    __syncthreads();
  }
  for (int i = tid; i < m; i += NT) out[i] = sx[i];
}

// rho = r - S dx (S lower, unfactored); S_RHO = ||rho||_inf, S_RNORM = ||r||_inf.
// Wave w walks columns c = w, w+4, ...: column entries below the diagonal feed
// y_r (into the wave's own partial vector) and the column dot feeds y_c.
// LDS (dynamic, sized to the largest member): dx, four partial vectors, NT reduction slots.
__global__ __launch_bounds__(NT) void residual_kernel(Arrays a, const int* list) {
  extern __shared__ double dyn[];
  const int ld = a.max_m;
  double* sdx = dyn;
  double* Y = dyn + ld;  // partial vector of wave w at Y + w * ld
  double* red = dyn + 5 * ld;
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  const int m = M.m;
  const double* S = a.S + M.oS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* yw = Y + wave * ld;
  for (int i = threadIdx.x; i < m; i += NT) {
    sdx[i] = a.dx[M.oM + i];
    Y[i] = Y[ld + i] = Y[2 * ld + i] = Y[3 * ld + i] = 0.0;
  }
  __syncthreads();
  for (int c = wave; c < m; c += 4) {
    const double* col = S + (long long)c * m;
    const double xc = sdx[c];
    double dot = 0.0;
    for (int r = c + 1 + lane; r < m; r += 64) {
      const double s = col[r];
      yw[r] += s * xc;
      dot += s * sdx[r];
    }
    dot = wave_sum(dot);
    if (lane == 0) yw[c] += col[c] * xc + dot;
  }
  __syncthreads();
  double rm = 0.0, rr = 0.0;
  for (int i = threadIdx.x; i < m; i += NT) {
    const double sy = (Y[i] + Y[ld + i]) + (Y[2 * ld + i] + Y[3 * ld + i]);
    const double ri = a.r[M.oM + i];
    const double v = ri - sy;
    a.rho[M.oM + i] = v;
    rm = nan_max(rm, fabs(v));
    rr = nan_max(rr, fabs(ri));
  }
  rm = block_max(rm, red);
  rr = block_max(rr, red);
  if (threadIdx.x == 0) {
    a.sc[(long long)p * NSLOT + S_RHO] = rm;
    a.sc[(long long)p * NSLOT + S_RNORM] = rr;
  }
}

__global__ __launch_bounds__(NT) void vec_op_kernel(Arrays a, const int* list, int mode) {
  const int p = list[blockIdx.x];
  const Member M = a.mem[p];
  for (int i = threadIdx.x; i < M.m; i += NT) {
    const long long o = M.oM + i;
    if (mode == 0)
      a.dxb[o] = a.dx[o];
    else if (mode == 1)
      a.dx[o] = a.dx[o] + a.delta[o];
    else
      a.dx[o] = a.dxb[o];
  }
}

// one thread per constraint: A_j (upper COO) mirrored into its Av column
__global__ void coo_densify_kernel(Arrays a, const long long* cptr, const int* crow, const int* ccol,
                                   const double* cval, const long long* cons, int ncons) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncons) return;
  const long long off = cons[2 * j];  // Av offset of the column
  const int k = (int)cons[2 * j + 1];
  double* col = a.Av + off;
  for (long long e = cptr[j]; e < cptr[j + 1]; ++e) {
    const int r = crow[e], c = ccol[e];
    col[r + (long long)c * k] += cval[e];
    if (r != c) col[c + (long long)r * k] += cval[e];
  }
}

}  // namespace

#define SDPX_LAUNCH(kernel, grid, ...)                                                   \
  do {                                                                                   \
    if ((grid) > 0) {                                                                    \
      hipLaunchKernelGGL(kernel, dim3((unsigned)(grid)), dim3(NT), 0, st, __VA_ARGS__); \
      SDPX_HIP(hipGetLastError());                                                       \
    }                                                                                    \
    return kOk;                                                                          \
  } while (0)

int launch_gemv(const Arrays& a, const int* list, int n, int mode, hipStream_t st) {
  if (n <= 0) return kOk;
  if (a.sp_ptr)
    hipLaunchKernelGGL(sp_gemv_kernel, dim3((unsigned)n, (unsigned)ceil_div(KMAX * KMAX, kSpPositions)), dim3(NT), 0,
                       st, a, list, mode);
  else
    hipLaunchKernelGGL(gemv_part_kernel, dim3((unsigned)n, (unsigned)a.max_chunks), dim3(NT), 0, st, a, list, mode);
  SDPX_HIP(hipGetLastError());
  SDPX_LAUNCH(gemv_kernel, n, a, list, mode);
}
int launch_gemvt(const Arrays& a, const int* list, int n, int mode, hipStream_t st) {
  if (n <= 0) return kOk;
  if (a.sp_ptr)
    hipLaunchKernelGGL(sp_gemvt_kernel, dim3((unsigned)n, (unsigned)ceil_div(a.max_m, kSpConstraints)), dim3(NT), 0,
                       st, a, list, mode);
  else
    hipLaunchKernelGGL(gemvt_kernel, dim3((unsigned)n, (unsigned)a.max_chunks), dim3(NT), 0, st, a, list, mode);
  SDPX_HIP(hipGetLastError());
  if (mode == 1) return kOk;
  SDPX_LAUNCH(vec_norm_kernel, n, a, list, mode);
}
int launch_block_factor(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(block_factor_kernel, n, a, list);
}
// work items grouped by padded block size: off[c]..off[c+1] hold kp = 64, 48, 32, 16 for c = 0..3
int launch_panel(const Arrays& a, const int2* work, const int* off, int chunk, hipStream_t st) {
  const int n64 = off[1] - off[0], n48 = off[2] - off[1], n32 = off[3] - off[2], n16 = off[4] - off[3];
  if (n64 > 0) hipLaunchKernelGGL(panel_kernel<64>, dim3((unsigned)n64), dim3(PT), 0, st, a, work + off[0], chunk);
  if (n48 > 0) hipLaunchKernelGGL(panel_kernel<48>, dim3((unsigned)n48), dim3(PT), 0, st, a, work + off[1], chunk);
  if (n32 > 0) hipLaunchKernelGGL(panel_kernel<32>, dim3((unsigned)n32), dim3(PT), 0, st, a, work + off[2], chunk);
  if (n16 > 0) hipLaunchKernelGGL(panel_kernel<16>, dim3((unsigned)n16), dim3(PT), 0, st, a, work + off[3], chunk);
  SDPX_HIP(hipGetLastError());
  return kOk;
}
int launch_shift_copy(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(shift_copy_kernel, n, a, list);
}
int launch_factor_extrema(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(factor_extrema_kernel, n, a, list);
}
int launch_exact_block(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(exact_block_kernel, n, a, list);
}
int launch_rhs_block(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(rhs_block_kernel, n, a, list);
}
int launch_solve(const Arrays& a, const int* list, int n, int mode, hipStream_t st) {
  SDPX_LAUNCH(solve_kernel, n, a, list, mode);
}
int launch_residual(const Arrays& a, const int* list, int n, hipStream_t st) {
  if (n <= 0) return kOk;
  const size_t lds = sizeof(double) * (5 * (size_t)a.max_m + NT);
  hipLaunchKernelGGL(residual_kernel, dim3((unsigned)n), dim3(NT), lds, st, a, list);
  SDPX_HIP(hipGetLastError());
  return kOk;
}
int launch_vec_op(const Arrays& a, const int* list, int n, int mode, hipStream_t st) {
  SDPX_LAUNCH(vec_op_kernel, n, a, list, mode);
}
int launch_dir_block(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(dir_block_kernel, n, a, list);
}
int launch_comp(const Arrays& a, const int* list, int n, hipStream_t st) { SDPX_LAUNCH(comp_kernel, n, a, list); }
int launch_corrector_block(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(corrector_block_kernel, n, a, list);
}
// one launch per register width; members outside a width's range exit at once
// one kernel per block-size class present in the list (bit c of `classes`, step_class())
int launch_step_waves(const Arrays& a, const int* list, int n, int op, unsigned classes, hipStream_t st) {
  if (n <= 0) return kOk;
  const dim3 block(NT);
  if (classes & 1u)
    hipLaunchKernelGGL(step_small_kernel, dim3((unsigned)ceil_div(n, NT)), block, 0, st, a, list, n, op);  // k <= 2
  if (classes & 2u) hipLaunchKernelGGL(step_quad_kernel<16>, dim3((unsigned)n), block, 0, st, a, list, n, op);
  if (classes & 4u) hipLaunchKernelGGL(step_quad_kernel<32>, dim3((unsigned)n), block, 0, st, a, list, n, op);
  if (classes & 8u)
    hipLaunchKernelGGL(step_quad_wide_kernel<48>, dim3((unsigned)(2 * n)), block, 0, st, a, list, n, op);
  if (classes & 16u)
    hipLaunchKernelGGL(step_quad_wide_kernel<64>, dim3((unsigned)(2 * n)), block, 0, st, a, list, n, op);
  SDPX_HIP(hipGetLastError());
  return kOk;
}
int launch_ftb(const Arrays& a, const int* list, int n, unsigned classes, hipStream_t st) {
  return launch_step_waves(a, list, n, 0, classes, st);
}
int launch_backtrack(const Arrays& a, const int* list, int n, unsigned classes, hipStream_t st) {
  return launch_step_waves(a, list, n, 1, classes, st);
}
int launch_update(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(update_kernel, n, a, list);
}
int launch_accepted(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(accepted_kernel, n, a, list);
}
int launch_objectives(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(objectives_kernel, n, a, list);
}
int launch_scale(const Arrays& a, const int* list, int n, int which, hipStream_t st) {
  SDPX_LAUNCH(scale_kernel, n, a, list, which);
}
int launch_best(const Arrays& a, const int* list, int n, int dir, hipStream_t st) {
  SDPX_LAUNCH(best_kernel, n, a, list, dir);
}

int launch_adj_w(const Arrays& a, const int* list, int n, hipStream_t st) { SDPX_LAUNCH(adj_w_kernel, n, a, list); }
int launch_adj_final(const Arrays& a, const int* list, int n, hipStream_t st) {
  SDPX_LAUNCH(adj_final_kernel, n, a, list);
}

int launch_set_int(const Arrays& a, const int* list, int n, int slot, int value, hipStream_t st) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(set_int_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, a, list, n, slot, value);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int launch_coo_densify(const Arrays& a, const long long* cptr, const int* crow, const int* ccol, const double* cval,
                       const long long* cons, int ncons, hipStream_t st) {
  if (ncons <= 0) return kOk;
  hipLaunchKernelGGL(coo_densify_kernel, dim3((unsigned)ceil_div(ncons, 256)), dim3(256), 0, st, a, cptr, crow, ccol,
                     cval, cons, ncons);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

}  // namespace batch
}  // namespace sdpx
