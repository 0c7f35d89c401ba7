// C-ABI of the dense Schur route: one handle = one core-form SDP resident in
// HBM plus every workspace buffer of SDPX's dense Newton step, one stream.
//
// Device layout (FP64, column-major, per PSD block l of side k):
//   Av_l   k^2 x m   constraint panel (column i = vec(A_i))       [problem]
//   T, Ph  k^2 x m   Schur panels A_i M_Y and L_X^-1 A_i M_Y        [max over blocks]
//   S, Sf  m x m     Schur matrix (lower authoritative) and its Cholesky factor
//   X, Y, LX, MY, LXinv, Xinv, P, R, dX, dY, W1, W2   k x k blocks
//   trial  cap x k^2 batched line-search trial matrices
// Host code only sequences kernels and reads back the scalars of each phase.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <vector>

#include "../../include/sdpx_mi355x.h"
#include "blas_f64.hpp"

namespace sdpx {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

namespace {

constexpr int kTrialCap = 256;  // batched line-search candidates per launch
constexpr double kEps = 2.220446049250313e-16;

struct Block {
  int k = 0;
  long long k2 = 0;
  int nblk = 0;  // diagonal blocks of the block Cholesky
  double *Av = nullptr, *C = nullptr;
  double *X = nullptr, *Y = nullptr;
  double *LX = nullptr, *MY = nullptr, *LXinv = nullptr, *Xinv = nullptr;
  double *dinvX = nullptr, *dinvY = nullptr;
  double *P = nullptr, *R = nullptr, *dX = nullptr, *dY = nullptr, *W1 = nullptr, *W2 = nullptr;
  double *trial = nullptr, *trial_dinv = nullptr;
  // FP32 copy of Av for the FP32 P-hat chain (sdpx_dense_set_panel_precision) and the
  // version of Av it was converted from
  float* Av32 = nullptr;
  long long av32_ver = -1;
  // sparse block (sdpx_dense_set_block_sparse): SparseBlockCOO over na positions, 0-based
  int na = 0;
  int *ids = nullptr, *sptr = nullptr, *srow = nullptr, *scol = nullptr, *slin = nullptr;
  double *sval = nullptr, *W1T = nullptr;
};

// closed forms for blocks of side <= 2 (kernels/generic.jl:133-154, 292-340, 350-420)
#pragma clang fp contract(off)
__device__ bool small_isposdef(const double* X, const double* dX, int k, double t) {
  if (k == 1) return X[0] + t * dX[0] > 0.0;
  const double a = X[0] + t * dX[0];
  if (!(a > 0.0)) return false;
  const double b = X[1] + t * dX[1];
  const double d = X[3] + t * dX[3];
  return a * d > b * b;
}

__device__ bool small_margin(const double* X, const double* dX, int k, double t, double ratio) {
  if (!(ratio > 0.0)) return small_isposdef(X, dX, k, t);
  double mn, mx;
  if (k == 1) {
    const double d11 = X[0] + t * dX[0];
    if (!(d11 > 0.0)) return false;
    mn = mx = fabs(sqrt(d11));
  } else {
    const double d11 = X[0] + t * dX[0];
    if (!(d11 > 0.0)) return false;
    const double l11 = sqrt(d11);
    const double l21 = (X[1] + t * dX[1]) / l11;
    const double d22 = (X[3] + t * dX[3]) - l21 * l21;
    if (!(d22 > 0.0)) return false;
    const double l22 = sqrt(d22);
    mn = fmin(fabs(l11), fabs(l22));
    mx = fmax(fabs(l11), fabs(l22));
  }
  return isfinite(mn) && isfinite(mx) && mx > 0.0 && mn > ratio * mx;
}

__device__ double small_ftb(const double* X, const double* dX, int k) {
  if (k == 1) {
    const double d = dX[0];
    return d < 0.0 ? fmin(1.0, -X[0] / d) : 1.0;
  }
  const double a = X[0], b = X[2], d = X[3];
  const double da = dX[0], db = (dX[2] + dX[1]) / 2.0, dd = dX[3];
  const double a1 = a + da, b1 = b + db, d1 = d + dd;
  if (a1 > 0.0 && a1 * d1 > b1 * b1) return 1.0;
  const double c0 = a * d - b * b;
  const double c1 = da * d + a * dd - 2.0 * b * db;
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
#pragma clang fp contract(on)

__global__ void small_trials_kernel(const double* X, const double* dX, int k, const double* ts, int B, double ratio,
                                    int* ok) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B) ok[b] = small_margin(X, dX, k, ts[b], ratio) ? 1 : 0;
}

__global__ void small_ftb_kernel(const double* X, const double* dX, int k, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = small_ftb(X, dX, k);
}

// trial[b] = X + ts[b] dX for b < n0, Y + ts[b] dY otherwise
__global__ void trial_build_kernel(const double* X, const double* dX, const double* Y, const double* dY, long long k2,
                                   const double* ts, int n0, int B, double* trial) {
  const long long n = k2 * B;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long long)gridDim.x * 256) {
    const int b = (int)(idx / k2);
    const long long i = idx % k2;
    trial[idx] = b < n0 ? X[i] + ts[b] * dX[i] : Y[i] + ts[b] * dY[i];
  }
}

// B identity matrices (k x k each): stand-in trials for timing a line-search round
__global__ void identity_trials_kernel(long long k2, int k, int B, double* trial) {
  const long long n = k2 * B;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long long)gridDim.x * 256) {
    const long long i = idx % k2;
    trial[idx] = i % (k + 1) == 0 ? 1.0 : 0.0;
  }
}

// Dense lower-triangular k x k stand-in for L_X^-1 and M_Y while the autotuner
// times the panel and Gram products: entries (0.5 + u) / sqrt(k) with u in
// [0, 1) from an integer hash of (i, j), zeros above the diagonal.  With these
// the panels and the Gram operands hold O(|A_i|) values, as in a solve (the
// power drawn, and so the clock, depends on the operand bits: products of
// zeros run several percent faster than the same launch on real data).
__global__ void tune_lower_kernel(int k, double* a) {
  const long long n = (long long)k * k;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long long)gridDim.x * 256) {
    const int i = (int)(idx % k), j = (int)(idx / k);
    unsigned h = (unsigned)i * 2654435761u ^ ((unsigned)j + 0x9e3779b9u) * 40503u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    a[idx] = i >= j ? (0.5 + (double)(h >> 8) * (1.0 / 16777216.0)) * rsqrt((double)k) : 0.0;
  }
}

// corrector target (threaded.jl:839-854 / 943-1003) from the block dot in *dotv:
//   mode 0 (fixed):     max(rl < 1 ? rl^2 : rl, beta) * mu,  rl = dot / mu / k
//   mode 1 (local):     sigma * dot / k
//   mode 2 (global):    sigma * mu
__global__ void corrector_target_kernel(const double* dotv, int mode, double mu, double beta, double sigma, int k,
                                        double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double target;
  if (mode == 0) {
    const double rl = *dotv / mu / (double)k;
    const double gl = fmax(rl < 1.0 ? rl * rl : rl, beta);
    target = gl * mu;
  } else if (mode == 1) {
    target = sigma * (*dotv) / (double)k;
  } else {
    target = sigma * mu;
  }
  *out = target;
}

__global__ void neg_sum_kernel(const double* a, const double* b, double* out, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = -(a[i] + b[i]);
}

// ---- Ruiz equilibration (ingest.jl:1287-1521, dense path :1432-1521) ----------
// One pass over a block's panel (k x k x m, every A_i column-major) and C: with
// e given, every entry is scaled in place, a(r, c) *= e[r] * e[c] (the
// reference's order: the factor product first); then, or without e, the row
// maxima |a(r, .)| over every A_i and C are folded into rowmax (the bit
// patterns of non-negative doubles order like unsigned integers, so the
// device-wide max is exact and order-free).  A wavefront owns a fixed slab of
// kRuizRows rows (8 per lane, kept in registers) and walks the columns.
constexpr int kRuizRows = 512;

__global__ void __launch_bounds__(256) ruiz_pass_kernel(double* Av, double* C, int k, int m, const double* e,
                                                        unsigned long long* rowmax, int nslab) {
  const int lane = threadIdx.x & 63;
  const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long waves = (long long)gridDim.x * 4;  // a multiple of nslab
  const int slab = (int)(wave % nslab);
  const long long ncol = (long long)k * m + k;  // the m panels' columns, then C's
  const int r0 = slab * kRuizRows + lane;
  const int rend = min(k, (slab + 1) * kRuizRows);
  double mx[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) mx[j] = 0.0;
  for (long long col = wave / nslab; col < ncol; col += waves / nslab) {
    const int c = (int)(col % k);
    double* p = col < (long long)k * m ? Av + col * k : C + (col - (long long)k * m) * k;
    const double ec = e ? e[c] : 1.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int r = r0 + 64 * j;
      if (r < rend) {
        double v = p[r];
        if (e) {
          v = v * (e[r] * ec);
          p[r] = v;
        }
        mx[j] = fmax(mx[j], fabs(v));
      }
    }
  }
  if (rowmax) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int r = r0 + 64 * j;
      if (r < rend) atomicMax(rowmax + r, (unsigned long long)__double_as_longlong(mx[j]));
    }
  }
}

// maxima[i] = max(maxima[i], max |A_i|) over one block (one work-group per constraint;
// first != 0 starts from 0)
__global__ void __launch_bounds__(256) ruiz_colmax_kernel(const double* Av, long long k2, double* maxima,
                                                          int first) {
  __shared__ double red[256];
  const long long i = blockIdx.x;
  const double* a = Av + i * k2;
  double mx = 0.0;
  for (long long j = threadIdx.x; j < k2; j += 256) mx = fmax(mx, fabs(a[j]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) maxima[i] = first ? red[0] : fmax(maxima[i], red[0]);
}

// ---- sparse blocks (schur.jl:566-640, 936-1032, 1229-1318) --------------------------
// A block given as the reference's SparseBlockCOO (0-based here): positions 0..na-1 map
// to constraints ids[p]; entries ptr[p]..ptr[p+1]-1 with row, col, lin = row + col k, val,
// every stored entry of the symmetric A_p (both triangles).
//
// densify: Av[:, ids[p]] = sum of the position's entries (one thread per position, in order)
__global__ void sparse_densify_kernel(int na, const int* ids, const int* ptr, const int* lin, const double* val,
                                      long long k2, double* Av) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= na) return;
  double* a = Av + (long long)ids[p] * k2;
  for (int t = ptr[p]; t < ptr[p + 1]; ++t) a[lin[t]] += val[t];
}

// W1_p = Y A_p W2 (_two_sided_coo_product!, schur.jl:601-640): entry (i, j) of W1_p is
// sum_t Y[i, row_t] * (val_t * W2[col_t, j]), the first entry writing and the rest
// accumulating in stored order.  Stored lin-major, W1T[lin * na + p] (lin = i + j k), so the
// pair kernel reads consecutive positions.  Grid: x over lin, y over groups of 256 positions.
__global__ void sparse_w1_kernel(int k, int na, const int* ptr, const int* row, const int* col, const double* val,
                                 const double* Y, const double* W2, double* W1T) {
  const int p = blockIdx.y * 256 + threadIdx.x;
  if (p >= na) return;
  const int li = blockIdx.x;
  const int i = li % k, j = li / k;
  double acc = 0.0;
  const int t0 = ptr[p], t1 = ptr[p + 1];
  for (int t = t0; t < t1; ++t) {
    const double scaled = val[t] * W2[col[t] + (long long)j * k];
    const double term = Y[i + (long long)row[t] * k] * scaled;
    acc = t == t0 ? term : acc + term;
  }
  W1T[(long long)li * na + p] = acc;
}

// S_pr = <W1_p, A_r> for p <= r (_dot_dense_coo, schur.jl:566-579: the sum over r's entries
// of W1_p[lin_t] * val_t in stored order), accumulated into the m x m buffer Acc at
// (ids[p], ids[r]) (upper triangle: ids are sorted).  One work-group per r, threads over p.
__global__ void __launch_bounds__(256) sparse_pair_kernel(int na, const int* ids, const int* ptr, const int* lin,
                                                          const double* val, const double* W1T, double* Acc,
                                                          long long m) {
  const int r = blockIdx.x;
  const int t0 = ptr[r], t1 = ptr[r + 1];
  const long long colr = (long long)ids[r] * m;
  for (int p = threadIdx.x; p <= r; p += 256) {
    double acc = 0.0;
    for (int t = t0; t < t1; ++t) acc += W1T[(long long)lin[t] * na + p] * val[t];
    Acc[ids[p] + colr] += acc;
  }
}

// S(i, j) = (dense ? S(i, j) : 0) + Acc(j, i) for i >= j: the sparse blocks' sums, held in the
// upper triangle of Acc, added to the lower triangle of S
__global__ void add_upper_to_lower_kernel(double* S, const double* Acc, int n, int has_dense) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  if (i < j) return;
  const double v = Acc[j + (long long)i * n];
  S[idx] = has_dense ? S[idx] + v : v;
}

// ---- triangular condition estimate (mixed_precision_kkt.jl:371-392) ----------------
// colsum[j] = sum_{i >= j} |L(i, j)| (one work-group per column)
__global__ void __launch_bounds__(256) tri_colsum_kernel(const double* L, long long ld, int n, double* colsum) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int i = j + threadIdx.x; i < n; i += 256) acc += fabs(L[i + j * ld]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) colsum[j] = red[0];
}

// part[s * n + i] = sum over the columns j <= i of chunk s of |L(i, j)| (one thread per row:
// the reads of one column are coalesced across the threads)
__global__ void __launch_bounds__(256) tri_rowsum_kernel(const double* L, long long ld, int n, int chunk,
                                                        double* part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int s = blockIdx.y;
  const int j0 = s * chunk, j1 = min(min(n, j0 + chunk), i + 1);
  double acc = 0.0;
  for (int j = j0; j < j1; ++j) acc += fabs(L[i + (long long)j * ld]);
  part[(long long)s * n + i] = acc;
}

// ---- block-arrow elimination (kkt.jl:2101-2349, 2977-3096) ----------------------
constexpr int kArrowMaxLocal = 128;  // local block side held in LDS

// S(u, v) of the symmetric Schur matrix from its lower triangle
__device__ __forceinline__ double sym_lower(const double* S, long long m, int u, int v) {
  return u >= v ? S[u + v * m] : S[v + u * m];
}

// One work-group per listed local block b (q <= kArrowMaxLocal): D_b = S[U_b, U_b] gathered
// through perm, its diagonal shifted by reg[b] max(|D_ii|, 1) when reg[b] > 0 (the ladder of
// factor_arrow_kkt!, kkt.jl:2232-2260), Cholesky in LDS (info[b] = first failing column + 1,
// a non-positive, infinite or NaN pivot failing), L_b stored to Lloc; then the coupling rows
// AT[g, U_b] = C_b L_b^-T (x L_b' = C[g, U_b] by forward substitution, one thread per global g;
// the reference's W_b = D_b^-1 C_b is L_b^-T of these rows).
__global__ void __launch_bounds__(256) arrow_local_factor_kernel(const double* S, int m, const int* perm,
                                                                 const int* loc_off, const long long* loc_foff,
                                                                 const int* list, const double* reg, double* Lloc,
                                                                 int* info, double* AT, int ng, int mloc) {
  extern __shared__ double A[];
  __shared__ int fail;
  const int b = list[blockIdx.x];
  const int o = loc_off[b], q = loc_off[b + 1] - o, pq = q + 1;
  const double shift = reg[b];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < q * q; idx += 256) {
    const int i = idx % q, j = idx / q;
    if (i < j) continue;
    double a = sym_lower(S, m, perm[o + i], perm[o + j]);
    if (i == j && shift > 0.0) a += shift * fmax(fabs(a), 1.0);
    A[i + j * pq] = a;
  }
  if (tid == 0) fail = 0;
  __syncthreads();
  for (int j = 0; j < q; ++j) {
    if (tid == 0) {
      const double d = A[j + j * pq];
      if (!(d > 0.0) || isinf(d)) {
        fail = j + 1;
      } else {
        A[j + j * pq] = sqrt(d);
      }
    }
    __syncthreads();
    if (fail) break;
    const double djj = A[j + j * pq];
    for (int i = j + 1 + tid; i < q; i += 256) A[i + j * pq] /= djj;
    __syncthreads();
    const int r = q - j - 1;  // trailing rows / columns j+1 .. q-1, lower entries
    for (int idx = tid; idx < r * r; idx += 256) {
      const int i = j + 1 + idx % r, c = j + 1 + idx / r;
      if (i >= c) A[i + c * pq] -= A[i + j * pq] * A[c + j * pq];
    }
    __syncthreads();
  }
  if (tid == 0) info[b] = fail;
  if (fail) return;
  double* Lb = Lloc + loc_foff[b];
  for (int idx = tid; idx < q * q; idx += 256) {
    const int i = idx % q, j = idx / q;
    Lb[idx] = i >= j ? A[i + j * pq] : 0.0;
  }
  for (int g = tid; g < ng; g += 256) {
    const int vg = perm[mloc + g];
    double* row = AT + g;  // row g of AT: entries (g, o + j) at row[(o + j) * ng]
    for (int j = 0; j < q; ++j) {
      double t = sym_lower(S, m, vg, perm[o + j]);
      for (int i = 0; i < j; ++i) t -= row[(long long)(o + i) * ng] * A[j + i * pq];
      row[(long long)(o + j) * ng] = t / A[j + j * pq];
    }
  }
}

// Sred(a, c) = S(G_a, G_c) for a >= c (ld ng)
__global__ void arrow_gather_global_kernel(const double* S, int m, const int* perm, int mloc, int ng,
                                           double* Sred) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)ng * ng) return;
  const int a = (int)(idx % ng), c = (int)(idx / ng);
  if (a < c) return;
  Sred[idx] = sym_lower(S, m, perm[mloc + a], perm[mloc + c]);
}

// xp[i] = x[perm[i]] (dir 0) or x[perm[i]] = xp[i] (dir 1)
__global__ void arrow_permute_kernel(double* x, const int* perm, double* xp, int m, int dir) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  if (dir == 0)
    xp[i] = x[perm[i]];
  else
    x[perm[i]] = xp[i];
}

// One work-group per local block: xp[U_b] <- L_b^-1 xp[U_b] (trans 0) or L_b^-T xp[U_b] (trans 1),
// the vector in LDS, one column (or row) of L_b per step
__global__ void __launch_bounds__(256) arrow_local_solve_kernel(const double* Lloc, const int* loc_off,
                                                                const long long* loc_foff, double* xp, int trans) {
  __shared__ double xs[kArrowMaxLocal];
  const int b = blockIdx.x;
  const int o = loc_off[b], q = loc_off[b + 1] - o;
  const double* L = Lloc + loc_foff[b];
  const int tid = threadIdx.x;
  if (tid < q) xs[tid] = xp[o + tid];
  __syncthreads();
  if (trans == 0) {
    for (int j = 0; j < q; ++j) {
      if (tid == 0) xs[j] /= L[j + (long long)j * q];
      __syncthreads();
      if (tid > j && tid < q) xs[tid] -= L[tid + (long long)j * q] * xs[j];
      __syncthreads();
    }
  } else {
    for (int j = q - 1; j >= 0; --j) {
      if (tid == 0) xs[j] /= L[j + (long long)j * q];
      __syncthreads();
      if (tid < j) xs[tid] -= L[j + (long long)tid * q] * xs[j];
      __syncthreads();
    }
  }
  if (tid < q) xp[o + tid] = xs[tid];
}

// ---- equality-constrained KKT (kkt.jl:1466-1665, 2570-2595) ----------------------
// dst(r, c) = src(r, c) for an rows x cols block
__global__ void copy_block_kernel(const double* src, long long lds, double* dst, long long ldd, int rows, int cols) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)rows * cols) return;
  const int r = (int)(idx % rows), c = (int)(idx / rows);
  dst[r + c * ldd] = src[r + c * lds];
}

// _normalize_equality_panel_columns! (kkt.jl:281-304): per column, scale = 1 / ||col||_inf
// (1 for a zero column or a non-finite reciprocal), applied in place
__global__ void __launch_bounds__(256) normalize_columns_kernel(double* B, long long ld, int rows, double* scale) {
  __shared__ double red[256];
  double* col = B + (long long)blockIdx.x * ld;
  double mx = 0.0;
  for (int r = threadIdx.x; r < rows; r += 256) mx = fmax(mx, fabs(col[r]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  double sc = red[0] == 0.0 ? 1.0 : 1.0 / red[0];
  if (!isfinite(sc)) sc = 1.0;
  if (threadIdx.x == 0) scale[blockIdx.x] = sc;
  if (sc == 1.0) return;
  for (int r = threadIdx.x; r < rows; r += 256) col[r] *= sc;
}

// q = scale .* p - q   (_form_scaled_equality_rhs!, kkt.jl:346-357), or dy = scale .* dy
__global__ void scaled_rhs_kernel(const double* scale, const double* p, double* q, int n, int mode) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (mode == 0) q[i] = scale[i] * p[i] - q[i];
  else q[i] = q[i] * scale[i];
}

// A_i ./= s[i] (the per-variable scale)
__global__ void div_columns_kernel(double* Av, long long k2, int m, const double* s) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= k2 * m) return;
  Av[idx] = Av[idx] / s[idx / k2];
}

}  // namespace
}  // namespace sdpx

using namespace sdpx;

struct sdpx_dense {
  int device = 0;
  int m = 0;
  int L = 0;
  std::vector<Block> blk;
  hipStream_t stream = nullptr;
  double *c = nullptr, *x = nullptr, *dx = nullptr, *dx_best = nullptr, *d = nullptr, *v = nullptr, *r = nullptr;
  double *rho = nullptr, *delta = nullptr, *v2 = nullptr;
  double *S = nullptr, *Sf = nullptr, *dinvS = nullptr, *trsv_y = nullptr;
  double *T = nullptr, *Ph = nullptr;  // Schur panels (max over blocks)
  double* splitws = nullptr;
  int split = 1;
  int2* tile_order = nullptr;  // Schur Gram lower-tile visiting order (FP32 / BF16x3 kernels)
  int2* gram_items = nullptr;  // FP64 Schur Gram work list (tiles x split_max entries)
  // mixed precision (sdpx_dense_set_precision): FP32 copy of the Schur panel
  // (precision 1) or its bf16 hi / lo planes (precision 2)
  int precision = 0;
  // tile-shape flags of the two P-hat panel products (autotuned).  The
  // create-time default is 64 x 64 for both: K <= k is short and the triangular
  // operand ends every K range on a 64 boundary (at n = 400 it ran 5.5 ms
  // against 10.5 ms on 128 x 128 tiles, within 8% of the fastest shape)
  unsigned panel_tiles[2] = {kGemmSmallTiles, kGemmSmallTiles};
  int tile_group = 8;        // super-block side of the Gram tile order
  int ftb_depth = 7;         // bisection levels per batched trial round (fraction to the boundary)
  int split_created = 1;     // split-K factor chosen at create
  int split_max = 1;         // partial sums the workspace holds (autotuner candidates up to this)
  float tune_panel_ms = 0.f, tune_gram_ms = 0.f;
  float tune_ms[2][4] = {};  // panel products x tile shapes (128x128, 64x64, 128x64, 64x128)
  float tune_ftb_ms[4] = {};  // modelled fraction-to-the-boundary search, D = 4..7 levels per round
  float* Pf = nullptr;
  long long ldPf = 0, maxk2 = 0;
  // FP32 factor of the Schur matrix (sdpx_dense_set_factor_precision): the FP32 matrix (ld ldS32),
  // a 64 x 64 FP64 block and a 64 x 64 FP32 inverse for the diagonal-block steps
  int factor_precision = 0;
  float* S32 = nullptr;
  long long ldS32 = 0;
  double* blk64 = nullptr;
  float* dinv32 = nullptr;
  __bf16 *Ph_hi = nullptr, *Ph_lo = nullptr;
  long long ldPb = 0;
  // FP32 P-hat chain of the BF16x3 assembly (sdpx_dense_set_panel_precision): FP32 copies of
  // M_Y and L_X^-1 (max k^2 floats each); av_ver counts the writes to any block's Av
  int panel_precision = 0;
  float *MY32 = nullptr, *LX32 = nullptr;
  long long av_ver = 0;
  double* gemv_part = nullptr;
  long long gemv_part_cap = 0;
  double* symv_part = nullptr;
  double* red_part = nullptr;  // reduction partials
  double* dsc = nullptr;       // device scalars
  double* hsc = nullptr;       // pinned host scalars
  int* dinfo = nullptr;
  int* hinfo = nullptr;
  double* dts = nullptr;  // line-search candidates
  double* hts = nullptr;
  int* dok = nullptr;
  long long bytes = 0;
  std::vector<void*> allocs;
  int last_reg_attempts = 0;
  // equality constraints B' x = p (n_eq columns of B, m x n_eq): B, B-tilde = L_S^-1 B D,
  // Q = B-tilde' B-tilde and its factor, the column scales D, solve vectors
  int neq = 0;
  bool eq_factored = false;
  double *B = nullptr, *Bt = nullptr, *Btmp = nullptr, *Q = nullptr, *Qf = nullptr, *dinvQ = nullptr;
  double *eqscale = nullptr, *qrhs = nullptr, *qy = nullptr, *rtil = nullptr, *peq = nullptr;
  // block-arrow elimination (sdpx_dense_set_arrow): permuted order = the local variables of each
  // block (ascending, block by block), then the global ones.  Lloc: local factors packed
  // (q_b x q_b column-major each); AT = C L_U^-T (ng x mloc, ld ng); Sred = S_GG - AT AT' (ng x ng,
  // ld ng), factored into Sf with ld ng
  bool arrow = false;
  int ng = 0, mloc = 0, nloc = 0, max_q = 0;
  int *perm = nullptr, *loc_off = nullptr, *loc_list = nullptr, *loc_info = nullptr;
  long long* loc_foff = nullptr;
  double *loc_reg = nullptr, *Lloc = nullptr, *AT = nullptr, *Sred = nullptr, *xp = nullptr;
  double* arrow_part = nullptr;
  long long arrow_part_cap = 0;
  double* cond_part = nullptr;  // row-sum partials of the condition estimate (32 x m)
  // device-event kernel timing (sdpx_dense_set_profiling)
  bool prof = false;
  hipEvent_t ev[6] = {};
  double t_panels = 0.0, t_gram = 0.0, t_reduce = 0.0, t_factor = 0.0;
  long long n_gram = 0, n_factor = 0;
};

namespace {

constexpr int kScalars = 1024;

int dmalloc(sdpx_dense_t* h, void** p, size_t bytes) {
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
int dalloc(sdpx_dense_t* h, T** p, size_t count) {
  return dmalloc(h, reinterpret_cast<void**>(p), count * sizeof(T));
}

// Gram visiting orders for `group` x `group` super-blocks and `split` K ranges:
// the lower-tile order of the FP32 / BF16x3 kernels and the FP64 kernel's work
// list (full tiles of each XCD's run first, shorter items last).
int set_gram_order(sdpx_dense_t* h, int group, int split) {
  const int nt = ceil_div(h->m, 128);
  std::vector<int2> order((size_t)nt * (nt + 1) / 2), items(order.size() * (size_t)split);
  lower_tile_order(nt, group, order.data());
  gram_item_order(h->m, group, split, items.data());
  SDPX_HIP(hipStreamSynchronize(h->stream));  // no launch in flight reads the old lists
  SDPX_HIP(hipMemcpy(h->tile_order, order.data(), order.size() * sizeof(int2), hipMemcpyHostToDevice));
  SDPX_HIP(hipMemcpy(h->gram_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
  return kOk;
}

int sync_scalars(sdpx_dense_t* h, int ndoubles, int nints = 0) {
  if (ndoubles > 0)
    SDPX_HIP(hipMemcpyAsync(h->hsc, h->dsc, sizeof(double) * ndoubles, hipMemcpyDeviceToHost, h->stream));
  if (nints > 0) SDPX_HIP(hipMemcpyAsync(h->hinfo, h->dinfo, sizeof(int) * nints, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

GemmArgs gemm(int M, int N, int K, double alpha, const double* A, long long lda, const double* B, long long ldb,
              double beta, double* C, long long ldc, unsigned flags) {
  GemmArgs g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.batch = 1;
  g.flags = flags;
  return g;
}

// C = alpha A B + beta C for k x k blocks
int mm(sdpx_dense_t* h, int k, double alpha, const double* A, bool ta, const double* B, bool tb, double beta,
       double* C, unsigned extra = 0) {
  return gemm_f64(gemm(k, k, k, alpha, A, k, B, k, beta, C, k,
                       (ta ? kGemmTransA : 0u) | (tb ? kGemmTransB : 0u) | extra),
                  h->stream);
}

int check_handle(sdpx_dense_t* h) {
  if (!h) {
    set_last_error("null handle");
    return kInvalidArgument;
  }
  SDPX_HIP(hipSetDevice(h->device));
  return kOk;
}

// Z = X^-1 (P Y - R) for block b into W2 (predictor/corrector RHS, step.jl:95-106)
int block_rhs_z(sdpx_dense_t* h, Block& b, const double* Y) {
  SDPX_CHECK(mm(h, b.k, 1.0, b.P, false, Y, false, 0.0, b.W1));
  SDPX_CHECK(axpbyz_f64(b.k2, 1.0, b.W1, -1.0, b.R, b.W1, h->stream));
  SDPX_CHECK(mm(h, b.k, 1.0, b.Xinv, false, b.W1, false, 0.0, b.W2));
  return kOk;
}

int rhs_from_blocks(sdpx_dense_t* h) {
  for (int l = 0; l < h->L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(block_rhs_z(h, b, b.Y));
    SDPX_CHECK(avt_f64(b.Av, b.k2, h->m, b.W2, 1.0, l == 0 ? 0.0 : 1.0, h->v, h->stream));
  }
  hipLaunchKernelGGL(neg_sum_kernel, dim3(ceil_div(h->m, 256)), dim3(256), 0, h->stream, h->d, h->v, h->r, h->m);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

// rho = r - S dx; dsc[slot] = ||rho||_inf, dsc[slot+1] = ||r||_inf
int direction_residual_dev(sdpx_dense_t* h, int slot) {
  SDPX_CHECK(symv_lower_f64(h->S, h->m, h->dx, -1.0, 1.0, h->r, h->rho, h->symv_part, h->stream));
  SDPX_CHECK(absmax_f64(h->rho, h->m, h->red_part, h->dsc + slot, h->stream));
  SDPX_CHECK(absmax_f64(h->r, h->m, h->red_part, h->dsc + slot + 1, h->stream));
  return kOk;
}

// solve_arrow_kkt! (kkt.jl:2977-3096) with the factors of arrow_factor: x = P r; x_U <- L_U^-1 x_U;
// x_G -= AT x_U; x_G <- Sred^-1 x_G; x_U -= AT' x_G; x_U <- L_U^-T x_U; out = P' x
int arrow_solve(sdpx_dense_t* h, double* out) {
  const int m = h->m, mloc = h->mloc, ng = h->ng;
  hipLaunchKernelGGL(arrow_permute_kernel, dim3(ceil_div(m, 256)), dim3(256), 0, h->stream, out, h->perm, h->xp, m, 0);
  SDPX_HIP(hipGetLastError());
  hipLaunchKernelGGL(arrow_local_solve_kernel, dim3(h->nloc), dim3(256), 0, h->stream, h->Lloc, h->loc_off,
                     h->loc_foff, h->xp, 0);
  SDPX_HIP(hipGetLastError());
  if (ng > 0) {
    double* xg = h->xp + mloc;
    SDPX_CHECK(avx_f64(h->AT, ng, mloc, h->xp, -1.0, xg, 1.0, xg, 0.0, nullptr, h->arrow_part, h->arrow_part_cap,
                       h->stream));
    SDPX_CHECK(trsv_f64_steps(h->Sf, ng, ng, h->dinvS, xg, h->trsv_y, 0, h->stream));
    SDPX_CHECK(trsv_f64_steps(h->Sf, ng, ng, h->dinvS, xg, h->trsv_y, 1, h->stream));
    SDPX_CHECK(avt_f64(h->AT, ng, mloc, xg, -1.0, 1.0, h->xp, h->stream));
  }
  hipLaunchKernelGGL(arrow_local_solve_kernel, dim3(h->nloc), dim3(256), 0, h->stream, h->Lloc, h->loc_off,
                     h->loc_foff, h->xp, 1);
  SDPX_HIP(hipGetLastError());
  hipLaunchKernelGGL(arrow_permute_kernel, dim3(ceil_div(m, 256)), dim3(256), 0, h->stream, out, h->perm, h->xp, m, 1);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

// factor_arrow_kkt! (kkt.jl:2132-2349): every local block factored with its own ladder (the
// failing ones re-launched with the next shift), the coupling rows solved in the same launch,
// Sred = S_GG - AT AT' on MFMA, then the ladder of _factor_with_relative_regularization!
// (kkt.jl:1695-1716) on Sred.  *attempts = the sum of every ladder's attempts.
int arrow_factor(sdpx_dense_t* h, int* attempts_out, bool* ok_out) {
  const int m = h->m, nloc = h->nloc, ng = h->ng, mloc = h->mloc;
  std::vector<int> list(nloc), attempts(nloc, 0), info(nloc, 0);
  std::vector<double> reg(nloc, 0.0);
  for (int b = 0; b < nloc; ++b) list[b] = b;
  int total = 0;
  bool ok = true;
  const size_t lds = sizeof(double) * (size_t)h->max_q * (h->max_q + 1);
  while (!list.empty()) {
    SDPX_HIP(hipMemcpyAsync(h->loc_list, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, h->stream));
    SDPX_HIP(hipMemcpyAsync(h->loc_reg, reg.data(), sizeof(double) * nloc, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(arrow_local_factor_kernel, dim3((unsigned)list.size()), dim3(256), lds, h->stream, h->S, m,
                       h->perm, h->loc_off, h->loc_foff, h->loc_list, h->loc_reg, h->Lloc, h->loc_info, h->AT, ng,
                       mloc);
    SDPX_HIP(hipGetLastError());
    SDPX_HIP(hipMemcpyAsync(info.data(), h->loc_info, sizeof(int) * nloc, hipMemcpyDeviceToHost, h->stream));
    SDPX_HIP(hipStreamSynchronize(h->stream));
    std::vector<int> next;
    for (int b : list) {
      if (info[b] == 0) continue;
      if (attempts[b] >= 6) {
        ok = false;
        continue;
      }
      attempts[b] += 1;
      total += 1;
      reg[b] = attempts[b] == 1 ? std::sqrt(kEps) : reg[b] * 10.0;
      next.push_back(b);
    }
    if (!ok) break;
    list.swap(next);
  }
  if (ok && ng > 0) {
    hipLaunchKernelGGL(arrow_gather_global_kernel, dim3(ceil_div((long long)ng * ng, 256)), dim3(256), 0, h->stream,
                       h->S, m, h->perm, mloc, ng, h->Sred);
    SDPX_HIP(hipGetLastError());
    SDPX_CHECK(gemm_f64(gemm(ng, ng, mloc, -1.0, h->AT, ng, h->AT, ng, 1.0, h->Sred, ng,
                             kGemmTransB | kGemmLowerOut),
                        h->stream));
    int att = 0;
    double shift = 0.0;
    while (true) {
      SDPX_CHECK(copy_lower_shift_f64(h->Sred, h->Sf, ng, att == 0 ? 0.0 : shift, h->stream));
      SDPX_CHECK(potrf_f64(h->Sf, ng, 0, ng, 1, h->dinvS, 0, h->dinfo, h->stream));
      SDPX_CHECK(sync_scalars(h, 0, 1));
      if (h->hinfo[0] == 0) break;
      if (att >= 6) {
        ok = false;
        break;
      }
      att += 1;
      shift = att == 1 ? std::sqrt(kEps) : shift * 10.0;
    }
    total += att;
  }
  *attempts_out = total;
  *ok_out = ok;
  return kOk;
}

// Higham's one-norm estimator (the LAPACK xLACN2 iteration behind xTRCON; N. J. Higham,
// ACM TOMS 14 (1988) 381-396, Algorithm 4.1): an estimate of ||B||_1 from products with B and
// B' only, at most five rounds plus the alternating-sign test vector.  apply(x, t) overwrites
// the host vector x with B x (t = 0) or B' x (t = 1).
template <typename Apply>
int onenorm_estimate(int n, Apply apply, double* est_out) {
  std::vector<double> x(n, 1.0 / n), v;
  std::vector<int> sgn(n);
  auto sign = [](double a) { return a >= 0.0 ? 1 : -1; };
  auto asum = [](const std::vector<double>& a) {
    double t = 0.0;
    for (double e : a) t += std::fabs(e);
    return t;
  };
  auto argmax = [&](const std::vector<double>& a) {
    int j = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(a[i]) > std::fabs(a[j])) j = i;
    return j;
  };
  SDPX_CHECK(apply(x, 0));
  if (n == 1) {
    *est_out = std::fabs(x[0]);
    return kOk;
  }
  double est = asum(x);
  for (int i = 0; i < n; ++i) {
    sgn[i] = sign(x[i]);
    x[i] = sgn[i];
  }
  SDPX_CHECK(apply(x, 1));
  int j = argmax(x);
  for (int iter = 2;; ++iter) {
    std::fill(x.begin(), x.end(), 0.0);
    x[j] = 1.0;
    SDPX_CHECK(apply(x, 0));
    const double old = est;
    est = asum(x);
    bool same = true;
    for (int i = 0; i < n && same; ++i) same = sign(x[i]) == sgn[i];
    if (same || est <= old) break;
    for (int i = 0; i < n; ++i) {
      sgn[i] = sign(x[i]);
      x[i] = sgn[i];
    }
    SDPX_CHECK(apply(x, 1));
    const int jlast = j;
    j = argmax(x);
    if (!(x[jlast] != std::fabs(x[j]) && iter < 5)) break;
  }
  double alt = 1.0;
  for (int i = 0; i < n; ++i) {
    x[i] = alt * (1.0 + (double)i / (n - 1));
    alt = -alt;
  }
  SDPX_CHECK(apply(x, 0));
  const double temp = 2.0 * asum(x) / (3.0 * n);
  *est_out = std::max(est, temp);
  return kOk;
}

int solve_into(sdpx_dense_t* h, const double* rhs, double* out) {
  if (rhs != out) SDPX_HIP(hipMemcpyAsync(out, rhs, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream));
  if (h->arrow) return arrow_solve(h, out);
  SDPX_CHECK(trsv_f64_steps(h->Sf, h->m, h->m, h->dinvS, out, h->trsv_y, 0, h->stream));
  SDPX_CHECK(trsv_f64_steps(h->Sf, h->m, h->m, h->dinvS, out, h->trsv_y, 1, h->stream));
  return kOk;
}

// Batched trial PSD tests for one block: candidates t0[0..n0) on X + t dX,
// then t1[0..n1) on Y + t dY, all factored in one batched launch sequence.
// pass[j] = 1 when the trial passes trial_has_cholesky_margin!
// (kernels/generic.jl:318-340; ratio = 0 is plain positive definiteness).
int trial_eval(sdpx_dense_t* h, Block& b, const std::vector<double>& t0, const std::vector<double>& t1, double ratio,
               std::vector<int>& pass) {
  const int n0 = (int)t0.size(), n1 = (int)t1.size(), B = n0 + n1;
  pass.assign(B, 0);
  if (B == 0) return kOk;
  if (B > kTrialCap) {
    set_last_error("trial_eval: too many line-search candidates");
    return kInvalidArgument;
  }
  for (int j = 0; j < n0; ++j) h->hts[j] = t0[j];
  for (int j = 0; j < n1; ++j) h->hts[n0 + j] = t1[j];
  SDPX_HIP(hipMemcpyAsync(h->dts, h->hts, sizeof(double) * B, hipMemcpyHostToDevice, h->stream));
  if (b.k <= 2) {
    if (n0) hipLaunchKernelGGL(small_trials_kernel, dim3(1), dim3(kTrialCap), 0, h->stream, b.X, b.dX, b.k, h->dts, n0,
                               ratio, h->dok);
    if (n1) hipLaunchKernelGGL(small_trials_kernel, dim3(1), dim3(kTrialCap), 0, h->stream, b.Y, b.dY, b.k,
                               h->dts + n0, n1, ratio, h->dok + n0);
    SDPX_HIP(hipGetLastError());
    SDPX_HIP(hipMemcpyAsync(h->hinfo, h->dok, sizeof(int) * B, hipMemcpyDeviceToHost, h->stream));
    SDPX_HIP(hipStreamSynchronize(h->stream));
    for (int j = 0; j < B; ++j) pass[j] = h->hinfo[j];
    return kOk;
  }
  const long long n = b.k2 * B;
  hipLaunchKernelGGL(trial_build_kernel, dim3(std::min(8192, ceil_div(n, 256))), dim3(256), 0, h->stream, b.X, b.dX,
                     b.Y, b.dY, b.k2, h->dts, n0, B, b.trial);
  SDPX_HIP(hipGetLastError());
  const long long dstride = (long long)b.nblk * kCholNb * kCholNb;
  SDPX_CHECK(potrf_f64(b.trial, b.k, b.k2, b.k, B, b.trial_dinv, dstride, h->dok, h->stream));
  if (ratio > 0.0) SDPX_CHECK(diag_extrema_f64(b.trial, b.k, b.k2, b.k, B, h->dsc, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->hinfo, h->dok, sizeof(int) * B, hipMemcpyDeviceToHost, h->stream));
  if (ratio > 0.0) SDPX_HIP(hipMemcpyAsync(h->hsc, h->dsc, sizeof(double) * 2 * B, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  for (int j = 0; j < B; ++j) {
    bool ok = h->hinfo[j] == 0;
    if (ok && ratio > 0.0) {
      const double mn = h->hsc[2 * j], mx = h->hsc[2 * j + 1];
      ok = std::isfinite(mn) && std::isfinite(mx) && mx > 0.0 && mn > ratio * mx;
    }
    pass[j] = ok ? 1 : 0;
  }
  return kOk;
}

// Heap-ordered tree of the midpoints the sequential bisection can visit in the
// next `depth` levels from (lo, hi), computed with the same FP64 arithmetic.
void bisection_tree(double lo, double hi, int depth, std::vector<double>& mids) {
  const int nodes = (1 << depth) - 1;
  std::vector<double> nlo(2 * nodes + 2), nhi(2 * nodes + 2);
  mids.assign(nodes, 0.0);
  nlo[1] = lo;
  nhi[1] = hi;
  for (int node = 1; node <= nodes; ++node) {
    const double mid = (nlo[node] + nhi[node]) / 2.0;
    mids[node - 1] = mid;
    nlo[2 * node] = nlo[node];
    nhi[2 * node] = mid;
    nlo[2 * node + 1] = mid;
    nhi[2 * node + 1] = nhi[node];
  }
}

// Once lo and hi are adjacent doubles, every further bisection level tests
// (lo + hi) / 2, which is lo (known to pass) or hi (known to fail): the
// remaining levels change nothing and the search can stop with lo.
bool bounds_adjacent(double lo, double hi) {
  const double mid = (lo + hi) / 2.0;
  return mid == lo || mid == hi;
}

// Replays `depth` bisection levels from the tested tree (pass flags at offset).
void bisection_replay(const std::vector<double>& mids, const std::vector<int>& pass, int offset, int depth, double& lo,
                      double& hi) {
  int node = 1;
  for (int lvl = 0; lvl < depth; ++lvl) {
    const double mid = mids[node - 1];
    if (pass[offset + node - 1]) {
      lo = mid;
      node = 2 * node + 1;
    } else {
      hi = mid;
      node = 2 * node;
    }
  }
}

// fraction_to_boundary_bound! (kernels/generic.jl:350-420) of X + t dX and
// Y + t dY for one block: the full-step test, then the 64-step bisection.  The
// midpoints of several levels of both sides are tested in one batched launch
// and the sequential path is replayed from the outcomes, so the result equals
// the sequential bisection's (the first round speculatively includes the
// first six levels next to the full step).
int ftb_block(sdpx_dense_t* h, Block& b, double* bx, double* by) {
  if (b.k <= 2) {
    hipLaunchKernelGGL(small_ftb_kernel, dim3(1), dim3(64), 0, h->stream, b.X, b.dX, b.k, h->dsc);
    hipLaunchKernelGGL(small_ftb_kernel, dim3(1), dim3(64), 0, h->stream, b.Y, b.dY, b.k, h->dsc + 1);
    SDPX_HIP(hipGetLastError());
    SDPX_CHECK(sync_scalars(h, 2));
    *bx = h->hsc[0];
    *by = h->hsc[1];
    return kOk;
  }
  // D levels (2^D - 1 candidates per side) per round; round 0 takes the full
  // step and D - 1 levels.  D is a pure throughput choice (the replay gives
  // the sequential search's result for any D): 7 unless the autotuner set it.
  const int D = h->ftb_depth, D0 = h->ftb_depth - 1;
  double lo[2] = {0.0, 0.0}, hi[2] = {1.0, 1.0}, res[2] = {1.0, 1.0};
  int remaining[2] = {64, 64};
  bool active[2] = {true, true};
  std::vector<double> mids[2], cand[2];
  std::vector<int> pass;
  // round 0: full step + first D0 levels
  for (int sd = 0; sd < 2; ++sd) {
    bisection_tree(0.0, 1.0, D0, mids[sd]);
    cand[sd].assign(1, 1.0);
    cand[sd].insert(cand[sd].end(), mids[sd].begin(), mids[sd].end());
  }
  SDPX_CHECK(trial_eval(h, b, cand[0], cand[1], 0.0, pass));
  for (int sd = 0; sd < 2; ++sd) {
    const int off = sd == 0 ? 0 : (int)cand[0].size();
    if (pass[off]) {
      active[sd] = false;
      res[sd] = 1.0;
    } else {
      bisection_replay(mids[sd], pass, off + 1, D0, lo[sd], hi[sd]);
      remaining[sd] -= D0;
      if (bounds_adjacent(lo[sd], hi[sd])) {
        active[sd] = false;
        res[sd] = lo[sd];
      }
    }
  }
  while (active[0] || active[1]) {
    int depth[2] = {0, 0};
    for (int sd = 0; sd < 2; ++sd) {
      cand[sd].clear();
      if (!active[sd]) continue;
      depth[sd] = std::min(D, remaining[sd]);
      bisection_tree(lo[sd], hi[sd], depth[sd], mids[sd]);
      cand[sd] = mids[sd];
    }
    SDPX_CHECK(trial_eval(h, b, cand[0], cand[1], 0.0, pass));
    for (int sd = 0; sd < 2; ++sd) {
      if (!active[sd]) continue;
      const int off = sd == 0 ? 0 : (int)cand[0].size();
      bisection_replay(mids[sd], pass, off, depth[sd], lo[sd], hi[sd]);
      remaining[sd] -= depth[sd];
      if (remaining[sd] <= 0 || bounds_adjacent(lo[sd], hi[sd])) {
        active[sd] = false;
        res[sd] = lo[sd];
      }
    }
  }
  *bx = res[0];
  *by = res[1];
  return kOk;
}

// line_search! (step.jl:1025-1075) for both sides at once: candidates
// t_{j+1} = t_j * gamma (sequential arithmetic) are tested in batches; a side
// stops at its first passing candidate or its first candidate below min_step.
int backtrack_both(sdpx_dense_t* h, double gamma, double min_step, double ratio, double* tX, double* tY) {
  double next[2] = {1.0, 1.0}, res[2] = {1.0, 1.0};
  bool active[2] = {true, true};
  int batch = 8;  // the first candidates usually pass: start small, then widen
  std::vector<double> cand[2];
  std::vector<int> pass, all;
  while (active[0] || active[1]) {
    for (int sd = 0; sd < 2; ++sd) {
      cand[sd].clear();
      if (!active[sd]) continue;
      double tc = next[sd];
      for (int j = 0; j < batch; ++j) {
        cand[sd].push_back(tc);
        if (tc < min_step) break;
        tc *= gamma;
      }
    }
    const int n0 = (int)cand[0].size(), B = n0 + (int)cand[1].size();
    all.assign(B, 1);
    for (int l = 0; l < h->L; ++l) {
      SDPX_CHECK(trial_eval(h, h->blk[l], cand[0], cand[1], ratio, pass));
      for (int j = 0; j < B; ++j) all[j] = all[j] && pass[j];
    }
    for (int sd = 0; sd < 2; ++sd) {
      if (!active[sd]) continue;
      const int off = sd == 0 ? 0 : n0;
      for (size_t j = 0; j < cand[sd].size(); ++j) {
        if (all[off + j] || cand[sd][j] < min_step) {
          res[sd] = cand[sd][j];
          active[sd] = false;
          break;
        }
      }
      if (active[sd]) next[sd] = cand[sd].back() * gamma;
    }
    batch = kTrialCap / 2;
  }
  *tX = res[0];
  *tY = res[1];
  return kOk;
}

// ---- cold-start helpers on host (cold_start.jl:24-49, 141-312, 314-428) --------------
namespace cold {

double safety(double scale) { return std::sqrt(2.220446049250313e-16) * std::max(1.0, scale); }
double rounding_slack(double scale) { return 8.0 * 2.220446049250313e-16 * std::max(1.0, scale); }

// certified lower bound on lambda_min from the lower triangle (cold_start.jl:141-181)
double psd_lower_bound(const double* A, int n) {
  if (n == 1) return A[0];
  if (n == 2) {
    const double first = A[0], second = A[3], off = A[1];
    const double center = (first + second) / 2.0;
    const double radius = std::hypot((first - second) / 2.0, off);
    return center - radius;
  }
  double bound = 0.0;
  for (int row = 0; row < n; ++row) {
    double radius = 0.0;
    for (int col = 0; col < n; ++col) {
      if (col == row) continue;
      const int lr = std::max(row, col), lc = std::min(row, col);
      radius += std::fabs(A[lr + (long long)lc * n]);
    }
    const double cand = A[row + (long long)row * n] - radius;
    bound = row == 0 ? cand : std::min(bound, cand);
  }
  return bound;
}

// finite max-norm of the lower triangle and the certified margin (cold_start.jl:184-201)
bool psd_stats(const double* A, int n, double& scale, double& margin) {
  scale = 0.0;
  for (int c = 0; c < n; ++c)
    for (int r = c; r < n; ++r) {
      const double v = A[r + (long long)c * n];
      if (!std::isfinite(v)) {
        margin = 0.0;
        return false;
      }
      scale = std::max(scale, std::fabs(v));
    }
  margin = psd_lower_bound(A, n);
  return true;
}

// _cold_start_psd_shift! (cold_start.jl:288-312): ok, applied shift, margin, scale
bool psd_shift(double* A, int n, double& applied, double& margin, double& scale) {
  applied = 0.0;
  if (n == 0) {
    margin = 0.0;
    scale = 1.0;
    return true;
  }
  if (!psd_stats(A, n, scale, margin)) return false;
  double sf = safety(scale);
  const double slack = rounding_slack(scale);
  if (margin > sf) return true;
  for (int it = 0; it < 16; ++it) {
    const double shift = sf - margin + slack;
    if (!std::isfinite(shift)) return false;
    for (int i = 0; i < n; ++i) A[i + (long long)i * n] += shift;
    applied += shift;
    if (!psd_stats(A, n, scale, margin)) return false;
    sf = safety(scale);
    if (margin > sf) return true;
  }
  return false;
}

// _cold_start_identity_mass_shifts (cold_start.jl:314-350)
bool mass_shifts(double primal, double dual, long long rho, double& dp, double& dd) {
  dp = dd = 0.0;
  if (!(rho > 0 && std::isfinite(primal) && std::isfinite(dual))) return false;
  const double target = (double)rho;
  const double p = std::max(0.0, (target - primal) / target);
  const double d = std::max(0.0, (target - dual) / target);
  if (!(std::isfinite(p) && std::isfinite(d))) return false;
  dp = p;
  dd = d;
  return true;
}

// _cold_start_centering_shifts (cold_start.jl:409-428)
bool centering_shifts(double kappa, double primal, double dual, double& ps, double& ds) {
  bool ok = std::isfinite(kappa) && kappa > 0.0 && std::isfinite(primal) && primal > 0.0 && std::isfinite(dual) &&
            dual > 0.0;
  ps = ok ? kappa / (2.0 * dual) : 0.0;
  ds = ok ? kappa / (2.0 * primal) : 0.0;
  ok = ok && std::isfinite(ps) && std::isfinite(ds);
  if (!ok) ps = ds = 0.0;
  return ok;
}

double kdot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

double trace(const std::vector<double>& a, int n) {
  double t = 0.0;
  for (int i = 0; i < n; ++i) t += a[i + (long long)i * n];
  return t;
}

}  // namespace cold

}  // namespace

extern "C" {

const char* sdpx_last_error(void) { return sdpx::last_error(); }

int sdpx_version(void) { return 100; }

int sdpx_device_count(int* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  if (count) *count = n;
  return kOk;
}

int sdpx_dense_destroy(sdpx_dense_t* h) {
  if (!h) return kOk;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto& e : h->ev)
    if (e) hipEventDestroy(e);
  for (void* p : h->allocs) hipFree(p);
  if (h->hsc) hipHostFree(h->hsc);
  if (h->hinfo) hipHostFree(h->hinfo);
  if (h->hts) hipHostFree(h->hts);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return kOk;
}

int sdpx_dense_create(int device, int L, const int* k, int m, sdpx_dense_t** out) {
  if (!out || L <= 0 || m <= 0 || !k) {
    set_last_error("sdpx_dense_create: invalid dimensions");
    return kInvalidArgument;
  }
  *out = nullptr;
  for (int l = 0; l < L; ++l)
    if (k[l] <= 0) {
      set_last_error("sdpx_dense_create: block sizes must be positive");
      return kInvalidArgument;
    }
  SDPX_HIP(hipSetDevice(device));
  auto* h = new sdpx_dense();
  h->device = device;
  h->m = m;
  h->L = L;
  h->blk.resize(L);
  auto fail = [&](int s) {
    sdpx_dense_destroy(h);
    return s;
  };
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    set_last_error("hipStreamCreate failed");
    return fail(kHipError);
  }
  long long maxk2 = 0;
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    b.k = k[l];
    b.k2 = (long long)b.k * b.k;
    b.nblk = ceil_div(b.k, kCholNb);
    maxk2 = std::max(maxk2, b.k2);
    h->maxk2 = maxk2;
    const long long dn = (long long)b.nblk * kCholNb * kCholNb;
    int s = kOk;
    if ((s = dalloc(h, &b.Av, (size_t)(b.k2 * m)))) return fail(s);
    double** mats[] = {&b.C, &b.X, &b.Y, &b.LXinv, &b.Xinv, &b.P, &b.R, &b.dX, &b.dY, &b.W1, &b.W2};
    for (double** p : mats)
      if ((s = dalloc(h, p, (size_t)b.k2))) return fail(s);
    // L_X and M_Y (and their diagonal-block inverses) side by side: both
    // factors run as one batch of two
    if ((s = dalloc(h, &b.LX, (size_t)(2 * b.k2)))) return fail(s);
    b.MY = b.LX + b.k2;
    if ((s = dalloc(h, &b.dinvX, (size_t)(2 * dn)))) return fail(s);
    b.dinvY = b.dinvX + dn;
    if (b.k > 2) {
      if ((s = dalloc(h, &b.trial, (size_t)(b.k2 * kTrialCap)))) return fail(s);
      if ((s = dalloc(h, &b.trial_dinv, (size_t)(dn * kTrialCap)))) return fail(s);
    }
  }
  int s = kOk;
  double** vecs[] = {&h->c, &h->x, &h->dx, &h->dx_best, &h->d, &h->v, &h->r, &h->rho, &h->delta, &h->v2};
  for (double** p : vecs)
    if ((s = dalloc(h, p, (size_t)m))) return fail(s);
  const long long mm2 = (long long)m * m;
  if ((s = dalloc(h, &h->S, (size_t)mm2))) return fail(s);
  if ((s = dalloc(h, &h->Sf, (size_t)mm2))) return fail(s);
  if ((s = dalloc(h, &h->dinvS, (size_t)ceil_div(m, kCholNb) * kCholNb * kCholNb))) return fail(s);
  if ((s = dalloc(h, &h->trsv_y, (size_t)m))) return fail(s);
  if ((s = dalloc(h, &h->T, (size_t)(maxk2 * m)))) return fail(s);
  if ((s = dalloc(h, &h->Ph, (size_t)(maxk2 * m)))) return fail(s);
  // split-K for the Schur Gram product.  All lower tiles cost the same, so the
  // launch runs ceil(tiles*split / slots) rounds of resident work-groups
  // (two 128 x 128 groups per CU); pick the split that fills the rounds best,
  // keeping >= 64 K-slices per group and the partial sums within 4 GiB.
  {
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
    const long long slots = 2LL * cus;
    const long long nt = ceil_div(m, 128);
    const long long tiles = nt * (nt + 1) / 2;
    const long long stages = ceil_div(maxk2, 16);
    const long long mem_cap = (4LL << 30) / (8 * mm2);
    int best = 1;
    double best_cost = 1e300;
    for (int sp = 1; sp <= 32; ++sp) {
      if (sp > 1 && (stages / sp < 64 || sp > mem_cap)) break;
      const long long groups = tiles * sp;
      const long long rounds = (groups + slots - 1) / slots;
      // cost in units of one group's K-slice (~6 us): per-group slices plus a
      // prologue/epilogue allowance, and the split-K reduction pass that reads
      // sp partial m x m matrices (~8 bytes * sp * m^2 at ~4 TB/s)
      const double cost = (double)rounds * ((double)stages / sp + 24.0) + (sp > 1 ? 3.1e-7 * sp * mm2 : 0.0);
      if (cost < best_cost * 0.995) {
        best_cost = cost;
        best = sp;
      }
    }
    h->split = best;
    h->split_created = best;
    // room for the autotuner's larger candidates (up to 2x), within the same cap
    h->split_max = best > 1 ? (int)std::min<long long>(2LL * best, std::min<long long>(mem_cap, std::max(1LL, stages / 64)))
                            : 1;
    if (h->split_max < best) h->split_max = best;
    if (h->split_max > 1 && (s = dalloc(h, &h->splitws, (size_t)(h->split_max * mm2)))) return fail(s);
    // Gram tiles visited in 8 x 8 super-blocks: the co-resident groups of one
    // XCD then read few distinct Phat column panels, shared in its L2.
    if ((s = dalloc(h, &h->tile_order, (size_t)tiles))) return fail(s);
    if ((s = dalloc(h, &h->gram_items, (size_t)(tiles * h->split_max)))) return fail(s);
    if ((s = set_gram_order(h, 8, h->split))) return fail(s);
  }
  h->gemv_part_cap = std::max<long long>(maxk2 * 64, 4096);
  if ((s = dalloc(h, &h->gemv_part, (size_t)h->gemv_part_cap))) return fail(s);
  if ((s = dalloc(h, &h->symv_part, (size_t)ceil_div(m, 64) * m))) return fail(s);
  if ((s = dalloc(h, &h->red_part, 1024))) return fail(s);
  if ((s = dalloc(h, &h->dsc, kScalars))) return fail(s);
  if ((s = dalloc(h, &h->dinfo, 1024))) return fail(s);
  if ((s = dalloc(h, &h->dts, kTrialCap))) return fail(s);
  if ((s = dalloc(h, &h->dok, kTrialCap))) return fail(s);
  if (hipHostMalloc(reinterpret_cast<void**>(&h->hsc), sizeof(double) * kScalars) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&h->hinfo), sizeof(int) * 1024) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&h->hts), sizeof(double) * kTrialCap) != hipSuccess) {
    set_last_error("hipHostMalloc failed");
    return fail(kOutOfMemory);
  }
  SDPX_HIP(hipMemsetAsync(h->S, 0, sizeof(double) * mm2, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  *out = h;
  return kOk;
}

int sdpx_dense_memory(sdpx_dense_t* h, long long* bytes) {
  SDPX_CHECK(check_handle(h));
  if (bytes) *bytes = h->bytes;
  return kOk;
}

int sdpx_dense_set_block(sdpx_dense_t* h, int l, const double* Av, const double* C) {
  SDPX_CHECK(check_handle(h));
  if (l < 0 || l >= h->L || !Av || !C) {
    set_last_error("sdpx_dense_set_block: bad block index or null pointer");
    return kInvalidArgument;
  }
  Block& b = h->blk[l];
  if (b.na != 0) {
    set_last_error("sdpx_dense_set_block: the block was given as sparse coordinates");
    return kInvalidArgument;
  }
  SDPX_HIP(hipMemcpyAsync(b.Av, Av, sizeof(double) * b.k2 * h->m, hipMemcpyHostToDevice, h->stream));
  h->av_ver += 1;
  SDPX_HIP(hipMemcpyAsync(b.C, C, sizeof(double) * b.k2, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_set_objective(sdpx_dense_t* h, const double* c) {
  SDPX_CHECK(check_handle(h));
  SDPX_HIP(hipMemcpyAsync(h->c, c, sizeof(double) * h->m, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_set_iterate(sdpx_dense_t* h, const double* x, const double* X, const double* Y) {
  SDPX_CHECK(check_handle(h));
  if (x) SDPX_HIP(hipMemcpyAsync(h->x, x, sizeof(double) * h->m, hipMemcpyHostToDevice, h->stream));
  long long off = 0;
  for (auto& b : h->blk) {
    if (X) SDPX_HIP(hipMemcpyAsync(b.X, X + off, sizeof(double) * b.k2, hipMemcpyHostToDevice, h->stream));
    if (Y) SDPX_HIP(hipMemcpyAsync(b.Y, Y + off, sizeof(double) * b.k2, hipMemcpyHostToDevice, h->stream));
    off += b.k2;
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_get_iterate(sdpx_dense_t* h, double* x, double* X, double* Y) {
  SDPX_CHECK(check_handle(h));
  if (x) SDPX_HIP(hipMemcpyAsync(x, h->x, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
  long long off = 0;
  for (auto& b : h->blk) {
    if (X) SDPX_HIP(hipMemcpyAsync(X + off, b.X, sizeof(double) * b.k2, hipMemcpyDeviceToHost, h->stream));
    if (Y) SDPX_HIP(hipMemcpyAsync(Y + off, b.Y, sizeof(double) * b.k2, hipMemcpyDeviceToHost, h->stream));
    off += b.k2;
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_get_direction(sdpx_dense_t* h, double* dx, double* dX, double* dY) {
  SDPX_CHECK(check_handle(h));
  if (dx) SDPX_HIP(hipMemcpyAsync(dx, h->dx, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
  long long off = 0;
  for (auto& b : h->blk) {
    if (dX) SDPX_HIP(hipMemcpyAsync(dX + off, b.dX, sizeof(double) * b.k2, hipMemcpyDeviceToHost, h->stream));
    if (dY) SDPX_HIP(hipMemcpyAsync(dY + off, b.dY, sizeof(double) * b.k2, hipMemcpyDeviceToHost, h->stream));
    off += b.k2;
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_get_schur(sdpx_dense_t* h, double* S) {
  SDPX_CHECK(check_handle(h));
  SDPX_HIP(hipMemcpyAsync(S, h->S, sizeof(double) * h->m * h->m, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_set_schur(sdpx_dense_t* h, const double* S) {
  SDPX_CHECK(check_handle(h));
  SDPX_HIP(hipMemcpyAsync(h->S, S, sizeof(double) * h->m * h->m, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_set_rhs(sdpx_dense_t* h, const double* r) {
  SDPX_CHECK(check_handle(h));
  SDPX_HIP(hipMemcpyAsync(h->r, r, sizeof(double) * h->m, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_get_rhs(sdpx_dense_t* h, double* r) {
  SDPX_CHECK(check_handle(h));
  SDPX_HIP(hipMemcpyAsync(r, h->r, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_set_dx(sdpx_dense_t* h, const double* dx) {
  SDPX_CHECK(check_handle(h));
  SDPX_HIP(hipMemcpyAsync(h->dx, dx, sizeof(double) * h->m, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_synchronize(sdpx_dense_t* h) {
  SDPX_CHECK(check_handle(h));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

// ---------------------------------------------------------------------------
int sdpx_dense_residuals(sdpx_dense_t* h, const double* mu, int affine, int factor, double* out) {
  SDPX_CHECK(check_handle(h));
  const int L = h->L;
  // P_l = Av_l x - X_l - C_l ; slot l = ||P_l||_inf
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(avx_f64(b.Av, b.k2, h->m, h->x, 1.0, b.P, -1.0, b.X, -1.0, b.C, h->gemv_part, h->gemv_part_cap,
                       h->stream));
    SDPX_CHECK(absmax_f64(b.P, b.k2, h->red_part, h->dsc + l, h->stream));
  }
  // d = c - sum_l Av_l' vec(Y_l)
  SDPX_HIP(hipMemcpyAsync(h->d, h->c, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream));
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(avt_f64(b.Av, b.k2, h->m, b.Y, -1.0, 1.0, h->d, h->stream));
  }
  SDPX_CHECK(absmax_f64(h->d, h->m, h->red_part, h->dsc + L, h->stream));
  // R_l = -X_l Y_l (+ mu_l I)
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(mm(h, b.k, -1.0, b.X, false, b.Y, false, 0.0, b.R));
    if (!affine) SDPX_CHECK(add_diag_f64(b.R, b.k, b.k, mu[l], nullptr, 0.0, h->stream));
  }
  int nint = 0;
  if (factor) {
    for (int l = 0; l < L; ++l) {
      Block& b = h->blk[l];
      SDPX_HIP(hipMemcpyAsync(b.LX, b.X, sizeof(double) * b.k2, hipMemcpyDeviceToDevice, h->stream));
      SDPX_HIP(hipMemcpyAsync(b.MY, b.Y, sizeof(double) * b.k2, hipMemcpyDeviceToDevice, h->stream));
      // L_X and M_Y in one batch of two (MY = LX + k2, dinvY = dinvX + dn; info 2l, 2l+1)
      SDPX_CHECK(potrf_f64(b.LX, b.k, b.k2, b.k, 2, b.dinvX, b.dinvY - b.dinvX, h->dinfo + 2 * l, h->stream));
    }
    nint = 2 * L;
  }
  SDPX_CHECK(sync_scalars(h, L + 1, nint));
  double p_res = 0.0;
  for (int l = 0; l < L; ++l) p_res = std::max(p_res, h->hsc[l]);
  for (int l = 0; l < L; ++l)
    if (std::isnan(h->hsc[l])) p_res = h->hsc[l];
  const double d_res = h->hsc[L];
  bool ok = true;
  for (int i = 0; i < nint; ++i) ok = ok && h->hinfo[i] == 0;
  if (factor && ok) {
    // L_X^-1 and X^-1 = L_X^-T L_X^-1 for the Schur panels and X^-1 M products;
    // M_Y is used as a full matrix operand, so its strict upper triangle is cleared
    for (int l = 0; l < L; ++l) {
      Block& b = h->blk[l];
      SDPX_CHECK(zero_upper_f64(b.MY, b.k, b.k, h->stream));
      SDPX_CHECK(trtri_lower_f64(b.LX, b.k, b.k, b.dinvX, b.LXinv, nullptr, h->stream));
      SDPX_CHECK(gemm_f64(gemm(b.k, b.k, b.k, 1.0, b.LXinv, b.k, b.LXinv, b.k, 0.0, b.Xinv, b.k,
                               kGemmTransA | kGemmAUpper | kGemmBLower),
                          h->stream));
    }
  }
  if (out) {
    out[0] = p_res;
    out[1] = d_res;
    out[2] = ok ? 1.0 : 0.0;
  }
  return ok ? kOk : kNotPositiveDefinite;
}

int sdpx_dense_schur_build(sdpx_dense_t* h) {
  SDPX_CHECK(check_handle(h));
  const int m = h->m;
  bool first_dense = true, any_sparse = false;
  for (int l = 0; l < h->L; ++l) {
    Block& b = h->blk[l];
    const int k = b.k;
    if (b.na > 0) {
      any_sparse = true;
      continue;
    }
    const double beta0 = first_dense ? 0.0 : 1.0;
    first_dense = false;
    if (h->prof) SDPX_HIP(hipEventRecord(h->ev[0], h->stream));
    if (h->precision == 2 && h->panel_precision == 1 && b.Av32) {
      // BF16x3 assembly from the FP32 P-hat chain: the bf16 planes straight from FP32 MFMA products
      if (b.av32_ver != h->av_ver) {
        SDPX_CHECK(convert_panel_f32(b.Av, b.k2, b.Av32, b.k2, b.k2, m, h->stream));
        b.av32_ver = h->av_ver;
      }
      SDPX_CHECK(convert_panel_f32(b.MY, k, h->MY32, k, k, k, h->stream));
      SDPX_CHECK(convert_panel_f32(b.LXinv, k, h->LX32, k, k, k, h->stream));
      SDPX_CHECK(panel_chain_bf16(b.Av32, h->MY32, h->LX32, k, m, reinterpret_cast<float*>(h->T), h->Ph_hi,
                                  h->Ph_lo, h->ldPb, h->stream));
      const long long kp = (b.k2 + kBf16Align - 1) / kBf16Align * kBf16Align;
      if (h->prof) SDPX_HIP(hipEventRecord(h->ev[1], h->stream));
      SDPX_CHECK(gram_bf16x3(h->Ph_hi, h->Ph_lo, h->ldPb, m, (int)kp, h->S, m, beta0, h->split, h->splitws,
                             h->tile_order, h->stream, h->prof ? h->ev[2] : nullptr));
      if (h->prof) {
        SDPX_HIP(hipEventRecord(h->ev[3], h->stream));
        SDPX_HIP(hipEventSynchronize(h->ev[3]));
        float a = 0.f, g = 0.f, r = 0.f;
        SDPX_HIP(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
        SDPX_HIP(hipEventElapsedTime(&g, h->ev[1], h->ev[2]));
        SDPX_HIP(hipEventElapsedTime(&r, h->ev[2], h->ev[3]));
        h->t_panels += a;
        h->t_gram += g;
        h->t_reduce += r;
        h->n_gram += 1;
      }
      continue;
    }
    // T_i = A_i M_Y for every i as one product over the stacked panel
    // [A_1; ...; A_m] = Av' (each A_i is symmetric), M = km rows, op(B) lower;
    // rows (i, r) are written back to the k^2 x m panel layout (block remap)
    GemmArgs w = gemm(k * m, k, k, 1.0, b.Av, k, b.MY, k, 0.0, h->T, k, kGemmTransA | kGemmBLower | h->panel_tiles[0]);
    w.block_rows = k;
    w.block_stride = b.k2;
    SDPX_CHECK(gemm_f64(w, h->stream));
    // Phat = L_X^-1 [T_1 ... T_m]   (k x km, op(A) lower; the partial row tile
    // comes first, where the triangle leaves it the shortest K range)
    SDPX_CHECK(gemm_f64(gemm(k, k * m, k, 1.0, b.LXinv, k, h->T, k, 0.0, h->Ph, k,
                             kGemmALower | kGemmRaggedFirst | h->panel_tiles[1]),
                        h->stream));
    // S (+)= Phat' Phat, lower triangle only (schur.jl:469-475 syrk 'L','T')
    GemmArgs s = gemm(m, m, (int)b.k2, 1.0, h->Ph, b.k2, h->Ph, b.k2, beta0, h->S, m,
                      kGemmTransA | kGemmLowerOut | kGemmSchurGram);
    s.tile_order = h->tile_order;
    s.item_order = h->gram_items;
    if (h->precision == 2) {
      // mixed precision: Phat split into bf16 hi + lo planes, Gram product as three
      // bf16 MFMA products accumulated in FP32, S kept in FP64
      const long long kp = (b.k2 + kBf16Align - 1) / kBf16Align * kBf16Align;
      SDPX_CHECK(split_panel_bf16(h->Ph, b.k2, h->Ph_hi, h->Ph_lo, h->ldPb, b.k2, m, h->stream));
      if (h->prof) SDPX_HIP(hipEventRecord(h->ev[1], h->stream));
      SDPX_CHECK(gram_bf16x3(h->Ph_hi, h->Ph_lo, h->ldPb, m, (int)kp, h->S, m, beta0, h->split,
                             h->splitws, h->tile_order, h->stream, h->prof ? h->ev[2] : nullptr));
    } else if (h->precision == 1) {
      // mixed precision: Phat rounded to FP32, Gram product on FP32 MFMA, S kept in FP64
      SDPX_CHECK(convert_panel_f32(h->Ph, b.k2, h->Pf, h->ldPf, b.k2, m, h->stream));
      if (h->prof) SDPX_HIP(hipEventRecord(h->ev[1], h->stream));
      SDPX_CHECK(gram_f32(h->Pf, h->ldPf, m, (int)b.k2, h->S, m, beta0, h->split, h->splitws,
                          h->tile_order, h->stream, h->prof ? h->ev[2] : nullptr));
    } else {
      if (h->prof) SDPX_HIP(hipEventRecord(h->ev[1], h->stream));
      SDPX_CHECK(gemm_f64(s, h->stream, h->split, h->splitws, h->prof ? h->ev[2] : nullptr));
    }
    if (h->prof) {
      SDPX_HIP(hipEventRecord(h->ev[3], h->stream));
      SDPX_HIP(hipEventSynchronize(h->ev[3]));
      float a = 0.f, g = 0.f, r = 0.f;
      SDPX_HIP(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
      SDPX_HIP(hipEventElapsedTime(&g, h->ev[1], h->ev[2]));
      SDPX_HIP(hipEventElapsedTime(&r, h->ev[2], h->ev[3]));
      h->t_panels += a;
      h->t_gram += g;
      h->t_reduce += r;
      h->n_gram += 1;
    }
  }
  if (any_sparse) {
    // sparse blocks (sparse_schur_block!, schur.jl:936-1032, then the scatter of
    // reduce_sparse_schur!, :1229-1318): pair sums accumulated in the upper triangle of the
    // factor buffer (free until the factor copies S), then added to the lower triangle of S
    SDPX_HIP(hipMemsetAsync(h->Sf, 0, sizeof(double) * (size_t)m * m, h->stream));
    for (int l = 0; l < h->L; ++l) {
      Block& b = h->blk[l];
      if (b.na == 0) continue;
      hipLaunchKernelGGL(sparse_w1_kernel, dim3((unsigned)b.k2, (unsigned)ceil_div(b.na, 256)), dim3(256), 0,
                         h->stream, b.k, b.na, b.sptr, b.srow, b.scol, b.sval, b.Y, b.Xinv, b.W1T);
      SDPX_HIP(hipGetLastError());
      hipLaunchKernelGGL(sparse_pair_kernel, dim3((unsigned)b.na), dim3(256), 0, h->stream, b.na, b.ids, b.sptr,
                         b.slin, b.sval, b.W1T, h->Sf, (long long)m);
      SDPX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(add_upper_to_lower_kernel, dim3((unsigned)ceil_div((long long)m * m, 256)), dim3(256), 0,
                       h->stream, h->S, h->Sf, m, first_dense ? 0 : 1);
    SDPX_HIP(hipGetLastError());
  }
  return kOk;
}

int sdpx_dense_schur_factor(sdpx_dense_t* h, int* reg_attempts) {
  SDPX_CHECK(check_handle(h));
  const int m = h->m;
  int attempts = 0;
  double reg = 0.0;
  bool ok = false;
  if (h->prof) SDPX_HIP(hipEventRecord(h->ev[4], h->stream));
  if (h->arrow) {
    SDPX_CHECK(arrow_factor(h, &attempts, &ok));
  } else {
    while (true) {
      if (h->factor_precision == 1) {
        // FP32 factor (panels and trailing updates on FP32 MFMA), returned in FP64 storage
        SDPX_CHECK(lower_to_f32(h->S, m, attempts == 0 ? 0.0 : reg, h->S32, h->ldS32, h->stream));
        SDPX_CHECK(potrf_f32(h->S32, h->ldS32, m, h->blk64, h->dinvS, h->dinv32, h->dinfo, h->stream));
        SDPX_CHECK(lower_to_f64(h->S32, h->ldS32, m, h->Sf, h->stream));
      } else {
        SDPX_CHECK(copy_lower_shift_f64(h->S, h->Sf, m, attempts == 0 ? 0.0 : reg, h->stream));
        SDPX_CHECK(potrf_f64(h->Sf, m, 0, m, 1, h->dinvS, 0, h->dinfo, h->stream));
      }
      SDPX_CHECK(sync_scalars(h, 0, 1));
      ok = h->hinfo[0] == 0;
      if (ok || attempts >= 6) break;
      attempts += 1;
      reg = attempts == 1 ? std::sqrt(kEps) : reg * 10.0;
    }
  }
  h->last_reg_attempts = attempts;
  if (reg_attempts) *reg_attempts = attempts;
  if (h->prof) {
    SDPX_HIP(hipEventRecord(h->ev[5], h->stream));
    SDPX_HIP(hipEventSynchronize(h->ev[5]));
    float f = 0.f;
    SDPX_HIP(hipEventElapsedTime(&f, h->ev[4], h->ev[5]));
    h->t_factor += f;
    h->n_factor += 1;
  }
  return ok ? kOk : kNotPositiveDefinite;
}

int sdpx_dense_set_profiling(sdpx_dense_t* h, int enabled) {
  SDPX_CHECK(check_handle(h));
  if (enabled && !h->ev[0]) {
    for (auto& e : h->ev) SDPX_HIP(hipEventCreate(&e));
  }
  h->prof = enabled != 0;
  return kOk;
}

int sdpx_dense_kernel_times(sdpx_dense_t* h, int reset, double* out) {
  SDPX_CHECK(check_handle(h));
  if (out) {
    out[0] = h->t_gram;
    out[1] = (double)h->n_gram;
    out[2] = h->t_panels;
    out[3] = h->t_reduce;
    out[4] = h->t_factor;
    out[5] = (double)h->n_factor;
  }
  if (reset) {
    h->t_gram = h->t_panels = h->t_reduce = h->t_factor = 0.0;
    h->n_gram = h->n_factor = 0;
  }
  return kOk;
}

int sdpx_gram_work_list(int m, int group, int split, int* out) {
  if (m < 1 || group < 1 || split < 1 || !out) {
    set_last_error("sdpx_gram_work_list: m, group and split must be positive, out non-null");
    return kInvalidArgument;
  }
  const long long nt = ceil_div(m, 128);
  if (nt > 32767 || nt * (nt + 1) / 2 * split > (1LL << 30)) {
    set_last_error("sdpx_gram_work_list: too many work items");
    return kInvalidArgument;
  }
  gram_item_order(m, group, split, reinterpret_cast<int2*>(out));
  return kOk;
}

int sdpx_dense_schur_split(sdpx_dense_t* h, int* split) {
  SDPX_CHECK(check_handle(h));
  if (split) *split = h->split;
  return kOk;
}

// Plan-time tuning of the Schur build for this handle's shape: every candidate
// is run on the device and timed with events, on the handle's A_i with dense
// stand-ins for L_X^-1 and M_Y (restored afterwards), so it can run before any
// iterate is set; it overwrites the panel and Schur buffers.  Candidates: each of the two P-hat panel
// products on 128 x 128, 64 x 64, 128 x 64 or 64 x 128 tiles; Gram split-K
// factor s/2, 3s/4, s, 3s/2 or 2s of the create-time choice s (the partial-sum
// workspace holds up to 2s, within the create-time memory cap); Gram tile
// order in 4 x 4, 8 x 8 or 16 x 16 super-blocks.
// Also the number of bisection levels per batched trial round of the fraction
// to the boundary (4..7; any value gives the same result).  out (optional, 20
// doubles): tile shape code of each panel product (0..3 in the order above),
// split, super-block side, panel ms, Gram ms, create-time split, the 2 x 4
// panel timings, the levels per round, and the modelled search ms for 4..7.
int sdpx_dense_autotune(sdpx_dense_t* h, double* out) {
  SDPX_CHECK(check_handle(h));
  const Block* bp = &h->blk[0];
  for (int l = 1; l < h->L; ++l)
    if (h->blk[l].k > bp->k) bp = &h->blk[l];
  const Block& b = *bp;
  const int m = h->m, k = b.k;
  hipEvent_t ev[2] = {};
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      for (auto& f : ev)
        if (f) (void)hipEventDestroy(f);
      set_last_error("sdpx_dense_autotune: hipEventCreate failed");
      return kHipError;
    }
  auto time_ms = [&](auto&& launch, float* ms) -> int {
    SDPX_CHECK(launch());  // warm
    SDPX_HIP(hipEventRecord(ev[0], h->stream));
    SDPX_CHECK(launch());
    SDPX_HIP(hipEventRecord(ev[1], h->stream));
    SDPX_HIP(hipEventSynchronize(ev[1]));
    SDPX_HIP(hipEventElapsedTime(ms, ev[0], ev[1]));
    return kOk;
  };
  // L_X^-1 and M_Y are replaced by dense stand-ins while timing, and restored
  double* saved = nullptr;
  const size_t kk = (size_t)k * k;
  if (hipMalloc(&saved, 2 * kk * sizeof(double)) != hipSuccess) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    set_last_error("sdpx_dense_autotune: hipMalloc failed");
    return kOutOfMemory;
  }
  const bool saved_ok =
      hipMemcpyAsync(saved, b.LXinv, kk * sizeof(double), hipMemcpyDeviceToDevice, h->stream) == hipSuccess &&
      hipMemcpyAsync(saved + kk, b.MY, kk * sizeof(double), hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
  auto body = [&]() -> int {
    if (!saved_ok) {
      set_last_error("sdpx_dense_autotune: saving L_X^-1 and M_Y failed");
      return kHipError;
    }
    hipLaunchKernelGGL(tune_lower_kernel, dim3(ceil_div((long long)kk, 256)), dim3(256), 0, h->stream, k, b.LXinv);
    hipLaunchKernelGGL(tune_lower_kernel, dim3(ceil_div((long long)kk, 256)), dim3(256), 0, h->stream, k, b.MY);
    SDPX_HIP(hipGetLastError());
    h->tune_panel_ms = 0.f;
    // the two panel products, each on four tile shapes
    const unsigned shapes[4] = {0u, (unsigned)kGemmSmallTiles, (unsigned)kGemmTallTiles, (unsigned)kGemmWideTiles};
    for (int prod = 0; prod < 2; ++prod) {
      float best = 0.f;
      // the create-time default (64 x 64) is the first incumbent
      for (int c : {1, 0, 2, 3}) {
        const unsigned tiles = shapes[c];
        float ms = 0.f;
        SDPX_CHECK(time_ms(
            [&]() -> int {
              if (prod == 0) {
                GemmArgs w = gemm(k * m, k, k, 1.0, b.Av, k, b.MY, k, 0.0, h->T, k,
                                  kGemmTransA | kGemmBLower | tiles);
                w.block_rows = k;
                w.block_stride = b.k2;
                return gemm_f64(w, h->stream);
              }
              return gemm_f64(gemm(k, k * m, k, 1.0, b.LXinv, k, h->T, k, 0.0, h->Ph, k,
                                   kGemmALower | kGemmRaggedFirst | tiles),
                              h->stream);
            },
            &ms));
        h->tune_ms[prod][c] = ms;
        if (c == 1 || ms < 0.99f * best) {  // 1% hysteresis, as for the Gram below
          best = ms;
          h->panel_tiles[prod] = tiles;
        }
      }
      h->tune_panel_ms += best;
    }
    // Gram product in the handle's precision: split-K factor and tile order
    const long long kp = (b.k2 + kBf16Align - 1) / kBf16Align * kBf16Align;
    if (h->precision == 2)
      SDPX_CHECK(split_panel_bf16(h->Ph, b.k2, h->Ph_hi, h->Ph_lo, h->ldPb, b.k2, m, h->stream));
    else if (h->precision == 1)
      SDPX_CHECK(convert_panel_f32(h->Ph, b.k2, h->Pf, h->ldPf, b.k2, m, h->stream));
    auto gram = [&](int sp) -> int {
      if (h->precision == 2)
        return gram_bf16x3(h->Ph_hi, h->Ph_lo, h->ldPb, m, (int)kp, h->S, m, 0.0, sp, h->splitws, h->tile_order,
                           h->stream, nullptr);
      if (h->precision == 1)
        return gram_f32(h->Pf, h->ldPf, m, (int)b.k2, h->S, m, 0.0, sp, h->splitws, h->tile_order, h->stream,
                        nullptr);
      GemmArgs g = gemm(m, m, (int)b.k2, 1.0, h->Ph, b.k2, h->Ph, b.k2, 0.0, h->S, m,
                        kGemmTransA | kGemmLowerOut | kGemmSchurGram);
      g.tile_order = h->tile_order;
      g.item_order = h->gram_items;
      return gemm_f64(g, h->stream, sp, h->splitws);
    };
    const int s0 = h->split_created;
    int best_split = s0, best_group = 8;
    bool first = true;
    // the create-time choice (split s0, 8 x 8 super-blocks) is timed first and is
    // replaced only by a candidate at least 1% faster (3% for another tile
    // order): run-to-run noise between near-equal candidates does not flip the
    // choice.  The kernel is MFMA-bound, so the orders time alike; the 64
    // co-resident groups of an XCD span one 8 x 8 super-block (16 panels of
    // Phat) or up to four 4 x 4 ones (32), and fewer panels mean fewer L2 misses.
    for (int group : {8, 4, 16}) {
      int last = 0;
      for (int sp : {s0, (3 * s0) / 4, s0 / 2, (3 * s0) / 2, 2 * s0}) {
        if (sp > h->split_max) continue;
        if (sp < 1 || sp == last) continue;
        last = sp;
        SDPX_CHECK(set_gram_order(h, group, sp));
        float ms = 0.f;
        SDPX_CHECK(time_ms([&]() -> int { return gram(sp); }, &ms));
        if (first || ms < (group == 8 ? 0.99f : 0.97f) * h->tune_gram_ms) {
          first = false;
          h->tune_gram_ms = ms;
          best_split = sp;
          best_group = group;
        }
      }
    }
    h->split = best_split;
    h->tile_group = best_group;
    SDPX_CHECK(set_gram_order(h, best_group, best_split));
    // fraction-to-the-boundary rounds: wall time of one batched trial round
    // (trial fill, batched Cholesky, read-back, host wait) at the batch sizes
    // of D = 4..7 levels per round; a search of ~53 levels costs round 0
    // (2 * 2^(D-1) trials) plus ceil((53 - (D-1)) / D) rounds of 2 (2^D - 1)
    if (b.k > 2) {
      const long long dstride = (long long)b.nblk * kCholNb * kCholNb;
      auto round_s = [&](int B, double* sec) -> int {
        for (int rep = 0; rep < 3; ++rep) {  // first run warms
          const auto t0 = std::chrono::steady_clock::now();
          hipLaunchKernelGGL(identity_trials_kernel, dim3(std::min(8192, ceil_div(b.k2 * B, 256))), dim3(256), 0,
                             h->stream, b.k2, b.k, B, b.trial);
          SDPX_HIP(hipGetLastError());
          SDPX_CHECK(potrf_f64(b.trial, b.k, b.k2, b.k, B, b.trial_dinv, dstride, h->dok, h->stream));
          SDPX_HIP(hipMemcpyAsync(h->hinfo, h->dok, sizeof(int) * B, hipMemcpyDeviceToHost, h->stream));
          SDPX_HIP(hipStreamSynchronize(h->stream));
          if (rep > 0) *sec += 0.5 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return kOk;
      };
      double best = 0.0;
      for (int D = 4; D <= 7; ++D) {
        double r0 = 0.0, r1 = 0.0;
        SDPX_CHECK(round_s(2 << (D - 1), &r0));
        SDPX_CHECK(round_s(2 * ((1 << D) - 1), &r1));
        const double cost = 1e3 * (r0 + std::ceil((53.0 - (D - 1)) / D) * r1);
        h->tune_ftb_ms[D - 4] = (float)cost;
        if (D == 4 || cost < best) {
          best = cost;
          h->ftb_depth = D;
        }
      }
    }
    return kOk;
  };
  const int rc = body();
  // L_X^-1 and M_Y back (also after a failed timing)
  const bool restored = saved_ok &&
      hipMemcpyAsync(b.LXinv, saved, kk * sizeof(double), hipMemcpyDeviceToDevice, h->stream) == hipSuccess &&
      hipMemcpyAsync(b.MY, saved + kk, kk * sizeof(double), hipMemcpyDeviceToDevice, h->stream) == hipSuccess &&
      hipStreamSynchronize(h->stream) == hipSuccess;
  (void)hipFree(saved);
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc != kOk) return rc;
  if (!restored) {
    set_last_error("sdpx_dense_autotune: restoring L_X^-1 and M_Y failed");
    return kHipError;
  }
  if (out) {
    for (int prod = 0; prod < 2; ++prod) {
      const unsigned t = h->panel_tiles[prod];
      out[prod] = t == kGemmSmallTiles ? 1.0 : t == kGemmTallTiles ? 2.0 : t == kGemmWideTiles ? 3.0 : 0.0;
      for (int c = 0; c < 4; ++c) out[7 + 4 * prod + c] = h->tune_ms[prod][c];
    }
    out[2] = h->split;
    out[3] = h->tile_group;
    out[4] = h->tune_panel_ms;
    out[5] = h->tune_gram_ms;
    out[6] = h->split_created;
    out[15] = h->ftb_depth;
    for (int d = 0; d < 4; ++d) out[16 + d] = h->tune_ftb_ms[d];
  }
  return kOk;
}

// Applies tiling choices (as reported by sdpx_dense_autotune) without timing:
// panel shape codes 0..3, split in [1, workspace capacity], super-block side >= 1.
int sdpx_dense_set_tiling(sdpx_dense_t* h, int panel0, int panel1, int split, int group, int ftb_depth) {
  SDPX_CHECK(check_handle(h));
  const unsigned shapes[4] = {0u, (unsigned)kGemmSmallTiles, (unsigned)kGemmTallTiles, (unsigned)kGemmWideTiles};
  if (panel0 < 0 || panel0 > 3 || panel1 < 0 || panel1 > 3 || split < 1 || split > h->split_max || group < 1 ||
      ftb_depth < 1 || ftb_depth > 7) {
    set_last_error("sdpx_dense_set_tiling: shape codes 0..3, split 1..workspace capacity, group >= 1, levels 1..7");
    return kInvalidArgument;
  }
  h->panel_tiles[0] = shapes[panel0];
  h->panel_tiles[1] = shapes[panel1];
  h->split = split;
  h->tile_group = group;
  h->ftb_depth = ftb_depth;
  return set_gram_order(h, group, split);
}

int sdpx_dense_predictor_rhs(sdpx_dense_t* h) {
  SDPX_CHECK(check_handle(h));
  return rhs_from_blocks(h);
}

int sdpx_dense_corrector_rhs(sdpx_dense_t* h, int mehrotra, double sigma, double mu_global, int block_local,
                             const double* mu, double beta) {
  SDPX_CHECK(check_handle(h));
  for (int l = 0; l < h->L; ++l) {
    Block& b = h->blk[l];
    double* dotv = h->dsc + 2 * l;
    double* target = h->dsc + 2 * l + 1;
    int mode;
    if (!mehrotra) {
      mode = 0;
      SDPX_CHECK(dot_f64(b.X, b.dX, 1.0, b.Y, b.dY, 1.0, b.k2, h->red_part, dotv, h->stream));
    } else if (block_local) {
      mode = 1;
      SDPX_CHECK(dot_f64(b.X, nullptr, 0.0, b.Y, nullptr, 0.0, b.k2, h->red_part, dotv, h->stream));
    } else {
      mode = 2;
    }
    hipLaunchKernelGGL(corrector_target_kernel, dim3(1), dim3(64), 0, h->stream, dotv, mode,
                       mehrotra ? mu_global : mu[l], beta, sigma, b.k, target);
    SDPX_HIP(hipGetLastError());
    // R = -X Y - dX dY + target I
    SDPX_CHECK(mm(h, b.k, -1.0, b.X, false, b.Y, false, 0.0, b.R));
    SDPX_CHECK(mm(h, b.k, -1.0, b.dX, false, b.dY, false, 1.0, b.R));
    SDPX_CHECK(add_diag_f64(b.R, b.k, b.k, 0.0, target, 1.0, h->stream));
  }
  return rhs_from_blocks(h);
}

int sdpx_dense_solve(sdpx_dense_t* h) {
  SDPX_CHECK(check_handle(h));
  return solve_into(h, h->r, h->dx);
}

int sdpx_dense_direction_residual(sdpx_dense_t* h, double* out) {
  SDPX_CHECK(check_handle(h));
  SDPX_CHECK(direction_residual_dev(h, 0));
  SDPX_CHECK(sync_scalars(h, 2));
  if (out) {
    out[0] = h->hsc[0];
    out[1] = h->hsc[1];
  }
  return kOk;
}

int sdpx_dense_refine(sdpx_dense_t* h, double abstol, int max_steps, int fixed_policy, int* steps_out,
                      double* residual_out) {
  SDPX_CHECK(check_handle(h));
  int steps = 0;
  double residual = 0.0;
  if (max_steps <= 0) {
    if (steps_out) *steps_out = 0;
    if (residual_out) *residual_out = 0.0;
    return kOk;
  }
  SDPX_CHECK(direction_residual_dev(h, 0));
  SDPX_CHECK(sync_scalars(h, 1));
  residual = h->hsc[0];
  for (int it = 0; it < max_steps; ++it) {
    if (!fixed_policy && residual <= abstol) break;
    // snapshot, correct: dx += S^-1 rho, re-evaluate (kkt.jl:3930-3960)
    SDPX_HIP(hipMemcpyAsync(h->dx_best, h->dx, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream));
    SDPX_CHECK(solve_into(h, h->rho, h->delta));
    SDPX_CHECK(axpbyz_f64(h->m, 1.0, h->dx, 1.0, h->delta, h->dx, h->stream));
    SDPX_CHECK(direction_residual_dev(h, 0));
    SDPX_CHECK(sync_scalars(h, 1));
    const double corrected = h->hsc[0];
    if (!std::isfinite(corrected) || corrected > residual) {
      SDPX_HIP(hipMemcpyAsync(h->dx, h->dx_best, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream));
      // rho must describe the retained direction
      SDPX_CHECK(direction_residual_dev(h, 0));
      SDPX_HIP(hipStreamSynchronize(h->stream));
      break;
    }
    steps += 1;
    const bool small = corrected > residual * 0.5;
    residual = corrected;
    if (!fixed_policy && small) break;
  }
  if (steps_out) *steps_out = steps;
  if (residual_out) *residual_out = residual;
  return kOk;
}

int sdpx_dense_set_factor_precision(sdpx_dense_t* h, int precision) {
  SDPX_CHECK(check_handle(h));
  if (precision != 0 && precision != 1) {
    set_last_error("sdpx_dense_set_factor_precision: precision must be 0 (FP64) or 1 (FP32)");
    return kInvalidArgument;
  }
  if (precision == 1 && (h->arrow || h->neq > 0)) {
    set_last_error("sdpx_dense_set_factor_precision: the FP32 factor serves the plain dense route only "
                   "(no block-arrow elimination, no equality border)");
    return kInvalidArgument;
  }
  if (precision == 1 && !h->S32) {
    h->ldS32 = ((long long)h->m + 3) / 4 * 4;
    int s = dalloc(h, &h->S32, (size_t)(h->ldS32 * h->m));
    if (!s) s = dalloc(h, &h->blk64, (size_t)kCholNb * kCholNb);
    if (!s) s = dalloc(h, &h->dinv32, (size_t)kCholNb * kCholNb);
    if (s) return s;
  }
  h->factor_precision = precision;
  return kOk;
}

int sdpx_dense_set_panel_precision(sdpx_dense_t* h, int precision) {
  SDPX_CHECK(check_handle(h));
  if (precision != 0 && precision != 1) {
    set_last_error("sdpx_dense_set_panel_precision: precision must be 0 (FP64) or 1 (FP32)");
    return kInvalidArgument;
  }
  if (precision == 1) {
    // FP32 copies: M_Y and L_X^-1 (max k^2 floats each) and every dense block's Av (k^2 x m floats)
    if (!h->MY32) {
      int s = dalloc(h, &h->MY32, (size_t)h->maxk2);
      if (!s) s = dalloc(h, &h->LX32, (size_t)h->maxk2);
      if (s) return s;
    }
    for (int l = 0; l < h->L; ++l) {
      Block& b = h->blk[l];
      if (b.na > 0 || b.k % 4 != 0 || b.Av32) continue;
      int s = dalloc(h, &b.Av32, (size_t)(b.k2 * h->m));
      if (s) return s;
      b.av32_ver = -1;
    }
  }
  h->panel_precision = precision;
  return kOk;
}

int sdpx_dense_set_precision(sdpx_dense_t* h, int precision) {
  SDPX_CHECK(check_handle(h));
  if (precision < 0 || precision > 2) {
    set_last_error("sdpx_dense_set_precision: precision must be 0 (FP64), 1 (FP32 Schur Gram) or 2 (BF16x3 Schur Gram)");
    return kInvalidArgument;
  }
  if (precision != 0 && h->arrow) {
    set_last_error("sdpx_dense_set_precision: the block-arrow route factors in FP64 only (precision 0)");
    return kInvalidArgument;
  }
  for (int l = 0; l < h->L && precision != 0; ++l) {
    if (h->blk[l].na > 0) {
      set_last_error("sdpx_dense_set_precision: sparse blocks assemble in FP64 only (precision 0)");
      return kInvalidArgument;
    }
  }
  if (precision == 1 && !h->Pf) {
    h->ldPf = (h->maxk2 + 3) / 4 * 4;
    int s = dalloc(h, &h->Pf, (size_t)(h->ldPf * h->m));
    if (s) return s;
  }
  if (precision == 2 && !h->Ph_hi) {
    h->ldPb = (h->maxk2 + kBf16Align - 1) / kBf16Align * kBf16Align;
    int s = dalloc(h, &h->Ph_hi, (size_t)(h->ldPb * h->m));
    if (!s) s = dalloc(h, &h->Ph_lo, (size_t)(h->ldPb * h->m));
    if (s) return s;
  }
  h->precision = precision;
  return kOk;
}

// rho = r - S dx with S applied matrix-free in FP64 (S dx)_i = <A_i, X^-1 (sum_j dx_j A_j) Y>;
// dsc[slot] = ||rho||_inf, dsc[slot + 1] = ||r||_inf.  Uses the W1 / W2 blocks as scratch.
static int exact_residual_dev(sdpx_dense_t* h, int slot) {
  for (int l = 0; l < h->L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(avx_f64(b.Av, b.k2, h->m, h->dx, 1.0, b.W1, 0.0, nullptr, 0.0, nullptr, h->gemv_part, h->gemv_part_cap,
                       h->stream));
    SDPX_CHECK(mm(h, b.k, 1.0, b.Xinv, false, b.W1, false, 0.0, b.W2));
    SDPX_CHECK(mm(h, b.k, 1.0, b.W2, false, b.Y, false, 0.0, b.W1));
    SDPX_CHECK(avt_f64(b.Av, b.k2, h->m, b.W1, 1.0, l == 0 ? 0.0 : 1.0, h->v2, h->stream));
  }
  SDPX_CHECK(axpbyz_f64(h->m, 1.0, h->r, -1.0, h->v2, h->rho, h->stream));
  SDPX_CHECK(absmax_f64(h->rho, h->m, h->red_part, h->dsc + slot, h->stream));
  SDPX_CHECK(absmax_f64(h->r, h->m, h->red_part, h->dsc + slot + 1, h->stream));
  return kOk;
}

int sdpx_dense_exact_residual(sdpx_dense_t* h, double* out) {
  SDPX_CHECK(check_handle(h));
  SDPX_CHECK(exact_residual_dev(h, 0));
  SDPX_CHECK(sync_scalars(h, 2));
  if (out) {
    out[0] = h->hsc[0];
    out[1] = h->hsc[1];
  }
  return kOk;
}

int sdpx_dense_refine_exact(sdpx_dense_t* h, double abstol, int max_steps, int* steps_out, double* residual_out) {
  SDPX_CHECK(check_handle(h));
  int steps = 0;
  SDPX_CHECK(exact_residual_dev(h, 0));
  SDPX_CHECK(sync_scalars(h, 1));
  double residual = h->hsc[0];
  for (int it = 0; it < max_steps && !(residual <= abstol); ++it) {
    SDPX_HIP(hipMemcpyAsync(h->dx_best, h->dx, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream));
    SDPX_CHECK(solve_into(h, h->rho, h->delta));
    SDPX_CHECK(axpbyz_f64(h->m, 1.0, h->dx, 1.0, h->delta, h->dx, h->stream));
    SDPX_CHECK(exact_residual_dev(h, 0));
    SDPX_CHECK(sync_scalars(h, 1));
    const double corrected = h->hsc[0];
    if (!std::isfinite(corrected) || corrected > residual) {
      SDPX_HIP(hipMemcpyAsync(h->dx, h->dx_best, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream));
      SDPX_CHECK(exact_residual_dev(h, 0));
      SDPX_HIP(hipStreamSynchronize(h->stream));
      break;
    }
    const double contraction = corrected / residual;
    steps += 1;
    residual = corrected;
    if (!(residual <= abstol) && contraction > 0.0) {
      // the refinement budget of _predicted_mixed_refinement_steps (mixed_precision_kkt.jl:394-415)
      // with the measured contraction in place of cond * eps * 8: stop as soon as the target
      // cannot be reached within max_steps
      const double more = std::ceil(std::log(abstol / residual) / std::log(contraction));
      if (!(steps + more <= max_steps)) break;
    }
  }
  if (steps_out) *steps_out = steps;
  if (residual_out) *residual_out = residual;
  return kOk;
}

int sdpx_dense_direction_blocks(sdpx_dense_t* h) {
  SDPX_CHECK(check_handle(h));
  for (int l = 0; l < h->L; ++l) {
    Block& b = h->blk[l];
    // dX = Av dx + P
    SDPX_CHECK(avx_f64(b.Av, b.k2, h->m, h->dx, 1.0, b.dX, 1.0, b.P, 0.0, nullptr, h->gemv_part, h->gemv_part_cap,
                       h->stream));
    // dY = sym(X^-1 (R - dX Y))
    SDPX_CHECK(mm(h, b.k, 1.0, b.dX, false, b.Y, false, 0.0, b.W1));
    SDPX_CHECK(axpbyz_f64(b.k2, 1.0, b.R, -1.0, b.W1, b.W1, h->stream));
    SDPX_CHECK(mm(h, b.k, 1.0, b.Xinv, false, b.W1, false, 0.0, b.dY));
    SDPX_CHECK(symmetrize_f64(b.dY, b.k, h->stream));
  }
  return kOk;
}

int sdpx_dense_complementarity(sdpx_dense_t* h, double tX, double tY, double* out) {
  SDPX_CHECK(check_handle(h));
  const int L = h->L;
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(dot_f64(b.X, nullptr, 0.0, b.Y, nullptr, 0.0, b.k2, h->red_part, h->dsc + l, h->stream));
    SDPX_CHECK(dot_f64(b.X, b.dX, tX, b.Y, b.dY, tY, b.k2, h->red_part, h->dsc + L + l, h->stream));
  }
  SDPX_CHECK(sync_scalars(h, 2 * L));
  for (int i = 0; i < 2 * L; ++i) out[i] = h->hsc[i];
  return kOk;
}

int sdpx_dense_fraction_to_boundary(sdpx_dense_t* h, double pf, double df, double* tX, double* tY) {
  SDPX_CHECK(check_handle(h));
  double bx = 1.0, by = 1.0;
  for (int l = 0; l < h->L; ++l) {
    double vx = 1.0, vy = 1.0;
    SDPX_CHECK(ftb_block(h, h->blk[l], &vx, &vy));
    bx = std::min(bx, vx);
    by = std::min(by, vy);
  }
  *tX = bx < 1.0 ? pf * bx : 1.0;
  *tY = by < 1.0 ? df * by : 1.0;
  return kOk;
}

int sdpx_dense_backtrack(sdpx_dense_t* h, double gamma, double min_step, double ratio, double* tX, double* tY) {
  SDPX_CHECK(check_handle(h));
  return backtrack_both(h, gamma, min_step, ratio, tX, tY);
}

int sdpx_dense_update(sdpx_dense_t* h, double tX, double tY, double* out) {
  SDPX_CHECK(check_handle(h));
  const int L = h->L;
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(axpbyz_f64(b.k2, 1.0, b.X, tX, b.dX, b.X, h->stream));
    SDPX_CHECK(axpbyz_f64(b.k2, 1.0, b.Y, tY, b.dY, b.Y, h->stream));
    SDPX_CHECK(dot_f64(b.X, nullptr, 0.0, b.Y, nullptr, 0.0, b.k2, h->red_part, h->dsc + 2 + l, h->stream));
    SDPX_CHECK(absmax_f64(b.X, b.k2, h->red_part, h->dsc + 2 + L + 2 * l, h->stream));
    SDPX_CHECK(absmax_f64(b.Y, b.k2, h->red_part, h->dsc + 2 + L + 2 * l + 1, h->stream));
  }
  SDPX_CHECK(axpbyz_f64(h->m, 1.0, h->x, tX, h->dx, h->x, h->stream));
  SDPX_CHECK(absmax_f64(h->x, h->m, h->red_part, h->dsc + 1, h->stream));
  SDPX_CHECK(sync_scalars(h, 2 + 3 * L));
  double comp = 0.0;
  bool finite = std::isfinite(h->hsc[1]);
  for (int l = 0; l < L; ++l) {
    comp += h->hsc[2 + l];
    out[2 + l] = h->hsc[2 + l];
    finite = finite && std::isfinite(h->hsc[2 + L + 2 * l]) && std::isfinite(h->hsc[2 + L + 2 * l + 1]);
  }
  out[0] = comp;
  out[1] = finite ? 1.0 : 0.0;
  return kOk;
}

int sdpx_dense_accepted_residuals(sdpx_dense_t* h, double tX, double tY, double bpr, double* out) {
  SDPX_CHECK(check_handle(h));
  SDPX_CHECK(axpbyz_f64(h->m, 1.0 - tY, h->d, -tY, h->rho, h->d, h->stream));
  SDPX_CHECK(absmax_f64(h->d, h->m, h->red_part, h->dsc, h->stream));
  SDPX_CHECK(sync_scalars(h, 1));
  out[0] = std::fabs(1.0 - tX) * bpr;
  out[1] = h->hsc[0];
  return kOk;
}

int sdpx_dense_objectives(sdpx_dense_t* h, double* pobj, double* dobj) {
  SDPX_CHECK(check_handle(h));
  SDPX_CHECK(dot_f64(h->c, nullptr, 0.0, h->x, nullptr, 0.0, h->m, h->red_part, h->dsc, h->stream));
  for (int l = 0; l < h->L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(dot_f64(b.C, nullptr, 0.0, b.Y, nullptr, 0.0, b.k2, h->red_part, h->dsc + 1 + l, h->stream));
  }
  SDPX_CHECK(sync_scalars(h, 1 + h->L));
  *pobj = h->hsc[0];
  double d = 0.0;
  for (int l = 0; l < h->L; ++l) d += h->hsc[1 + l];
  *dobj = d;
  return kOk;
}

int sdpx_dense_solution_residuals(sdpx_dense_t* h, double* p_res, double* d_res) {
  SDPX_CHECK(check_handle(h));
  const int L = h->L;
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(avx_f64(b.Av, b.k2, h->m, h->x, 1.0, b.W1, -1.0, b.X, -1.0, b.C, h->gemv_part, h->gemv_part_cap,
                       h->stream));
    SDPX_CHECK(absmax_f64(b.W1, b.k2, h->red_part, h->dsc + l, h->stream));
  }
  SDPX_HIP(hipMemcpyAsync(h->v2, h->c, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream));
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(avt_f64(b.Av, b.k2, h->m, b.Y, -1.0, 1.0, h->v2, h->stream));
  }
  SDPX_CHECK(absmax_f64(h->v2, h->m, h->red_part, h->dsc + L, h->stream));
  SDPX_CHECK(sync_scalars(h, L + 1));
  double p = 0.0;
  for (int l = 0; l < L; ++l) p = std::max(p, h->hsc[l]);
  *p_res = p;
  *d_res = h->hsc[L];
  return kOk;
}

int sdpx_dense_scale_blocks(sdpx_dense_t* h, int which, double factor) {
  SDPX_CHECK(check_handle(h));
  for (auto& b : h->blk) {
    if (which == 1) SDPX_CHECK(axpbyz_f64(b.k2, factor, b.X, 0.0, nullptr, b.X, h->stream));
    if (which == 2) SDPX_CHECK(axpbyz_f64(b.k2, factor, b.Y, 0.0, nullptr, b.Y, h->stream));
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

// Condition estimates of the Schur system from its factor L (S = L L'):
//  * xPOCON's: ||S||_1 of the unfactored S (its column sums from the lower triangle) times
//    Higham's estimate of ||S^-1||_1, S^-1 x applied as two triangular solves;
//  * _triangular_condition_estimate's (mixed_precision_kkt.jl:371-392): xTRCON's reciprocal
//    condition numbers of L in the one and infinity norms, 1 / (rcond_1 rcond_inf), an upper
//    bound of cond_1(S) that the reference calls deliberately conservative.
// out (4) = cond_1(S) estimate, the xTRCON product, rcond_1(L), rcond_inf(L) (infinite / 0 when an
// estimate is not positive and finite).
int sdpx_dense_factor_condition(sdpx_dense_t* h, int which, double* out) {
  SDPX_CHECK(check_handle(h));
  if (h->arrow || !out || which < 1 || which > 3) {
    set_last_error("sdpx_dense_factor_condition: which in 1..3 and out (4) required; not defined for the "
                   "block-arrow factor");
    return kInvalidArgument;
  }
  const bool want_s = (which & 1) != 0, want_l = (which & 2) != 0;
  const int m = h->m;
  const int nsplit = std::max(1, std::min(32, ceil_div(m, 256)));
  const int chunk = ceil_div(m, nsplit);
  if (!h->cond_part) SDPX_CHECK(dalloc(h, &h->cond_part, (size_t)33 * m));
  // column sums of the lower triangle (rows >= j) and row sums (columns <= i), for L and for S
  auto norms = [&](const double* A, std::vector<double>& col, std::vector<double>& row) -> int {
    hipLaunchKernelGGL(tri_colsum_kernel, dim3(m), dim3(256), 0, h->stream, A, (long long)m, m,
                       h->cond_part + (size_t)32 * m);
    SDPX_HIP(hipGetLastError());
    hipLaunchKernelGGL(tri_rowsum_kernel, dim3(ceil_div(m, 256), nsplit), dim3(256), 0, h->stream, A, (long long)m, m,
                       chunk, h->cond_part);
    SDPX_HIP(hipGetLastError());
    std::vector<double> part((size_t)nsplit * m);
    col.resize(m);
    row.assign(m, 0.0);
    SDPX_HIP(hipMemcpyAsync(col.data(), h->cond_part + (size_t)32 * m, sizeof(double) * m, hipMemcpyDeviceToHost,
                            h->stream));
    SDPX_HIP(hipMemcpyAsync(part.data(), h->cond_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
    SDPX_HIP(hipStreamSynchronize(h->stream));
    for (int sp = 0; sp < nsplit; ++sp)
      for (int i = 0; i < m; ++i) row[i] += part[(size_t)sp * m + i];
    return kOk;
  };
  double l1 = 0.0, linf = 0.0, s1 = 0.0;
  if (want_l) {
    std::vector<double> lc, lr;
    SDPX_CHECK(norms(h->Sf, lc, lr));
    for (int j = 0; j < m; ++j) {
      l1 = std::max(l1, lc[j]);
      linf = std::max(linf, lr[j]);
    }
  }
  if (want_s) {
    std::vector<double> sc, sr, sdiag(m);
    SDPX_CHECK(norms(h->S, sc, sr));
    SDPX_HIP(hipMemcpy2DAsync(sdiag.data(), sizeof(double), h->S, sizeof(double) * (m + 1), sizeof(double), m,
                              hipMemcpyDeviceToHost, h->stream));
    SDPX_HIP(hipStreamSynchronize(h->stream));
    for (int j = 0; j < m; ++j) s1 = std::max(s1, sc[j] + sr[j] - std::fabs(sdiag[j]));  // column j of S
  }
  // one triangular solve with L (trans 0) or L' (trans 1) on a host vector, through the device
  auto solve = [&](std::vector<double>& x, int trans) -> int {
    SDPX_HIP(hipMemcpyAsync(h->delta, x.data(), sizeof(double) * m, hipMemcpyHostToDevice, h->stream));
    SDPX_CHECK(trsv_f64_steps(h->Sf, m, m, h->dinvS, h->delta, h->trsv_y, trans, h->stream));
    SDPX_HIP(hipMemcpyAsync(x.data(), h->delta, sizeof(double) * m, hipMemcpyDeviceToHost, h->stream));
    SDPX_HIP(hipStreamSynchronize(h->stream));
    return kOk;
  };
  auto solve_s = [&](std::vector<double>& x, int) -> int {
    SDPX_CHECK(solve(x, 0));
    return solve(x, 1);
  };
  double sinv = 0.0, inv1 = 0.0, invinf = 0.0;
  if (want_s) SDPX_CHECK(onenorm_estimate(m, solve_s, &sinv));
  if (want_l) {
    // ||L^-1||_1: B = L^-1; ||L^-1||_inf = ||L^-T||_1: B = L^-T
    SDPX_CHECK(onenorm_estimate(m, [&](std::vector<double>& x, int t) { return solve(x, t); }, &inv1));
    SDPX_CHECK(onenorm_estimate(m, [&](std::vector<double>& x, int t) { return solve(x, 1 - t); }, &invinf));
  }
  const double rcs = (s1 > 0.0 && sinv > 0.0) ? (1.0 / sinv) / s1 : 0.0;
  const double rc1 = (l1 > 0.0 && inv1 > 0.0) ? (1.0 / l1) / inv1 : 0.0;
  const double rcinf = (linf > 0.0 && invinf > 0.0) ? (1.0 / linf) / invinf : 0.0;
  const double conds = rcs > 0.0 ? 1.0 / rcs : INFINITY;
  const double condl = (rc1 > 0.0 && rcinf > 0.0) ? 1.0 / (rc1 * rcinf) : INFINITY;
  out[0] = std::isfinite(conds) ? conds : INFINITY;
  out[1] = std::isfinite(condl) ? condl : INFINITY;
  out[2] = rc1;
  out[3] = rcinf;
  return kOk;
}

int sdpx_dense_factor_quality(sdpx_dense_t* h, double* out) {
  SDPX_CHECK(check_handle(h));
  const int L = h->L;
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(diag_extrema_f64(b.LX, b.k, 0, b.k, 1, h->dsc + 4 * l, h->stream));
    SDPX_CHECK(diag_extrema_f64(b.MY, b.k, 0, b.k, 1, h->dsc + 4 * l + 2, h->stream));
  }
  if (h->arrow) {
    // _kkt_factorization_quality with an arrow workspace (step.jl:266-300): the reduced factor
    // alone, 1 when there is no global variable
    if (h->ng > 0)
      SDPX_CHECK(diag_extrema_f64(h->Sf, h->ng, 0, h->ng, 1, h->dsc + 4 * L, h->stream));
    else
      SDPX_CHECK(fill_f64(2, 1.0, h->dsc + 4 * L, h->stream));
  } else {
    SDPX_CHECK(diag_extrema_f64(h->Sf, h->m, 0, h->m, 1, h->dsc + 4 * L, h->stream));
  }
  SDPX_CHECK(sync_scalars(h, 4 * L + 2));
  auto quality = [](double mn, double mx) {
    return (mx > 0.0 && std::isfinite(mx)) ? std::min(std::max(mn / mx, 0.0), 1.0) : 0.0;
  };
  double pm = 1.0, dm = 1.0;
  for (int l = 0; l < L; ++l) {
    pm = std::min(pm, quality(h->hsc[4 * l], h->hsc[4 * l + 1]));
    dm = std::min(dm, quality(h->hsc[4 * l + 2], h->hsc[4 * l + 3]));
  }
  out[0] = pm;
  out[1] = dm;
  out[2] = quality(h->hsc[4 * L], h->hsc[4 * L + 1]);
  return kOk;
}

int sdpx_dense_equilibrate(sdpx_dense_t* h, int ruiz_iters, double* E, double* s_out, double* objective_scale,
                           int* passes) {
  SDPX_CHECK(check_handle(h));
  h->av_ver += 1;  // the Ruiz passes rescale Av
  const int L = h->L, m = h->m;
  for (int l = 0; l < L; ++l) {
    if (h->blk[l].na > 0) {
      set_last_error("sdpx_dense_equilibrate: sparse blocks are taken as given (equilibrate before packing)");
      return kInvalidArgument;
    }
  }
  // _ruiz_control (ingest.jl:1141-1160): :auto = 2..8 passes, log tolerance 1/20
  const bool adaptive = ruiz_iters < 0;
  const int min_it = adaptive ? 2 : ruiz_iters, max_it = adaptive ? 8 : ruiz_iters;
  const double log_tol = adaptive ? 1.0 / 20.0 : 0.0;
  unsigned long long* rowmax = nullptr;
  double *e_dev = nullptr, *s_dev = nullptr, *maxima = nullptr;
  int kmax = 1;
  for (const Block& b : h->blk) kmax = std::max(kmax, b.k);
  SDPX_CHECK(dalloc(h, &rowmax, (size_t)kmax));
  SDPX_CHECK(dalloc(h, &e_dev, (size_t)kmax));
  SDPX_CHECK(dalloc(h, &s_dev, (size_t)m));
  SDPX_CHECK(dalloc(h, &maxima, (size_t)m));
  std::vector<unsigned long long> rn_bits(kmax);
  std::vector<double> rn(kmax), e(kmax);
  long long eoff = 0;
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    const int k = b.k;
    const int nslab = ceil_div(k, kRuizRows);
    const int groups = ceil_div(1024, nslab) * nslab;  // 4 waves each, so waves % nslab == 0
    std::vector<double> El(k, 1.0);
    // row maxima of the unscaled block (the first pass reads only)
    SDPX_HIP(hipMemsetAsync(rowmax, 0, sizeof(unsigned long long) * k, h->stream));
    hipLaunchKernelGGL(ruiz_pass_kernel, dim3(groups), dim3(256), 0, h->stream, b.Av, b.C, k, m, nullptr, rowmax,
                       nslab);
    SDPX_HIP(hipGetLastError());
    int done = 0;
    for (int it = 1; it <= max_it; ++it) {
      done = it;
      SDPX_HIP(hipMemcpyAsync(rn_bits.data(), rowmax, sizeof(unsigned long long) * k, hipMemcpyDeviceToHost,
                              h->stream));
      SDPX_HIP(hipStreamSynchronize(h->stream));
      // _ruiz_step / _ruiz_converged (ingest.jl:1162-1176, 1224-1231)
      bool converged = adaptive && it >= min_it;
      double dev = 0.0;
      for (int r = 0; r < k; ++r) {
        double v;
        std::memcpy(&v, &rn_bits[r], sizeof(double));
        rn[r] = v;
        if (!(v > 0.0)) continue;
        if (!std::isfinite(v)) converged = false;
        else dev = std::max(dev, std::fabs(std::log(v)));
      }
      converged = converged && dev <= log_tol;
      for (int r = 0; r < k; ++r) e[r] = rn[r] > 0.0 ? 1.0 / std::sqrt(rn[r]) : 1.0;
      for (int r = 0; r < k; ++r) El[r] *= e[r];
      // scale the block and C, and take the next pass's row maxima in the same sweep
      const bool last = converged || it == max_it;
      SDPX_HIP(hipMemcpyAsync(e_dev, e.data(), sizeof(double) * k, hipMemcpyHostToDevice, h->stream));
      if (!last) SDPX_HIP(hipMemsetAsync(rowmax, 0, sizeof(unsigned long long) * k, h->stream));
      hipLaunchKernelGGL(ruiz_pass_kernel, dim3(groups), dim3(256), 0, h->stream, b.Av, b.C, k, m, e_dev,
                         last ? nullptr : rowmax, nslab);
      SDPX_HIP(hipGetLastError());
      if (last) break;
    }
    if (max_it == 0) done = 0;
    if (passes) passes[l] = done;
    if (E) std::copy(El.begin(), El.end(), E + eoff);
    eoff += k;
    // per-constraint maxima over the blocks (ingest.jl:1470-1481)
    hipLaunchKernelGGL(ruiz_colmax_kernel, dim3((unsigned)m), dim3(256), 0, h->stream, b.Av, b.k2, maxima, l == 0);
    SDPX_HIP(hipGetLastError());
  }
  // _equilibration_variable_scales (ingest.jl:1252-1270)
  std::vector<double> mx(m), sv(m), cc(m);
  SDPX_HIP(hipMemcpyAsync(mx.data(), maxima, sizeof(double) * m, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipMemcpyAsync(cc.data(), h->c, sizeof(double) * m, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  double osc = 0.0;
  for (int i = 0; i < m; ++i) {
    sv[i] = mx[i] > 0.0 ? mx[i] : 1.0;
    cc[i] = cc[i] / sv[i];
    osc = std::max(osc, std::fabs(cc[i]));
  }
  if (!(osc > 0.0 && std::isfinite(osc))) osc = 1.0;
  for (int i = 0; i < m; ++i) cc[i] = cc[i] / osc;
  SDPX_HIP(hipMemcpyAsync(s_dev, sv.data(), sizeof(double) * m, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->c, cc.data(), sizeof(double) * m, hipMemcpyHostToDevice, h->stream));
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    hipLaunchKernelGGL(div_columns_kernel, dim3((unsigned)ceil_div(b.k2 * m, 256)), dim3(256), 0, h->stream, b.Av,
                       b.k2, m, s_dev);
    SDPX_HIP(hipGetLastError());
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  if (s_out) std::copy(sv.begin(), sv.end(), s_out);
  if (objective_scale) *objective_scale = osc;
  return kOk;
}

int sdpx_dense_get_data(sdpx_dense_t* h, double* C, double* c) {
  SDPX_CHECK(check_handle(h));
  long long off = 0;
  for (const Block& b : h->blk) {
    if (C) SDPX_HIP(hipMemcpyAsync(C + off, b.C, sizeof(double) * b.k2, hipMemcpyDeviceToHost, h->stream));
    off += b.k2;
  }
  if (c) SDPX_HIP(hipMemcpyAsync(c, h->c, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_cold_start(sdpx_dense_t* h, double beta, double eps_primal, double eps_dual, double* mu,
                          double* out) {
  SDPX_CHECK(check_handle(h));
  const int L = h->L, m = h->m;
  double rec[16] = {0.0};
  auto finish = [&](int reason) {
    rec[0] = reason == 0 ? 1.0 : 0.0;
    rec[1] = reason;
    if (out) std::copy(rec, rec + 16, out);
    return kOk;
  };
  long long degree = 0;
  for (const Block& b : h->blk) degree += b.k;
  const double threshold = std::max(std::sqrt(2.220446049250313e-16), std::max(eps_primal, eps_dual));
  // identity metric: X = Y = I, one Schur build (the Gram of the A_i) and one factor
  for (Block& b : h->blk) {
    SDPX_CHECK(fill_f64(b.k2, 0.0, b.X, h->stream));
    SDPX_CHECK(add_diag_f64(b.X, b.k, b.k, 1.0, nullptr, 0.0, h->stream));
    SDPX_HIP(hipMemcpyAsync(b.Y, b.X, sizeof(double) * b.k2, hipMemcpyDeviceToDevice, h->stream));
  }
  SDPX_CHECK(fill_f64(m, 0.0, h->x, h->stream));
  std::vector<double> zmu(L, 0.0);
  double tmp[3];
  SDPX_CHECK(sdpx_dense_residuals(h, zmu.data(), 1, 1, tmp));
  SDPX_CHECK(sdpx_dense_schur_build(h));
  int reg = 0;
  const int fst = sdpx_dense_schur_factor(h, &reg);
  if (fst < 0) return fst;
  rec[2] = reg;
  if (fst != 0) return finish(1);  // :identity_kkt_factorization
  // the two right-hand sides with the one factor, each checked against the unfactored
  // operator and refined when its normalized residual exceeds the gate
  int refine_total = 0;
  auto solve_gated = [&](double& normalized) -> int {
    SDPX_CHECK(solve_into(h, h->r, h->dx));
    SDPX_CHECK(direction_residual_dev(h, 0));
    SDPX_CHECK(sync_scalars(h, 2));
    const double scale = std::max(h->hsc[1], 1.0);
    normalized = h->hsc[0] / scale;
    if (std::isfinite(normalized) && normalized > threshold) {
      int steps = 0;
      double res = 0.0;
      SDPX_CHECK(sdpx_dense_refine(h, 64.0 * 2.220446049250313e-16 * scale, 8, 0, &steps, &res));
      refine_total += steps;
      SDPX_CHECK(direction_residual_dev(h, 0));
      SDPX_CHECK(sync_scalars(h, 1));
      normalized = h->hsc[0] / scale;
    }
    return kOk;
  };
  // primal: S x = [<A_i, C>]
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(avt_f64(b.Av, b.k2, m, b.C, 1.0, l == 0 ? 0.0 : 1.0, h->r, h->stream));
  }
  double pn = 0.0, dn = 0.0;
  SDPX_CHECK(solve_gated(pn));
  rec[3] = pn;
  if (!(std::isfinite(pn) && pn <= threshold)) return finish(2);  // :kkt_cold_start_solve
  SDPX_HIP(hipMemcpyAsync(h->x, h->dx, sizeof(double) * m, hipMemcpyDeviceToDevice, h->stream));
  // dual: S v = c
  SDPX_HIP(hipMemcpyAsync(h->r, h->c, sizeof(double) * m, hipMemcpyDeviceToDevice, h->stream));
  SDPX_CHECK(solve_gated(dn));
  rec[4] = dn;
  rec[5] = refine_total;
  if (!(std::isfinite(dn) && dn <= threshold)) return finish(2);
  // X = A(x) - C, Y = A(v)
  for (Block& b : h->blk) {
    SDPX_CHECK(avx_f64(b.Av, b.k2, m, h->x, 1.0, b.X, -1.0, b.C, 0.0, nullptr, h->gemv_part, h->gemv_part_cap,
                       h->stream));
    SDPX_CHECK(avx_f64(b.Av, b.k2, m, h->dx, 1.0, b.Y, 0.0, nullptr, 0.0, nullptr, h->gemv_part, h->gemv_part_cap,
                       h->stream));
  }
  // pre-shift residuals, normalized as for termination
  double p_res = 0.0, d_res = 0.0;
  SDPX_CHECK(sdpx_dense_solution_residuals(h, &p_res, &d_res));
  double cmax = 0.0;
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_CHECK(absmax_f64(b.C, b.k2, h->red_part, h->dsc + l, h->stream));
  }
  SDPX_CHECK(absmax_f64(h->c, m, h->red_part, h->dsc + L, h->stream));
  SDPX_CHECK(sync_scalars(h, L + 1));
  for (int l = 0; l < L; ++l) cmax = std::max(cmax, h->hsc[l]);
  const double npres = p_res / (1.0 + cmax), ndres = d_res / (1.0 + h->hsc[L]);
  rec[6] = npres;
  rec[7] = ndres;
  if (!(std::isfinite(npres) && std::isfinite(ndres) && npres <= threshold && ndres <= threshold))
    return finish(3);  // :kkt_cold_start_residuals
  // PSD shifts, identity mass floor and cross-centering on the host (k x k blocks)
  std::vector<std::vector<double>> X(L), Y(L);
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    X[l].resize(b.k2);
    Y[l].resize(b.k2);
    SDPX_HIP(hipMemcpyAsync(X[l].data(), b.X, sizeof(double) * b.k2, hipMemcpyDeviceToHost, h->stream));
    SDPX_HIP(hipMemcpyAsync(Y[l].data(), b.Y, sizeof(double) * b.k2, hipMemcpyDeviceToHost, h->stream));
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  double max_ps = 0.0, max_ds = 0.0, pmarg = 1.0, dmarg = 1.0;
  for (int l = 0; l < L; ++l) {
    const int k = h->blk[l].k;
    double ps, dsh, pm, dm, sc;
    const bool pok = cold::psd_shift(X[l].data(), k, ps, pm, sc);
    const bool dok = cold::psd_shift(Y[l].data(), k, dsh, dm, sc);
    if (!(pok && dok)) return finish(4);  // :psd_shift
    max_ps = std::max(max_ps, ps);
    max_ds = std::max(max_ds, dsh);
  }
  double pmass = 0.0, dmass = 0.0;
  for (int l = 0; l < L; ++l) {
    pmass += cold::trace(X[l], h->blk[l].k);
    dmass += cold::trace(Y[l], h->blk[l].k);
  }
  double pfloor = 0.0, dfloor = 0.0;
  if (!cold::mass_shifts(pmass, dmass, degree, pfloor, dfloor)) return finish(5);  // :identity_mass
  for (int l = 0; l < L; ++l) {
    const int k = h->blk[l].k;
    for (int i = 0; i < k; ++i) {
      if (pfloor > 0.0) X[l][i + (long long)i * k] += pfloor;
      if (dfloor > 0.0) Y[l][i + (long long)i * k] += dfloor;
    }
  }
  double kappa_floor = 0.0;
  pmass = dmass = 0.0;
  for (int l = 0; l < L; ++l) {
    kappa_floor += cold::kdot(X[l], Y[l]);
    pmass += cold::trace(X[l], h->blk[l].k);
    dmass += cold::trace(Y[l], h->blk[l].k);
  }
  double pc = 0.0, dc = 0.0;
  if (!cold::centering_shifts(kappa_floor, pmass, dmass, pc, dc)) return finish(6);  // :centering
  double kappa_after = 0.0;
  for (int l = 0; l < L; ++l) {
    const int k = h->blk[l].k;
    for (int i = 0; i < k; ++i) {
      if (pc > 0.0) X[l][i + (long long)i * k] += pc;
      if (dc > 0.0) Y[l][i + (long long)i * k] += dc;
    }
    double mg, sc;
    cold::psd_stats(X[l].data(), k, sc, mg);
    pmarg = l == 0 ? mg : std::min(pmarg, mg);
    cold::psd_stats(Y[l].data(), k, sc, mg);
    dmarg = l == 0 ? mg : std::min(dmarg, mg);
    const double kd = cold::kdot(X[l], Y[l]);
    kappa_after += kd;
    if (mu) mu[l] = beta * kd / k;
  }
  rec[8] = max_ps + std::max(pfloor, 0.0) + std::max(pc, 0.0);
  rec[9] = max_ds + std::max(dfloor, 0.0) + std::max(dc, 0.0);
  rec[10] = pmarg;
  rec[11] = dmarg;
  rec[12] = kappa_after / (double)std::max(degree, 1LL);
  rec[13] = kappa_floor / (double)std::max(degree, 1LL);
  for (int l = 0; l < L; ++l) {
    Block& b = h->blk[l];
    SDPX_HIP(hipMemcpyAsync(b.X, X[l].data(), sizeof(double) * b.k2, hipMemcpyHostToDevice, h->stream));
    SDPX_HIP(hipMemcpyAsync(b.Y, Y[l].data(), sizeof(double) * b.k2, hipMemcpyHostToDevice, h->stream));
  }
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return finish(0);
}

int sdpx_dense_set_equalities(sdpx_dense_t* h, int n, const double* B) {
  SDPX_CHECK(check_handle(h));
  if (n < 1 || !B || h->neq != 0) {
    set_last_error("sdpx_dense_set_equalities: n >= 1 and B (m x n) required, once per handle");
    return kInvalidArgument;
  }
  if (h->arrow) {
    set_last_error("sdpx_dense_set_equalities: the block-arrow route takes no equality border");
    return kInvalidArgument;
  }
  const long long m = h->m;
  const long long nbq = (long long)ceil_div(n, kCholNb) * kCholNb * kCholNb;
  SDPX_CHECK(dalloc(h, &h->B, (size_t)(m * n)));
  SDPX_CHECK(dalloc(h, &h->Bt, (size_t)(m * n)));
  SDPX_CHECK(dalloc(h, &h->Btmp, (size_t)kCholNb * n));
  SDPX_CHECK(dalloc(h, &h->Q, (size_t)n * n));
  SDPX_CHECK(dalloc(h, &h->Qf, (size_t)n * n));
  SDPX_CHECK(dalloc(h, &h->dinvQ, (size_t)nbq));
  SDPX_CHECK(dalloc(h, &h->eqscale, (size_t)n));
  SDPX_CHECK(dalloc(h, &h->qrhs, (size_t)n));
  SDPX_CHECK(dalloc(h, &h->qy, (size_t)n));
  SDPX_CHECK(dalloc(h, &h->peq, (size_t)n));
  SDPX_CHECK(dalloc(h, &h->rtil, (size_t)m));
  SDPX_HIP(hipMemcpyAsync(h->B, B, sizeof(double) * (size_t)(m * n), hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  h->neq = n;
  return kOk;
}

int sdpx_dense_equality_factor(sdpx_dense_t* h, double requested, double* out) {
  SDPX_CHECK(check_handle(h));
  const int m = h->m, n = h->neq;
  if (n == 0) {
    set_last_error("sdpx_dense_equality_factor: no equalities (sdpx_dense_set_equalities)");
    return kInvalidArgument;
  }
  h->eq_factored = false;
  // B-tilde = L_S^-1 B (la_trsm!, kkt.jl:1471): blocked forward substitution with the Schur
  // factor, each 64-row block = D_j (B_j - L(j, <j) B-tilde_<j) with the factor's diagonal-block inverses
  SDPX_HIP(hipMemcpyAsync(h->Bt, h->B, sizeof(double) * (size_t)m * n, hipMemcpyDeviceToDevice, h->stream));
  for (int j0 = 0; j0 < m; j0 += kCholNb) {
    const int jb = std::min(kCholNb, m - j0);
    if (j0 > 0)
      SDPX_CHECK(gemm_f64(gemm(jb, n, j0, -1.0, h->Sf + j0, m, h->Bt, m, 1.0, h->Bt + j0, m, 0), h->stream));
    SDPX_CHECK(gemm_f64(gemm(jb, n, jb, 1.0, h->dinvS + (long long)(j0 / kCholNb) * kCholNb * kCholNb, kCholNb,
                             h->Bt + j0, m, 0.0, h->Btmp, kCholNb, kGemmALower),
                        h->stream));
    hipLaunchKernelGGL(copy_block_kernel, dim3((unsigned)ceil_div((long long)jb * n, 256)), dim3(256), 0, h->stream,
                       h->Btmp, (long long)kCholNb, h->Bt + j0, (long long)m, jb, n);
    SDPX_HIP(hipGetLastError());
  }
  // D: per-column infinity-norm normalization (kkt.jl:1472)
  hipLaunchKernelGGL(normalize_columns_kernel, dim3((unsigned)n), dim3(256), 0, h->stream, h->Bt, (long long)m, m,
                     h->eqscale);
  SDPX_HIP(hipGetLastError());
  // Q = B-hat' B-hat (lower) and its Cholesky (kkt.jl:1495-1510)
  SDPX_CHECK(gemm_f64(gemm(n, n, m, 1.0, h->Bt, m, h->Bt, m, 0.0, h->Q, n, kGemmTransA | kGemmLowerOut), h->stream));
  SDPX_CHECK(copy_lower_shift_f64(h->Q, h->Qf, n, 0.0, h->stream));
  SDPX_CHECK(potrf_f64(h->Qf, n, 0, n, 1, h->dinvQ, 0, h->dinfo, h->stream));
  SDPX_CHECK(diag_extrema_f64(h->Qf, n, 0, n, 1, h->dsc, h->stream));
  SDPX_CHECK(sync_scalars(h, 2, 1));
  const int info = h->hinfo[0];
  const double mn = h->hsc[0], mx = h->hsc[1];
  // _equality_cholesky_rank_tau / _cholesky_has_numerical_rank (kkt.jl:76-113, 122-176)
  const double eps = 2.220446049250313e-16;
  double req = requested > 0.0 ? requested : std::sqrt(eps);
  req = std::min(std::sqrt(eps), req);
  const double tau = std::max((double)std::max(m, n) * eps, std::min(std::sqrt(eps), req));
  int status = 0;
  if (info != 0 || !std::isfinite(mn) || !std::isfinite(mx) || !(mx > 0.0)) {
    status = 1;  // the Gram factor failed
  } else if (!(mn > tau * mx)) {
    status = 2;  // numerically rank deficient
  } else if (mn <= std::sqrt(tau) * mx) {
    // suspicious regime: POTRF may round the zero pivot of exactly duplicated columns up
    // to O(sqrt(eps)); scan B-tilde for exact duplicate columns (kkt.jl:168-175, 406-440)
    std::vector<double> bt((size_t)m * n);
    SDPX_HIP(hipMemcpyAsync(bt.data(), h->Bt, sizeof(double) * bt.size(), hipMemcpyDeviceToHost, h->stream));
    SDPX_HIP(hipStreamSynchronize(h->stream));
    for (int a = 0; a < n && status == 0; ++a)
      for (int b = a + 1; b < n && status == 0; ++b) {
        bool same = true;
        for (int r = 0; r < m && same; ++r) {
          const double x = bt[r + (size_t)a * m], y = bt[r + (size_t)b * m];
          same = (x == y) || (x == 0.0 && y == 0.0);
        }
        if (same) status = 2;
      }
  }
  h->eq_factored = status == 0;
  if (out) {
    out[0] = status;
    out[1] = mn;
    out[2] = mx;
    out[3] = tau;
  }
  return status == 0 ? kOk : kNotPositiveDefinite;
}

int sdpx_dense_solve_equality(sdpx_dense_t* h, const double* p, double* dy) {
  SDPX_CHECK(check_handle(h));
  const int m = h->m, n = h->neq;
  if (!h->eq_factored || !p) {
    set_last_error("sdpx_dense_solve_equality: a successful sdpx_dense_equality_factor and p are required");
    return kInvalidArgument;
  }
  // _solve_dense_kkt_owned! (kkt.jl:2570-2595): r~ = L^-1 r; q = D p - B-hat' r~;
  // dy^ = Q^-1 q; dx = L^-T (r~ + B-hat dy^); dy = D dy^
  SDPX_HIP(hipMemcpyAsync(h->rtil, h->r, sizeof(double) * m, hipMemcpyDeviceToDevice, h->stream));
  SDPX_CHECK(trsv_f64_steps(h->Sf, m, m, h->dinvS, h->rtil, h->trsv_y, 0, h->stream));
  SDPX_CHECK(gemm_f64(gemm(n, 1, m, 1.0, h->Bt, m, h->rtil, m, 0.0, h->qrhs, n, kGemmTransA), h->stream));
  SDPX_HIP(hipMemcpyAsync(h->peq, p, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(scaled_rhs_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, h->eqscale, h->peq, h->qrhs,
                     n, 0);
  SDPX_HIP(hipGetLastError());
  SDPX_CHECK(trsv_f64_steps(h->Qf, n, n, h->dinvQ, h->qrhs, h->qy, 0, h->stream));
  SDPX_CHECK(trsv_f64_steps(h->Qf, n, n, h->dinvQ, h->qrhs, h->qy, 1, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->dx, h->rtil, sizeof(double) * m, hipMemcpyDeviceToDevice, h->stream));
  SDPX_CHECK(gemm_f64(gemm(m, 1, n, 1.0, h->Bt, m, h->qrhs, n, 1.0, h->dx, m, 0), h->stream));
  SDPX_CHECK(trsv_f64_steps(h->Sf, m, m, h->dinvS, h->dx, h->trsv_y, 1, h->stream));
  hipLaunchKernelGGL(scaled_rhs_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, h->eqscale, nullptr,
                     h->qrhs, n, 1);
  SDPX_HIP(hipGetLastError());
  if (dy) SDPX_HIP(hipMemcpyAsync(dy, h->qrhs, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_get_equality(sdpx_dense_t* h, double* Bt, double* Q, double* scale) {
  SDPX_CHECK(check_handle(h));
  const long long m = h->m, n = h->neq;
  if (Bt) SDPX_HIP(hipMemcpyAsync(Bt, h->Bt, sizeof(double) * m * n, hipMemcpyDeviceToHost, h->stream));
  if (Q) SDPX_HIP(hipMemcpyAsync(Q, h->Q, sizeof(double) * n * n, hipMemcpyDeviceToHost, h->stream));
  if (scale) SDPX_HIP(hipMemcpyAsync(scale, h->eqscale, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_dense_set_block_sparse(sdpx_dense_t* h, int l, int na, const int* ids, const int* ptr, const int* lin,
                                const int* row, const int* col, const double* val, const double* C) {
  SDPX_CHECK(check_handle(h));
  if (l < 0 || l >= h->L || na < 1 || na > h->m || !ids || !ptr || !C) {
    set_last_error("sdpx_dense_set_block_sparse: block index, 1 <= na <= m, ids, ptr and C required");
    return kInvalidArgument;
  }
  Block& b = h->blk[l];
  if (b.na != 0) {
    set_last_error("sdpx_dense_set_block_sparse: the block is already sparse (set once)");
    return kInvalidArgument;
  }
  if (h->precision != 0) {
    set_last_error("sdpx_dense_set_block_sparse: sparse blocks assemble in FP64 only");
    return kInvalidArgument;
  }
  const int k = b.k;
  // SparseBlockCOO as stored (types/constraints.jl:59-109): 1-based ptr / row / col / lin
  if (ptr[0] != 1) {
    set_last_error("sdpx_dense_set_block_sparse: ptr[1] must be 1 (1-based positions)");
    return kInvalidArgument;
  }
  for (int p = 0; p < na; ++p) {
    if (ptr[p + 1] < ptr[p] || ids[p] < 0 || ids[p] >= h->m || (p > 0 && ids[p] <= ids[p - 1])) {
      set_last_error("sdpx_dense_set_block_sparse: ptr non-decreasing, ids strictly increasing in [0, m)");
      return kInvalidArgument;
    }
  }
  const int nnz = ptr[na] - 1;
  if (nnz > 0 && (!lin || !row || !col || !val)) {
    set_last_error("sdpx_dense_set_block_sparse: null lin, row, col or val with entries present");
    return kInvalidArgument;
  }
  std::vector<int> p0(na + 1), r0(nnz), c0(nnz), l0(nnz);
  for (int p = 0; p <= na; ++p) p0[p] = ptr[p] - 1;
  for (int t = 0; t < nnz; ++t) {
    if (row[t] < 1 || row[t] > k || col[t] < 1 || col[t] > k || lin[t] != (col[t] - 1) * k + row[t]) {
      set_last_error("sdpx_dense_set_block_sparse: entry " + std::to_string(t + 1) +
                     " outside the block or lin != (col-1)k+row");
      return kInvalidArgument;
    }
    r0[t] = row[t] - 1;
    c0[t] = col[t] - 1;
    l0[t] = lin[t] - 1;
  }
  SDPX_CHECK(dalloc(h, &b.ids, (size_t)na));
  SDPX_CHECK(dalloc(h, &b.sptr, (size_t)na + 1));
  SDPX_CHECK(dalloc(h, &b.srow, (size_t)std::max(nnz, 1)));
  SDPX_CHECK(dalloc(h, &b.scol, (size_t)std::max(nnz, 1)));
  SDPX_CHECK(dalloc(h, &b.slin, (size_t)std::max(nnz, 1)));
  SDPX_CHECK(dalloc(h, &b.sval, (size_t)std::max(nnz, 1)));
  SDPX_CHECK(dalloc(h, &b.W1T, (size_t)b.k2 * na));
  SDPX_HIP(hipMemcpyAsync(b.ids, ids, sizeof(int) * na, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipMemcpyAsync(b.sptr, p0.data(), sizeof(int) * (na + 1), hipMemcpyHostToDevice, h->stream));
  if (nnz > 0) {
    SDPX_HIP(hipMemcpyAsync(b.srow, r0.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, h->stream));
    SDPX_HIP(hipMemcpyAsync(b.scol, c0.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, h->stream));
    SDPX_HIP(hipMemcpyAsync(b.slin, l0.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, h->stream));
    SDPX_HIP(hipMemcpyAsync(b.sval, val, sizeof(double) * nnz, hipMemcpyHostToDevice, h->stream));
  }
  // the dense panel the Av passes (residuals, right-hand sides, direction recovery) read
  SDPX_CHECK(fill_f64(b.k2 * h->m, 0.0, b.Av, h->stream));
  h->av_ver += 1;
  hipLaunchKernelGGL(sparse_densify_kernel, dim3((unsigned)ceil_div(na, 256)), dim3(256), 0, h->stream, na, b.ids,
                     b.sptr, b.slin, b.sval, b.k2, b.Av);
  SDPX_HIP(hipGetLastError());
  SDPX_HIP(hipMemcpyAsync(b.C, C, sizeof(double) * b.k2, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  b.na = na;
  return kOk;
}

int sdpx_dense_set_arrow(sdpx_dense_t* h, const int* owner) {
  SDPX_CHECK(check_handle(h));
  const int m = h->m, L = h->L;
  if (!owner || h->arrow || h->neq != 0 || h->precision != 0) {
    set_last_error("sdpx_dense_set_arrow: owner (m) required, once per handle, with no equality border and "
                   "precision 0");
    return kInvalidArgument;
  }
  // permuted order: the local variables of block 0, 1, ... (ascending within a block), then the globals
  std::vector<std::vector<int>> loc(L);
  std::vector<int> glob;
  for (int i = 0; i < m; ++i) {
    if (owner[i] < -1 || owner[i] >= L) {
      set_last_error("sdpx_dense_set_arrow: owner[" + std::to_string(i) + "] outside [-1, L)");
      return kInvalidArgument;
    }
    if (owner[i] < 0)
      glob.push_back(i);
    else
      loc[owner[i]].push_back(i);
  }
  std::vector<int> perm, off{0};
  std::vector<long long> foff;
  long long fsz = 0;
  int max_q = 0;
  for (int l = 0; l < L; ++l) {
    const int q = (int)loc[l].size();
    if (q == 0) continue;
    if (q > kArrowMaxLocal) {
      set_last_error("sdpx_dense_set_arrow: block " + std::to_string(l) + " has " + std::to_string(q) +
                     " local variables; at most " + std::to_string(kArrowMaxLocal) + " are supported");
      return kInvalidArgument;
    }
    perm.insert(perm.end(), loc[l].begin(), loc[l].end());
    off.push_back((int)perm.size());
    foff.push_back(fsz);
    fsz += (long long)q * q;
    max_q = std::max(max_q, q);
  }
  const int mloc = (int)perm.size(), ng = (int)glob.size(), nloc = (int)foff.size();
  if (mloc == 0) {
    set_last_error("sdpx_dense_set_arrow: no local variable (every variable touches several blocks)");
    return kInvalidArgument;
  }
  perm.insert(perm.end(), glob.begin(), glob.end());
  const size_t lds = sizeof(double) * (size_t)max_q * (max_q + 1);
  if (lds > 64 * 1024)
    SDPX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(arrow_local_factor_kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  SDPX_CHECK(dalloc(h, &h->perm, (size_t)m));
  SDPX_CHECK(dalloc(h, &h->loc_off, (size_t)nloc + 1));
  SDPX_CHECK(dalloc(h, &h->loc_foff, (size_t)nloc));
  SDPX_CHECK(dalloc(h, &h->loc_list, (size_t)nloc));
  SDPX_CHECK(dalloc(h, &h->loc_info, (size_t)nloc));
  SDPX_CHECK(dalloc(h, &h->loc_reg, (size_t)nloc));
  SDPX_CHECK(dalloc(h, &h->Lloc, (size_t)fsz));
  SDPX_CHECK(dalloc(h, &h->AT, (size_t)std::max<long long>((long long)ng * mloc, 1)));
  SDPX_CHECK(dalloc(h, &h->Sred, (size_t)std::max<long long>((long long)ng * ng, 1)));
  SDPX_CHECK(dalloc(h, &h->xp, (size_t)m));
  h->arrow_part_cap = std::max<long long>((long long)ng * 64, 4096);
  SDPX_CHECK(dalloc(h, &h->arrow_part, (size_t)h->arrow_part_cap));
  SDPX_HIP(hipMemcpyAsync(h->perm, perm.data(), sizeof(int) * m, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->loc_off, off.data(), sizeof(int) * (nloc + 1), hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->loc_foff, foff.data(), sizeof(long long) * nloc, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  h->ng = ng;
  h->mloc = mloc;
  h->nloc = nloc;
  h->max_q = max_q;
  h->arrow = true;
  return kOk;
}

int sdpx_dense_get_arrow(sdpx_dense_t* h, int* ng, int* nloc, int* max_q) {
  SDPX_CHECK(check_handle(h));
  if (ng) *ng = h->arrow ? h->ng : -1;
  if (nloc) *nloc = h->nloc;
  if (max_q) *max_q = h->max_q;
  return kOk;
}

}  // extern "C"
