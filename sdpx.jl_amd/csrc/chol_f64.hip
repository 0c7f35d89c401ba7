// Dense FP64 Cholesky (batched, right-looking, blocked) and triangular solves.
//
// The diagonal nb x nb block of every panel is factored by one wavefront in
// registers; the kernel also inverts that block (kept in `dinv`), so that the
// panel solve, the trailing update and every later triangular solve are MFMA
// GEMMs (gemm_f64.hip).  A non-positive or non-finite pivot sets the 1-based
// LAPACK-style `info` and every later kernel of that batch entry returns
// immediately (the failure is read once, after the factorization).
#include "chol_diag.hpp"
#include "common.hpp"

namespace sdpx {
namespace {

using chol::kDiagLds;
using chol::kDiagThreads;
constexpr int NB = kCholNb;

__global__ __launch_bounds__(kDiagThreads) void potrf_diag_kernel(double* A, long long lda, long long stride, int j, int jb,
                                                         double* dinv, long long dinv_stride, int* info) {
  const int b = blockIdx.x;
  if (info[b] != 0) return;
  __shared__ __align__(16) double lds[kDiagLds];
  chol::potrf_diag_body(A + (long long)b * stride + j + (long long)j * lda, lda, j, jb,
                        dinv + (long long)b * dinv_stride + (long long)(j / NB) * NB * NB, info + b, lds);
}

// one work-group per listed matrix: diagonal block j of matrix d[list[w]]
__global__ __launch_bounds__(kDiagThreads) void potrf_diag_grouped_kernel(const CholDesc* __restrict__ d,
                                                                 const int* __restrict__ list, int j) {
  const CholDesc c = d[list[blockIdx.x]];
  if (*c.info != 0 || j >= c.n) return;
  const int jb = min(NB, c.n - j);
  __shared__ __align__(16) double lds[kDiagLds];
  chol::potrf_diag_body(c.A + j + (long long)j * c.lda, c.lda, j, jb, c.dinv + (long long)(j / NB) * NB * NB, c.info,
                        lds);
}

__global__ void zero_int_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

int potrf_diag_grouped(const CholDesc* d, const int* list, int count, int j, hipStream_t stream) {
  if (count <= 0) return kOk;
  hipLaunchKernelGGL(potrf_diag_grouped_kernel, dim3(count), dim3(kDiagThreads), 0, stream, d, list, j);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int potrf_f64(double* A, long long lda, long long stride, int n, int batch, double* dinv, long long dinv_stride,
              int* info, hipStream_t stream) {
  if (batch <= 0) return kOk;
  hipLaunchKernelGGL(zero_int_kernel, dim3(ceil_div(batch, 256)), dim3(256), 0, stream, info, batch);
  SDPX_HIP(hipGetLastError());
  if (n <= 0) return kOk;
  // diagonal block 0 alone; every later one is factored inside the trailing
  // update that produces it (potrf_trailing_diag)
  hipLaunchKernelGGL(potrf_diag_kernel, dim3(batch), dim3(kDiagThreads), 0, stream, A, lda, stride, 0, std::min(NB, n),
                     dinv, dinv_stride, info);
  SDPX_HIP(hipGetLastError());
  for (int j = 0; j < n; j += NB) {
    const int jb = std::min(NB, n - j);
    const int rest = n - j - jb;
    if (rest <= 0) break;
    // panel: A21 <- A21 * inv(L11)'   (in place: each work-group owns whole rows)
    GemmArgs p{};
    p.M = rest;
    p.N = jb;
    p.K = jb;
    p.alpha = 1.0;
    p.beta = 0.0;
    p.A = A + j + jb + (long long)j * lda;
    p.lda = lda;
    p.strideA = stride;
    p.B = dinv + (long long)(j / NB) * NB * NB;
    p.ldb = NB;
    p.strideB = dinv_stride;
    p.C = A + j + jb + (long long)j * lda;
    p.ldc = lda;
    p.strideC = stride;
    p.batch = batch;
    p.flags = kGemmTransB;
    p.skip = info;
    p.skip_stride = 1;
    SDPX_CHECK(gemm_f64(p, stream));
    // trailing: A22 <- A22 - A21 A21'  (lower tiles only)
    GemmArgs t{};
    t.M = rest;
    t.N = rest;
    t.K = jb;
    t.alpha = -1.0;
    t.beta = 1.0;
    t.A = A + j + jb + (long long)j * lda;
    t.lda = lda;
    t.strideA = stride;
    t.B = A + j + jb + (long long)j * lda;
    t.ldb = lda;
    t.strideB = stride;
    t.C = A + j + jb + (long long)(j + jb) * lda;
    t.ldc = lda;
    t.strideC = stride;
    t.batch = batch;
    t.flags = kGemmTransB | kGemmLowerOut;
    t.skip = info;
    t.skip_stride = 1;
    SDPX_CHECK(potrf_trailing_diag(t, j + jb, std::min(NB, rest), dinv, dinv_stride, info, stream));
  }
  return kOk;
}

namespace {

constexpr int kTrsvRows = 64;  // rows (forward) or columns of L (backward) per work-group of one step

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One block step of a triangular solve with L (lower, diagonal-block inverses
// in dinv), as a single launch: every work-group applies the block's inverse
// (or its transpose) to x_b itself, group 0 writes the solved block to y
// (x_b is only read during the launch), and each group updates its 64 rows of
// x: forward x_i -= L(i, b) y_b for the rows below the block (lanes along
// rows, coalesced); backward x_i -= L(b, i)' y_b for the rows above it (one
// wave per column of L, 64 contiguous entries reduced across the wave).  All
// global loads of a thread are issued before the first barrier.
__global__ __launch_bounds__(256) void trsv_step_kernel(const double* __restrict__ L, long long lda, int n,
                                                        const double* __restrict__ D, int j0, int jb,
                                                        double* x, double* y, int trans) {
  constexpr int kPer = NB / 4;  // entries per thread in each 64 x 64 product
  __shared__ double xb[NB];
  __shared__ double yb[NB];
  __shared__ double red[4][NB];
  const int tid = threadIdx.x, r = tid & 63, part = tid >> 6;
  double dv[kPer], lv[kPer];
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    const int c = part + 4 * t;
    dv[t] = c < jb && r < jb ? (trans ? D[c + r * NB] : D[r + c * NB]) : 0.0;
  }
  const int i_fwd = j0 + jb + blockIdx.x * kTrsvRows + r;
  const int i_bwd0 = blockIdx.x * kTrsvRows + part;
  const int i_end = min(j0, (blockIdx.x + 1) * kTrsvRows);
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    if (!trans) {
      const int c = part + 4 * t;
      lv[t] = (i_fwd < n && c < jb) ? L[i_fwd + (long long)(j0 + c) * lda] : 0.0;
    } else {
      const int i = i_bwd0 + 4 * t;
      lv[t] = (i < i_end && r < jb) ? L[j0 + r + (long long)i * lda] : 0.0;
    }
  }
  if (tid < NB) xb[tid] = tid < jb ? x[j0 + tid] : 0.0;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < kPer; ++t) s += dv[t] * xb[part + 4 * t];
  red[part][r] = s;
  __syncthreads();
  if (tid < NB) {
    const double v = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
    yb[tid] = tid < jb ? v : 0.0;
    if (blockIdx.x == 0 && tid < jb) y[j0 + tid] = v;
  }
  __syncthreads();
  if (!trans) {
    double a = 0.0;
#pragma unroll
    for (int t = 0; t < kPer; ++t) a += lv[t] * yb[part + 4 * t];
    red[part][r] = a;
    __syncthreads();
    if (part == 0 && i_fwd < n) x[i_fwd] -= (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
  } else {
    const double yl = yb[r];
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
      const int i = i_bwd0 + 4 * t;
      const double tsum = wave_sum64(lv[t] * yl);
      if (r == 0 && i < i_end) x[i] -= tsum;
    }
  }
}

}  // namespace

int trsv_f64_steps(const double* L, long long lda, int n, const double* dinv, double* x, double* y, int trans,
                   hipStream_t stream) {
  if (n <= 0) return kOk;
  const int nblk = ceil_div(n, NB);
  for (int step = 0; step < nblk; ++step) {
    const int bi = trans ? (nblk - 1 - step) : step;
    const int j0 = bi * NB;
    const int jb = std::min(NB, n - j0);
    const int rows = trans ? j0 : n - j0 - jb;
    const int groups = std::max(1, ceil_div(rows, kTrsvRows));
    hipLaunchKernelGGL(trsv_step_kernel, dim3(groups), dim3(256), 0, stream, L, lda, n,
                       dinv + (long long)bi * NB * NB, j0, jb, x, y, trans);
    SDPX_HIP(hipGetLastError());
  }
  // y now holds the solution (every block was written once by its step)
  SDPX_HIP(hipMemcpyAsync(x, y, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, stream));
  return kOk;
}

int trsv_f64(const double* L, long long lda, int n, const double* dinv, double* x, int trans, hipStream_t stream) {
  const int nblk = ceil_div(n, NB);
  for (int step = 0; step < nblk; ++step) {
    const int bi = trans ? (nblk - 1 - step) : step;
    const int j = bi * NB;
    const int jb = std::min(NB, n - j);
    // x_j <- inv(L_jj) x_j   or   inv(L_jj)' x_j   (in place, one work-group owns the column)
    GemmArgs d{};
    d.M = jb;
    d.N = 1;
    d.K = jb;
    d.alpha = 1.0;
    d.beta = 0.0;
    d.A = dinv + (long long)bi * NB * NB;
    d.lda = NB;
    d.B = x + j;
    d.ldb = n;
    d.C = x + j;
    d.ldc = n;
    d.batch = 1;
    d.flags = trans ? kGemmTransA : 0u;
    SDPX_CHECK(gemm_f64(d, stream));
    GemmArgs u{};
    u.N = 1;
    u.K = jb;
    u.alpha = -1.0;
    u.beta = 1.0;
    u.B = x + j;
    u.ldb = n;
    u.ldc = n;
    u.batch = 1;
    u.lda = lda;
    if (!trans) {
      // x_{>j} -= L_{>j, j} x_j
      u.M = n - j - jb;
      u.A = L + j + jb + (long long)j * lda;
      u.C = x + j + jb;
      u.flags = 0;
    } else {
      // x_{<j} -= L_{j, <j}' x_j
      u.M = j;
      u.A = L + j;
      u.C = x;
      u.flags = kGemmTransA;
    }
    if (u.M > 0) SDPX_CHECK(gemm_f64(u, stream));
  }
  return kOk;
}

namespace {
__global__ void identity_kernel(double* W, int n) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < (long long)n * n) W[idx] = (idx % n == idx / n) ? 1.0 : 0.0;
}
}  // namespace

int trtri_lower_f64(const double* L, long long lda, int n, const double* dinv, double* out, double* /*work*/,
                    hipStream_t stream) {
  hipLaunchKernelGGL(identity_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, stream, out, n);
  SDPX_HIP(hipGetLastError());
  const int nblk = ceil_div(n, NB);
  for (int bi = 0; bi < nblk; ++bi) {
    const int j = bi * NB;
    const int jb = std::min(NB, n - j);
    // W[j:j+jb, 0:j+jb] <- inv(L_jj) W[j:j+jb, 0:j+jb]   (in place per column tile)
    GemmArgs d{};
    d.M = jb;
    d.N = j + jb;
    d.K = jb;
    d.alpha = 1.0;
    d.beta = 0.0;
    d.A = dinv + (long long)bi * NB * NB;
    d.lda = NB;
    d.B = out + j;
    d.ldb = n;
    d.C = out + j;
    d.ldc = n;
    d.batch = 1;
    SDPX_CHECK(gemm_f64(d, stream));
    const int rest = n - j - jb;
    if (rest <= 0) break;
    // W[j+jb:, 0:j+jb] -= L[j+jb:, j:j+jb] W[j:j+jb, 0:j+jb]
    GemmArgs u{};
    u.M = rest;
    u.N = j + jb;
    u.K = jb;
    u.alpha = -1.0;
    u.beta = 1.0;
    u.A = L + j + jb + (long long)j * lda;
    u.lda = lda;
    u.B = out + j;
    u.ldb = n;
    u.C = out + j + jb;
    u.ldc = n;
    u.batch = 1;
    SDPX_CHECK(gemm_f64(u, stream));
  }
  return kOk;
}

}  // namespace sdpx
