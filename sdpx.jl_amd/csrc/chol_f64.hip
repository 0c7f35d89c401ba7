// Dense FP64 Cholesky (batched, right-looking, blocked) and triangular solves.
//
// The diagonal nb x nb block of every panel is factored inside one work-group
// in LDS; the kernel also inverts that block (kept in `dinv`), so that the
// panel solve, the trailing update and every later triangular solve are MFMA
// GEMMs (gemm_f64.hip).  A non-positive or non-finite pivot sets the 1-based
// LAPACK-style `info` and every later kernel of that batch entry returns
// immediately (the failure is read once, after the factorization).
#include "common.hpp"

namespace sdpx {
namespace {

constexpr int NB = kCholNb;

// Factors the jb x jb diagonal block at a (column j of its matrix), writes L
// over its lower triangle and the block's inverse to di (NB x NB, zero
// padded); a bad pivot sets *info = j + c + 1.
__device__ void potrf_diag_body(double* a, long long lda, int j, int jb, double* di, int* info) {
  __shared__ double s[NB][NB + 1];
  __shared__ double inv[NB][NB + 1];
  __shared__ double red[4][NB];
  const int tid = threadIdx.x;
  const int row = tid % NB;   // row owned by this thread in the updates
  const int part = tid / NB;  // which quarter of the columns / sum terms
  for (int idx = tid; idx < jb * jb; idx += 256) {
    const int r = idx % jb, c = idx / jb;
    s[r][c] = (r >= c) ? a[r + (long long)c * lda] : 0.0;
  }
  __syncthreads();
  // Right-looking elimination on the unscaled columns: step c subtracts
  // a_r a_c2 / p from the trailing lower triangle (one barrier per column);
  // the columns are scaled by 1/sqrt(p) afterwards, all at once.
  for (int c = 0; c < jb; ++c) {
    const double p = s[c][c];
    if (!(p > 0.0)) {
      if (tid == 0) *info = j + c + 1;
      return;  // uniform: every thread read the same pivot
    }
    if (row > c && row < jb) {
      const double f = s[row][c] / p;
      for (int c2 = c + 1 + part; c2 <= row; c2 += 4) s[row][c2] -= f * s[c2][c];
    }
    __syncthreads();
  }
  if (tid < jb) red[0][tid] = sqrt(s[tid][tid]);
  __syncthreads();
  for (int idx = tid; idx < jb * jb; idx += 256) {
    const int r = idx % jb, c = idx / jb;
    if (r > c) s[r][c] = s[r][c] / red[0][c];
  }
  __syncthreads();
  if (tid < jb) s[tid][tid] = red[0][tid];
  __syncthreads();
  for (int idx = tid; idx < jb * jb; idx += 256) {
    const int r = idx % jb, c = idx / jb;
    if (r >= c) a[r + (long long)c * lda] = s[r][c];
  }
  // Inverse of the lower factor block by forward substitution on the
  // identity, one row r per step; the four threads of a column split the sum.
  for (int r = 0; r < jb; ++r) {
    const int c = row;
    double partial = 0.0;
    if (c < r)
      for (int q = c + part; q < r; q += 4) partial += s[r][q] * inv[q][c];
    red[part][c] = partial;
    __syncthreads();
    if (part == 0 && c <= r) {
      inv[r][c] = (c == r) ? 1.0 / s[r][r] : -((red[0][c] + red[1][c]) + (red[2][c] + red[3][c])) / s[r][r];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int r = idx % NB, c = idx / NB;
    di[idx] = (r < jb && c < jb && r >= c) ? inv[r][c] : 0.0;
  }
}

__global__ __launch_bounds__(256) void potrf_diag_kernel(double* A, long long lda, long long stride, int j, int jb,
                                                         double* dinv, long long dinv_stride, int* info) {
  const int b = blockIdx.x;
  if (info[b] != 0) return;
  potrf_diag_body(A + (long long)b * stride + j + (long long)j * lda, lda, j, jb,
                  dinv + (long long)b * dinv_stride + (long long)(j / NB) * NB * NB, info + b);
}

// one work-group per listed matrix: diagonal block j of matrix d[list[w]]
__global__ __launch_bounds__(256) void potrf_diag_grouped_kernel(const CholDesc* __restrict__ d,
                                                                 const int* __restrict__ list, int j) {
  const CholDesc c = d[list[blockIdx.x]];
  if (*c.info != 0 || j >= c.n) return;
  const int jb = min(NB, c.n - j);
  potrf_diag_body(c.A + j + (long long)j * c.lda, c.lda, j, jb, c.dinv + (long long)(j / NB) * NB * NB, c.info);
}

__global__ void zero_int_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

int potrf_diag_grouped(const CholDesc* d, const int* list, int count, int j, hipStream_t stream) {
  if (count <= 0) return kOk;
  hipLaunchKernelGGL(potrf_diag_grouped_kernel, dim3(count), dim3(256), 0, stream, d, list, j);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int potrf_f64(double* A, long long lda, long long stride, int n, int batch, double* dinv, long long dinv_stride,
              int* info, hipStream_t stream) {
  if (batch <= 0) return kOk;
  hipLaunchKernelGGL(zero_int_kernel, dim3(ceil_div(batch, 256)), dim3(256), 0, stream, info, batch);
  SDPX_HIP(hipGetLastError());
  for (int j = 0; j < n; j += NB) {
    const int jb = std::min(NB, n - j);
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(batch), dim3(256), 0, stream, A, lda, stride, j, jb, dinv,
                       dinv_stride, info);
    SDPX_HIP(hipGetLastError());
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
    SDPX_CHECK(gemm_f64(t, stream));
  }
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
