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

__global__ __launch_bounds__(256) void potrf_diag_kernel(double* A, long long lda, long long stride, int j, int jb,
                                                         double* dinv, long long dinv_stride, int* info) {
  const int b = blockIdx.x;
  if (info[b] != 0) return;
  double* a = A + (long long)b * stride + j + (long long)j * lda;
  __shared__ double s[NB][NB + 1];
  __shared__ double dgi[NB];
  __shared__ int fail;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < jb * jb; idx += 256) {
    const int r = idx % jb, c = idx / jb;
    s[r][c] = (r >= c) ? a[r + (long long)c * lda] : 0.0;
  }
  if (tid == 0) fail = 0;
  __syncthreads();
  for (int c = 0; c < jb; ++c) {
    if (tid == 0) {
      const double p = s[c][c];
      if (!(p > 0.0)) {
        fail = 1;
        info[b] = j + c + 1;
      } else {
        s[c][c] = sqrt(p);
      }
    }
    __syncthreads();
    if (fail) return;
    const double d = s[c][c];
    for (int r = c + 1 + tid; r < jb; r += 256) s[r][c] = s[r][c] / d;
    __syncthreads();
    const int nrem = jb - c - 1;
    for (int idx = tid; idx < nrem * nrem; idx += 256) {
      const int r = c + 1 + idx % nrem, c2 = c + 1 + idx / nrem;
      if (r >= c2) s[r][c2] = s[r][c2] - s[r][c] * s[c2][c];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < jb * jb; idx += 256) {
    const int r = idx % jb, c = idx / jb;
    if (r >= c) a[r + (long long)c * lda] = s[r][c];
  }
  // Inverse of the lower factor block, one column c per thread.  inv(r, c)
  // for r > c is kept transposed in the (unused) upper triangle s[c][r].
  if (tid < jb) dgi[tid] = 1.0 / s[tid][tid];
  __syncthreads();
  if (tid < jb) {
    const int c = tid;
    for (int r = c + 1; r < jb; ++r) {
      double acc = s[r][c] * dgi[c];
      for (int q = c + 1; q < r; ++q) acc += s[r][q] * s[c][q];
      s[c][r] = -acc / s[r][r];
    }
  }
  __syncthreads();
  double* di = dinv + (long long)b * dinv_stride + (long long)(j / NB) * NB * NB;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int r = idx % NB, c = idx / NB;
    double v = 0.0;
    if (r < jb && c < jb) v = (r == c) ? dgi[c] : (r > c ? s[c][r] : 0.0);
    di[idx] = v;
  }
}

__global__ void zero_int_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

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
