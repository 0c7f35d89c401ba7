// C-ABI of the LA provider seam's host-buffer operations (la_backend.jl:1529-1769):
// la_chol!, la_trsm!, la_trsv_lower!, la_trsv_transpose!, la_syrk!, la_mul_owned!,
// la_axpby_owned!, la_dot and la_norminf on caller-owned host arrays (Float64,
// column-major with a leading dimension).  Each call uploads its operands to
// device scratch owned by the handle (grown on demand, kept between calls), runs
// the library's device kernels (MFMA GEMM, blocked Cholesky, triangular solves on
// 64 x 64 diagonal-block inverses, reductions) on the handle's stream, writes the
// result back and synchronises: the seam's calling convention.  The dense Newton
// step never goes through here (its operands stay in HBM); these entry points
// serve SDPX code that calls the provider's la_* methods directly.
#include <cmath>
#include <vector>

#include "../../include/sdpx_mi355x.h"
#include "blas_f64.hpp"

struct sdpx_la {
  int device = 0;
  hipStream_t stream = nullptr;
  double* buf[3] = {nullptr, nullptr, nullptr};
  size_t cap[3] = {0, 0, 0};
  double* dinv = nullptr;
  size_t dinv_cap = 0;
  double* tmp = nullptr;  // one 64-row block of right-hand sides
  size_t tmp_cap = 0;
  double* red = nullptr;  // reduction partials and the result
  int* dinfo = nullptr;
};

namespace sdpx {
namespace {

constexpr int NB = kCholNb;

int grow(double** p, size_t* cap, size_t count) {
  if (count <= *cap) return kOk;
  if (*p) SDPX_HIP(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  const size_t want = std::max<size_t>(count, 64);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), want * sizeof(double));
  if (e != hipSuccess) {
    set_last_error(std::string("sdpx_la: hipMalloc of ") + std::to_string(want * sizeof(double)) +
                   " bytes: " + hipGetErrorString(e));
    return kOutOfMemory;
  }
  *cap = want;
  return kOk;
}

int check_la(sdpx_la_t* h) {
  if (!h) {
    set_last_error("sdpx_la: null handle");
    return kInvalidArgument;
  }
  SDPX_HIP(hipSetDevice(h->device));
  return kOk;
}

// rows x cols block between a host array with leading dimension ld and device storage with ld rows
int upload(sdpx_la_t* h, double* dst, const double* src, long long ld, int rows, int cols) {
  if (rows <= 0 || cols <= 0) return kOk;
  SDPX_HIP(hipMemcpy2DAsync(dst, sizeof(double) * rows, src, sizeof(double) * ld, sizeof(double) * rows, cols,
                            hipMemcpyHostToDevice, h->stream));
  return kOk;
}
int download(sdpx_la_t* h, double* dst, long long ld, const double* src, int rows, int cols) {
  if (rows <= 0 || cols <= 0) return kOk;
  SDPX_HIP(hipMemcpy2DAsync(dst, sizeof(double) * ld, src, sizeof(double) * rows, sizeof(double) * rows, cols,
                            hipMemcpyDeviceToHost, h->stream));
  return kOk;
}

// Inverse of every 64 x 64 diagonal block of the lower-triangular n x n matrix L (ld n), in the
// layout potrf_f64 writes (block b at dinv + b * 64 * 64, column-major, ld 64, zero above the
// diagonal and beyond the block's last row / column).  One work-group per block; thread c solves
// L_bb z = e_c by forward substitution from the block held in LDS.
__global__ void __launch_bounds__(NB) tri_block_inverse_kernel(const double* L, long long ld, int n, double* dinv) {
  __shared__ double blk[NB][NB + 1];
  const int b = blockIdx.x, j0 = b * NB, jb = min(NB, n - j0), c = threadIdx.x;
  for (int r = 0; r < NB; ++r) blk[r][c] = (r < jb && c < jb && r >= c) ? L[(j0 + r) + (long long)(j0 + c) * ld] : 0.0;
  __syncthreads();
  double* out = dinv + (long long)b * NB * NB + (long long)c * NB;  // column c, written by this thread only
  for (int i = 0; i < NB; ++i) {
    double v = 0.0;
    if (c < jb && i < jb && i >= c) {
      v = i == c ? 1.0 : 0.0;
      for (int t = c; t < i; ++t) v -= blk[i][t] * out[t];
      v /= blk[i][i];
    }
    out[i] = v;
  }
}

int block_inverses(sdpx_la_t* h, const double* L, int n) {
  const int nblk = ceil_div(n, NB);
  SDPX_CHECK(grow(&h->dinv, &h->dinv_cap, (size_t)nblk * NB * NB));
  hipLaunchKernelGGL(tri_block_inverse_kernel, dim3(nblk), dim3(NB), 0, h->stream, L, (long long)n, n, h->dinv);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

__global__ void copy_rows_kernel(const double* src, long long lds, double* dst, long long ldd, int rows, int cols) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)rows * cols) return;
  const int r = (int)(idx % rows), c = (int)(idx / rows);
  dst[r + c * ldd] = src[r + c * lds];
}

GemmArgs gemm_args(int M, int N, int K, double alpha, const double* A, long long lda, const double* B, long long ldb,
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

// the finite guard of la_chol! (la_backend.jl:1621-1631): every lower entry finite
bool lower_finite(const double* A, long long ld, int n) {
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i)
      if (!std::isfinite(A[i + (long long)j * ld])) return false;
  return true;
}

}  // namespace
}  // namespace sdpx

using namespace sdpx;

extern "C" {

int sdpx_la_create(int device, sdpx_la_t** out) {
  if (!out) {
    set_last_error("sdpx_la_create: out required");
    return kInvalidArgument;
  }
  *out = nullptr;
  int count = 0;
  SDPX_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    set_last_error("sdpx_la_create: no device " + std::to_string(device));
    return kInvalidArgument;
  }
  SDPX_HIP(hipSetDevice(device));
  auto* h = new sdpx_la_t();
  h->device = device;
  if (hipStreamCreate(&h->stream) != hipSuccess || hipMalloc(reinterpret_cast<void**>(&h->red), 520 * 8) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&h->dinfo), sizeof(int)) != hipSuccess) {
    set_last_error("sdpx_la_create: stream or scratch allocation failed");
    sdpx_la_destroy(h);
    return kHipError;
  }
  *out = h;
  return kOk;
}

int sdpx_la_destroy(sdpx_la_t* h) {
  if (!h) return kOk;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (double* p : h->buf)
    if (p) hipFree(p);
  if (h->dinv) hipFree(h->dinv);
  if (h->tmp) hipFree(h->tmp);
  if (h->red) hipFree(h->red);
  if (h->dinfo) hipFree(h->dinfo);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return kOk;
}

int sdpx_la_chol(sdpx_la_t* h, int n, double* A, long long lda) {
  SDPX_CHECK(check_la(h));
  if (n < 0 || (n > 0 && (!A || lda < n))) {
    set_last_error("sdpx_la_chol: A (lda >= n) required");
    return kInvalidArgument;
  }
  if (n == 0) return kOk;
  if (!lower_finite(A, lda, n)) return kNotPositiveDefinite;
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)n * n));
  SDPX_CHECK(grow(&h->dinv, &h->dinv_cap, (size_t)ceil_div(n, NB) * NB * NB));
  SDPX_CHECK(upload(h, h->buf[0], A, lda, n, n));
  SDPX_CHECK(potrf_f64(h->buf[0], n, 0, n, 1, h->dinv, 0, h->dinfo, h->stream));
  int info = 0;
  SDPX_HIP(hipMemcpyAsync(&info, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  if (info != 0) return kNotPositiveDefinite;
  std::vector<double> L((size_t)n * n);
  SDPX_CHECK(download(h, L.data(), n, h->buf[0], n, n));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  if (!lower_finite(L.data(), n, n)) return kNotPositiveDefinite;
  // cholesky!(Symmetric(A, :L)): the factor replaces the lower triangle, the upper one is kept
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) A[i + (long long)j * lda] = L[i + (size_t)j * n];
  return kOk;
}

int sdpx_la_trsm(sdpx_la_t* h, int n, int nrhs, const double* L, long long ldl, double* X, long long ldx) {
  SDPX_CHECK(check_la(h));
  if (n < 0 || nrhs < 0 || (n > 0 && (!L || ldl < n)) || (n > 0 && nrhs > 0 && (!X || ldx < n))) {
    set_last_error("sdpx_la_trsm: L (ldl >= n) and X (ldx >= n) required");
    return kInvalidArgument;
  }
  if (n == 0 || nrhs == 0) return kOk;
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)n * n));
  SDPX_CHECK(grow(&h->buf[1], &h->cap[1], (size_t)n * nrhs));
  SDPX_CHECK(grow(&h->tmp, &h->tmp_cap, (size_t)NB * nrhs));
  SDPX_CHECK(upload(h, h->buf[0], L, ldl, n, n));
  SDPX_CHECK(upload(h, h->buf[1], X, ldx, n, nrhs));
  SDPX_CHECK(block_inverses(h, h->buf[0], n));
  // blocked forward substitution: X_j = D_j (X_j - L(j, <j) X_<j), D_j the block inverse
  double* dL = h->buf[0];
  double* dX = h->buf[1];
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = std::min(NB, n - j0);
    if (j0 > 0)
      SDPX_CHECK(gemm_f64(gemm_args(jb, nrhs, j0, -1.0, dL + j0, n, dX, n, 1.0, dX + j0, n, 0), h->stream));
    SDPX_CHECK(gemm_f64(gemm_args(jb, nrhs, jb, 1.0, h->dinv + (long long)(j0 / NB) * NB * NB, NB, dX + j0, n, 0.0,
                                  h->tmp, NB, kGemmALower),
                        h->stream));
    hipLaunchKernelGGL(copy_rows_kernel, dim3(ceil_div((long long)jb * nrhs, 256)), dim3(256), 0, h->stream, h->tmp,
                       (long long)NB, dX + j0, (long long)n, jb, nrhs);
    SDPX_HIP(hipGetLastError());
  }
  SDPX_CHECK(download(h, X, ldx, dX, n, nrhs));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_la_trsv(sdpx_la_t* h, int n, const double* L, long long ldl, double* x, int trans) {
  SDPX_CHECK(check_la(h));
  if (n < 0 || (n > 0 && (!L || !x || ldl < n)) || (trans != 0 && trans != 1)) {
    set_last_error("sdpx_la_trsv: L (ldl >= n), x and trans 0 / 1 required");
    return kInvalidArgument;
  }
  if (n == 0) return kOk;
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)n * n));
  SDPX_CHECK(grow(&h->buf[1], &h->cap[1], (size_t)2 * n));
  SDPX_CHECK(upload(h, h->buf[0], L, ldl, n, n));
  SDPX_HIP(hipMemcpyAsync(h->buf[1], x, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  SDPX_CHECK(block_inverses(h, h->buf[0], n));
  SDPX_CHECK(trsv_f64_steps(h->buf[0], n, n, h->dinv, h->buf[1], h->buf[1] + n, trans, h->stream));
  SDPX_HIP(hipMemcpyAsync(x, h->buf[1], sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_la_syrk(sdpx_la_t* h, int n, int k, double alpha, const double* P, long long ldp, double beta, double* S,
                 long long lds) {
  SDPX_CHECK(check_la(h));
  if (n < 0 || k < 0 || (n > 0 && (!S || lds < n)) || (n > 0 && k > 0 && (!P || ldp < k))) {
    set_last_error("sdpx_la_syrk: S (lds >= n) and P (k x n, ldp >= k) required");
    return kInvalidArgument;
  }
  if (n == 0) return kOk;
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)std::max(k, 1) * n));
  SDPX_CHECK(grow(&h->buf[1], &h->cap[1], (size_t)n * n));
  SDPX_CHECK(upload(h, h->buf[0], P, ldp, k, n));
  if (beta != 0.0) SDPX_CHECK(upload(h, h->buf[1], S, lds, n, n));
  SDPX_CHECK(gemm_f64(gemm_args(n, n, k, alpha, h->buf[0], std::max(k, 1), h->buf[0], std::max(k, 1), beta, h->buf[1], n,
                                kGemmTransA | kGemmLowerOut),
                      h->stream));
  std::vector<double> out((size_t)n * n);
  SDPX_CHECK(download(h, out.data(), n, h->buf[1], n, n));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  // BLAS syrk('L', 'T'): the lower triangle is written, the upper one is kept
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) S[i + (long long)j * lds] = out[i + (size_t)j * n];
  return kOk;
}

int sdpx_la_gemm(sdpx_la_t* h, int ta, int tb, int m, int n, int k, double alpha, const double* A, long long lda,
                 const double* B, long long ldb, double beta, double* C, long long ldc) {
  SDPX_CHECK(check_la(h));
  const int ar = ta ? k : m, ac = ta ? m : k, br = tb ? n : k, bc = tb ? k : n;
  if (m < 0 || n < 0 || k < 0 || (m > 0 && n > 0 && (!C || ldc < m)) ||
      (k > 0 && m > 0 && (!A || lda < ar)) || (k > 0 && n > 0 && (!B || ldb < br))) {
    set_last_error("sdpx_la_gemm: A, B, C with leading dimensions of op's storage required");
    return kInvalidArgument;
  }
  if (m == 0 || n == 0) return kOk;
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)std::max(ar, 1) * std::max(ac, 1)));
  SDPX_CHECK(grow(&h->buf[1], &h->cap[1], (size_t)std::max(br, 1) * std::max(bc, 1)));
  SDPX_CHECK(grow(&h->buf[2], &h->cap[2], (size_t)m * n));
  if (k > 0) {
    SDPX_CHECK(upload(h, h->buf[0], A, lda, ar, ac));
    SDPX_CHECK(upload(h, h->buf[1], B, ldb, br, bc));
  }
  if (beta != 0.0) SDPX_CHECK(upload(h, h->buf[2], C, ldc, m, n));
  SDPX_CHECK(gemm_f64(gemm_args(m, n, k, alpha, h->buf[0], std::max(ar, 1), h->buf[1], std::max(br, 1), beta, h->buf[2],
                                m, (ta ? kGemmTransA : 0u) | (tb ? kGemmTransB : 0u)),
                      h->stream));
  SDPX_CHECK(download(h, C, ldc, h->buf[2], m, n));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_la_axpby(sdpx_la_t* h, long long n, double alpha, const double* x, double beta, double* y) {
  SDPX_CHECK(check_la(h));
  if (n < 0 || (n > 0 && (!x || !y))) {
    set_last_error("sdpx_la_axpby: x and y required");
    return kInvalidArgument;
  }
  if (n == 0) return kOk;
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)2 * n));
  SDPX_HIP(hipMemcpyAsync(h->buf[0], x, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->buf[0] + n, y, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  SDPX_CHECK(axpbyz_f64(n, alpha, h->buf[0], beta, h->buf[0] + n, h->buf[0] + n, h->stream));
  SDPX_HIP(hipMemcpyAsync(y, h->buf[0] + n, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_la_dot(sdpx_la_t* h, long long n, const double* x, const double* y, double* out) {
  SDPX_CHECK(check_la(h));
  if (n < 0 || !out || (n > 0 && (!x || !y))) {
    set_last_error("sdpx_la_dot: x, y and out required");
    return kInvalidArgument;
  }
  if (n == 0) {
    *out = 0.0;
    return kOk;
  }
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)2 * n));
  SDPX_HIP(hipMemcpyAsync(h->buf[0], x, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  SDPX_HIP(hipMemcpyAsync(h->buf[0] + n, y, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  SDPX_CHECK(dot_f64(h->buf[0], nullptr, 0.0, h->buf[0] + n, nullptr, 0.0, n, h->red, h->red + 512, h->stream));
  SDPX_HIP(hipMemcpyAsync(out, h->red + 512, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int sdpx_la_norminf(sdpx_la_t* h, long long n, const double* x, double* out) {
  SDPX_CHECK(check_la(h));
  if (n < 0 || !out || (n > 0 && !x)) {
    set_last_error("sdpx_la_norminf: x and out required");
    return kInvalidArgument;
  }
  if (n == 0) {
    *out = 0.0;  // la_norminf of an empty vector (la_backend.jl:1537)
    return kOk;
  }
  SDPX_CHECK(grow(&h->buf[0], &h->cap[0], (size_t)n));
  SDPX_HIP(hipMemcpyAsync(h->buf[0], x, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  SDPX_CHECK(absmax_f64(h->buf[0], n, h->red, h->red + 512, h->stream));
  SDPX_HIP(hipMemcpyAsync(out, h->red + 512, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  SDPX_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

}  // extern "C"
