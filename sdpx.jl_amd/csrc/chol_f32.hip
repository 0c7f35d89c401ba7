// FP32 Cholesky of the Schur matrix for the mixed-precision route (the
// reduced-precision factor of kernels/mixed_precision_kkt.jl:972-1111 with
// FP64 as the target precision): blocked right-looking, nb = 64, with the
// panel products L21 = A21 L11^-T and the trailing updates A22 -= L21 L21'
// (the m^3/3 of the factor) on FP32 MFMA.  The 64 x 64 diagonal blocks are
// factored by the FP64 diagonal-block code of chol_diag.hpp on an FP64 copy
// and rounded back: that chain is latency-bound, FP32 arithmetic would not
// shorten it, and the FP64 block inverses it leaves are the ones the
// triangular solves use.  The factor is returned in FP64 storage (values
// rounded to FP32 by the trailing updates), so the FP64 triangular solves,
// condition estimates and refinement work on it unchanged.
#include "chol_diag.hpp"
#include "common.hpp"

namespace sdpx {
namespace {

using chol::kDiagLds;
using chol::kDiagThreads;
constexpr int NB = kCholNb;
constexpr int TPITCH = NB + 4;  // LDS row pitch of the GEMM tiles (floats)

typedef float float4_t __attribute__((ext_vector_type(4)));

// dst (FP32, ld ldf) = lower(src) (FP64, ld n), diagonal += shift max(|src_ii|, 1)
__global__ void lower_to_f32_kernel(const double* __restrict__ src, int n, double shift, float* __restrict__ dst,
                                    long long ldf) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r < c) return;
  double v = src[idx];
  if (r == c && shift != 0.0) v += shift * fmax(fabs(v), 1.0);
  dst[r + (long long)c * ldf] = (float)v;
}

// dst (FP64, ld n) = lower(src) (FP32, ld ldf)
__global__ void lower_to_f64_kernel(const float* __restrict__ src, long long ldf, int n, double* __restrict__ dst) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r < c) return;
  dst[idx] = (double)src[r + (long long)c * ldf];
}

// Diagonal block j (jb x jb) of the FP32 matrix: copied to the FP64 block blk (64 x 64, ld 64),
// factored there (L and its inverse into dinv, FP64), L rounded back into A and the inverse into
// dinv32 (64 x 64, ld 64, zero padded) for the FP32 panel product.
__global__ __launch_bounds__(kDiagThreads) void potrf_diag_f32_kernel(float* A, long long lda, int j, int jb,
                                                                      double* blk, double* dinv, float* dinv32,
                                                                      int* info) {
  if (*info != 0) return;
  __shared__ __align__(16) double lds[kDiagLds];
  const int tid = threadIdx.x;
  float* a = A + j + (long long)j * lda;
  for (int e = tid; e < NB * NB; e += kDiagThreads) {
    const int r = e % NB, c = e / NB;
    blk[e] = (r < jb && c < jb && r >= c) ? (double)a[r + (long long)c * lda] : 0.0;
  }
  __syncthreads();  // the block's FP64 copy is read by other threads of the group
  double* di = dinv + (long long)(j / NB) * NB * NB;
  chol::potrf_diag_body(blk, NB, j, jb, di, info, lds);
  __syncthreads();
  if (*info != 0) return;
  for (int e = tid; e < NB * NB; e += kDiagThreads) {
    const int r = e % NB, c = e / NB;
    if (r < jb && c < jb && r >= c) a[r + (long long)c * lda] = (float)blk[e];
    dinv32[e] = (float)di[e];
  }
}

// C = alpha A B' + beta C on 64 x 64 tiles, FP32, K <= 64: A is M x K (ld lda), B is N x K
// (ld ldb), C is M x N (ld ldc); `lower`: only tiles ti >= tj and entries row >= col.  The whole K
// range of both operand tiles is staged once (K-major, zero padded to a multiple of 4); 4 waves of
// 32 x 32 (2 x 2 blocks of v_mfma_f32_16x16x4_f32).  A work-group owns whole rows of its tiles,
// so the panel product may run in place (C == A with N <= 64).
__global__ __launch_bounds__(256) void gemm_nt_f32_kernel(const float* A, long long lda, const float* B, long long ldb,
                                                          float* C, long long ldc, int M, int N, int K, float alpha,
                                                          float beta, int lower, const int* info) {
  if (*info != 0) return;
  __shared__ float As[NB][TPITCH], Bs[NB][TPITCH];
  int ti, tj;
  if (lower) {
    const int tile = blockIdx.x;
    int t = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= tile) ++t;
    while (t * (t + 1) / 2 > tile) --t;
    ti = t;
    tj = tile - t * (t + 1) / 2;
  } else {
    const int mt = (M + NB - 1) / NB;
    ti = blockIdx.x % mt;
    tj = blockIdx.x / mt;
  }
  const int i0 = ti * NB, j0 = tj * NB;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 32, wn0 = (wave >> 1) * 32;
  const int Kp = (K + 3) & ~3;
  // each thread stages 4 consecutive rows of one K column of each operand per pass
  const int r4 = (tid & 15) * 4;
  for (int k = tid >> 4; k < Kp; k += 16) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int ra = i0 + r4 + e, rb = j0 + r4 + e;
      As[k][r4 + e] = (k < K && ra < M) ? A[ra + (long long)k * lda] : 0.f;
      Bs[k][r4 + e] = (k < K && rb < N) ? B[rb + (long long)k * ldb] : 0.f;
    }
  }
  __syncthreads();
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) acc[i][jj] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kp; k0 += 4) {
    const int kk = k0 + (lane >> 4);
    float fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) fa[i] = As[kk][wm0 + 16 * i + (lane & 15)];
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) fb[jj] = Bs[kk][wn0 + 16 * jj + (lane & 15)];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
        acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[jj], fa[i], acc[i][jj], 0, 0, 0);
  }
  __syncthreads();  // in place: every operand read of the group is done before the first store
  // acc[i][jj][r] = C(row = i0 + wm0 + 16 i + lane % 16, col = j0 + wn0 + 16 jj + 4 (lane / 16) + r)
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int row = i0 + wm0 + 16 * i + (lane & 15);
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = j0 + wn0 + 16 * jj + 4 * (lane >> 4) + r;
        if (row < M && col < N && (!lower || row >= col)) {
          float* cp = C + row + (long long)col * ldc;
          const float v = alpha * acc[i][jj][r];
          *cp = beta == 0.f ? v : v + beta * (*cp);
        }
      }
  }
}

__global__ void zero_info_kernel(int* info) { *info = 0; }

}  // namespace

int lower_to_f32(const double* src, int n, double shift, float* dst, long long ldf, hipStream_t stream) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(lower_to_f32_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, stream, src, n, shift,
                     dst, ldf);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int lower_to_f64(const float* src, long long ldf, int n, double* dst, hipStream_t stream) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(lower_to_f64_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, stream, src, ldf, n,
                     dst);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int potrf_f32(float* A, long long lda, int n, double* blk, double* dinv, float* dinv32, int* info,
              hipStream_t stream) {
  hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, stream, info);
  SDPX_HIP(hipGetLastError());
  for (int j = 0; j < n; j += NB) {
    const int jb = std::min(NB, n - j);
    hipLaunchKernelGGL(potrf_diag_f32_kernel, dim3(1), dim3(kDiagThreads), 0, stream, A, lda, j, jb, blk, dinv,
                       dinv32, info);
    SDPX_HIP(hipGetLastError());
    const int rest = n - j - jb;
    if (rest <= 0) break;
    float* A21 = A + (j + jb) + (long long)j * lda;
    float* A22 = A + (j + jb) + (long long)(j + jb) * lda;
    // panel, in place: A21 <- A21 inv(L11)'
    hipLaunchKernelGGL(gemm_nt_f32_kernel, dim3(ceil_div(rest, NB)), dim3(256), 0, stream, A21, lda, dinv32,
                       (long long)NB, A21, lda, rest, jb, jb, 1.0f, 0.0f, 0, info);
    SDPX_HIP(hipGetLastError());
    // trailing: A22 <- A22 - A21 A21' (lower tiles)
    const int nt = ceil_div(rest, NB);
    hipLaunchKernelGGL(gemm_nt_f32_kernel, dim3(nt * (nt + 1) / 2), dim3(256), 0, stream, A21, lda, A21, lda, A22, lda,
                       rest, rest, jb, -1.0f, 1.0f, 1, info);
    SDPX_HIP(hipGetLastError());
  }
  return kOk;
}

}  // namespace sdpx
