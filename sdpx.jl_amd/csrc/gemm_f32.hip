// FP32 Schur Gram product on CDNA4 matrix cores (v_mfma_f32_16x16x4_f32) for
// the mixed-precision route: S = Phat' Phat over lower 128 x 128 tiles, Phat
// in FP32 (columns K-contiguous), accumulated in FP32 per K range and written
// to FP64 S (or to FP64 split-K partials reduced in a fixed order).
//
// Same organisation as the FP64 kernel (gemm_f64.hip): 256 threads = 2 x 2
// waves of 64 x 64, BK = 16 slices double-buffered through LDS, the next slice
// prefetched into registers (16-byte loads), operands fed swapped so the
// accumulator holds C' and stores are contiguous, split-K work items ordered so
// each XCD owns whole K ranges.  The FP32 MFMA accumulator layout is
// D[4 (lane / 16) + r][lane % 16] in register r.
#include "common.hpp"

namespace sdpx {
namespace {

constexpr int BM = 128;
constexpr int BK = 16;
constexpr int PADF = 4;  // LDS row pitch BM + 4 floats

typedef float float4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256, 2) void schur_gram_f32_kernel(const float* __restrict__ P, long long ldp, int m,
                                                                int K, double* S, long long lds, double beta,
                                                                int split, double* ws, const int2* tile_order) {
  __shared__ float As[2][BK][BM + PADF];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;

  // XCD-aware order of the (split, tile) work items (see gemm_f64.hip)
  const int total = gridDim.x;
  const int bid = blockIdx.x;
  const int xcd = bid % kNumXcd, slot = bid / kNumXcd;
  const int q = total / kNumXcd, rq = total % kNumXcd;
  const int lin = xcd * q + (xcd < rq ? xcd : rq) + slot;
  const int per_split = total / split;
  const int s = lin / per_split;
  const int tile = lin % per_split;
  int ti, tj;
  if (tile_order) {
    ti = tile_order[tile].x;
    tj = tile_order[tile].y;
  } else {
    int t = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
    while ((long long)(t + 1) * (t + 2) / 2 <= tile) ++t;
    while ((long long)t * (t + 1) / 2 > tile) --t;
    ti = t;
    tj = tile - t * (t + 1) / 2;
  }
  const int m0 = ti * BM, n0 = tj * BM;
  const bool diag = ti == tj;  // the B tile equals the A tile: one staged copy serves both

  int kBegin = 0, kEnd = K;
  if (split > 1) {
    const int stages = (K + BK - 1) / BK;
    const int per = (stages + split - 1) / split;
    kBegin = s * per * BK;
    kEnd = min(K, kBegin + per * BK);
  }

  // Thread loads: rows (tid / 4) and (tid / 4 + 64) of each operand tile, k = 4 (tid % 4) .. +3.
  const int lr = tid >> 2, lk = (tid & 3) * 4;
  const long long ra0 = (long long)min(m0 + lr, m - 1) * ldp, ra1 = (long long)min(m0 + lr + 64, m - 1) * ldp;
  const long long rb0 = (long long)min(n0 + lr, m - 1) * ldp, rb1 = (long long)min(n0 + lr + 64, m - 1) * ldp;
  __shared__ float Bs[2][BK][BM + PADF];

  float4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  float4_t va0, va1, vb0, vb1;
  auto load = [&](int k0) {
    const int kk = k0 + lk;
    if (kk + 4 <= kEnd) {
      va0 = *reinterpret_cast<const float4_t*>(P + ra0 + kk);
      va1 = *reinterpret_cast<const float4_t*>(P + ra1 + kk);
      vb0 = *reinterpret_cast<const float4_t*>(P + rb0 + kk);
      vb1 = *reinterpret_cast<const float4_t*>(P + rb1 + kk);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool in = kk + e < kEnd;
        va0[e] = in ? P[ra0 + kk + e] : 0.f;
        va1[e] = in ? P[ra1 + kk + e] : 0.f;
        vb0[e] = in ? P[rb0 + kk + e] : 0.f;
        vb1[e] = in ? P[rb1 + kk + e] : 0.f;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      As[buf][lk + e][lr] = va0[e];
      As[buf][lk + e][lr + 64] = va1[e];
      Bs[buf][lk + e][lr] = vb0[e];
      Bs[buf][lk + e][lr + 64] = vb1[e];
    }
  };
  auto compute = [&](int cur) {
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      float fa[4], fb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) fa[i] = As[cur][kr][wm0 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) fb[j] = Bs[cur][kr][wn0 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j], fa[i], acc[i][j], 0, 0, 0);
    }
  };
  (void)diag;
  if (kBegin < kEnd) {
    load(kBegin);
    store(0);
    __syncthreads();
    int cur = 0;
    int k0 = kBegin;
    for (; k0 + BK < kEnd; k0 += BK) {
      load(k0 + BK);
      compute(cur);
      store(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
    compute(cur);
  }
  // acc[i][j][r] = C(row = m0 + wm0 + 16 i + lane % 16, col = n0 + wn0 + 16 j + 4 (lane / 16) + r)
  const bool part = split > 1;
  double* out = part ? ws + (long long)s * m * m : S;
  const long long ldo = part ? (long long)m : lds;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = m0 + wm0 + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + 4 * (lane >> 4) + r;
        if (row < m && col < m && row >= col) {
          double* cp = out + row + (long long)col * ldo;
          const double v = (double)acc[i][j][r];
          *cp = (part || beta == 0.0) ? v : v + beta * (*cp);
        }
      }
  }
}

// S (lower) = sum_s W[s] + beta S, split order
__global__ __launch_bounds__(256) void gram_reduce_kernel(const double* ws, int split, int m, double* S,
                                                          long long lds, double beta) {
  const long long mm = (long long)m * m;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < mm; idx += (long long)gridDim.x * 256) {
    const int r = (int)(idx % m), c = (int)(idx / m);
    if (r < c) continue;
    double sum = 0.0;
    for (int s = 0; s < split; ++s) sum += ws[s * mm + idx];
    double* cp = S + r + (long long)c * lds;
    *cp = beta == 0.0 ? sum : sum + beta * (*cp);
  }
}

__global__ void panel_f32_kernel(const double* __restrict__ src, long long ld_src, float* __restrict__ dst,
                                 long long ld_dst, long long rows, int cols) {
  const long long n = ld_dst * cols;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const long long r = i % ld_dst, c = i / ld_dst;
    dst[i] = r < rows ? (float)src[r + c * ld_src] : 0.f;
  }
}

}  // namespace

int gram_f32(const float* P, long long ldp, int m, int K, double* S, long long lds, double beta, int split,
             double* ws, const int2* tile_order, hipStream_t stream, hipEvent_t after_main) {
  if (m <= 0) return kOk;
  if ((ldp & 3) != 0 || (reinterpret_cast<uintptr_t>(P) & 15) != 0) {
    set_last_error("gram_f32: the FP32 panel needs a leading dimension divisible by 4 and a 16-byte aligned base");
    return kInvalidArgument;
  }
  if (split < 1) split = 1;
  const int nt = ceil_div(m, BM);
  const long long tiles = (long long)nt * (nt + 1) / 2;
  hipLaunchKernelGGL(schur_gram_f32_kernel, dim3((unsigned)(tiles * split)), dim3(256), 0, stream, P, ldp, m, K, S,
                     lds, beta, split, ws, tile_order);
  SDPX_HIP(hipGetLastError());
  if (after_main) SDPX_HIP(hipEventRecord(after_main, stream));
  if (split > 1) {
    const long long mm = (long long)m * m;
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)std::min<long long>(ceil_div(mm, 256), 4096)), dim3(256), 0,
                       stream, ws, split, m, S, lds, beta);
    SDPX_HIP(hipGetLastError());
  }
  return kOk;
}

int convert_panel_f32(const double* src, long long ld_src, float* dst, long long ld_dst, long long rows, int cols,
                      hipStream_t stream) {
  const long long n = ld_dst * cols;
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(panel_f32_kernel, dim3((unsigned)std::min<long long>(ceil_div(n, 256), 65536)), dim3(256), 0,
                     stream, src, ld_src, dst, ld_dst, rows, cols);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

}  // namespace sdpx
