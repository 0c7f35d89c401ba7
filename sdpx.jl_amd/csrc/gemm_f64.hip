// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64), column-major.
//
// One kernel template serves every dense FP64 product of the Newton step:
//   * the Schur Gram product S = Phat' Phat (lower tiles only, split-K),
//   * Phat panels L_X^-1 [A_1 ... A_m] and T_i M_Y (triangular operands),
//   * the right-looking Cholesky panel / trailing updates,
//   * k x k block products (X Y, P Y, dX dY, X^-1 M).
//
// Work-group tile BM x BN, 256 threads = 4 waves in a 2 x 2 grid, each wave
// owning a (BM/2) x (BN/2) sub-tile built from 16 x 16 MFMA blocks.  K is
// staged through LDS in slices of BK = 16 with a register prefetch of the
// next slice.  The MFMA is fed with (B-fragment, A-fragment) so the
// accumulator holds C transposed: lane%16 indexes a row of C, which makes
// the column-major epilogue stores 128-byte contiguous.
#include "common.hpp"

namespace sdpx {
namespace {

constexpr int BK = 16;
constexpr int PAD = 16;  // row stride = BM + 16 doubles: conflict-free 64-bit fragment reads

typedef double double4_t __attribute__((ext_vector_type(4)));

template <int BM, int BN, bool TA, bool TB>
__device__ __forceinline__ void gemm_f64_body(const GemmArgs& g, int split, double* ws) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int A_ELEMS = BM * BK / 256;
  constexpr int B_ELEMS = BN * BK / 256;
  __shared__ double As[BK][BM + PAD];
  __shared__ double Bs[BK][BN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;

  const int bz = blockIdx.y;
  const int b = bz / split;
  const int s = bz % split;
  if (g.skip && g.skip[(long long)b * g.skip_stride] != 0) return;

  // XCD-aware remap: blocks are dispatched round-robin over the 8 XCDs, so
  // give each XCD a contiguous run of tiles (shared A/B panels stay in its L2).
  const int total = gridDim.x;
  const int bid = blockIdx.x;
  const int xcd = bid % kNumXcd;
  const int slot = bid / kNumXcd;
  const int q = total / kNumXcd, r = total % kNumXcd;
  const int tile = xcd * q + (xcd < r ? xcd : r) + slot;

  int ti, tj;
  const bool lower_out = (g.flags & kGemmLowerOut) != 0;
  if (lower_out) {
    // linear index over lower tiles, row-major: tile = ti*(ti+1)/2 + tj
    int t = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
    while ((long long)(t + 1) * (t + 2) / 2 <= tile) ++t;
    while ((long long)t * (t + 1) / 2 > tile) --t;
    ti = t;
    tj = tile - t * (t + 1) / 2;
  } else {
    const int mt = (g.M + BM - 1) / BM;
    ti = tile % mt;
    tj = tile / mt;
  }
  const int m0 = ti * BM;
  const int n0 = tj * BN;

  const double* A = g.A + (long long)b * g.strideA;
  const double* B = g.B + (long long)b * g.strideB;

  int kBegin = 0, kEnd = g.K;
  if (g.flags & kGemmALower) kEnd = min(kEnd, m0 + BM);
  if (g.flags & kGemmBLower) kBegin = max(kBegin, n0);
  if (g.flags & kGemmAUpper) kBegin = max(kBegin, m0);
  kBegin = (kBegin / BK) * BK;
  if (split > 1) {
    const int stages = (kEnd - kBegin + BK - 1) / BK;
    const int per = (stages + split - 1) / split;
    const int sb = kBegin + s * per * BK;
    const int se = min(kEnd, sb + per * BK);
    kBegin = sb;
    kEnd = se;
  }

  double4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  double ra[A_ELEMS], rb[B_ELEMS];

  auto load_a = [&](int k0) {
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      int i, kk;
      if (!TA) {
        i = tid % BM;
        kk = tid / BM + e * (256 / BM);
      } else {
        kk = tid % BK;
        i = tid / BK + e * (256 / BK);
      }
      const int gi = m0 + i, gk = k0 + kk;
      bool ok = gi < g.M && gk < kEnd && gk >= kBegin;
      if ((g.flags & kGemmALower) && gk > gi) ok = false;
      if ((g.flags & kGemmAUpper) && gk < gi) ok = false;
      double v = 0.0;
      if (ok) v = TA ? A[gk + (long long)gi * g.lda] : A[gi + (long long)gk * g.lda];
      ra[e] = v;
    }
  };
  auto store_a = [&]() {
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      int i, kk;
      if (!TA) {
        i = tid % BM;
        kk = tid / BM + e * (256 / BM);
      } else {
        kk = tid % BK;
        i = tid / BK + e * (256 / BK);
      }
      As[kk][i] = ra[e];
    }
  };
  auto load_b = [&](int k0) {
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      int j, kk;
      if (TB) {
        j = tid % BN;
        kk = tid / BN + e * (256 / BN);
      } else {
        kk = tid % BK;
        j = tid / BK + e * (256 / BK);
      }
      const int gj = n0 + j, gk = k0 + kk;
      bool ok = gj < g.N && gk < kEnd && gk >= kBegin;
      if ((g.flags & kGemmBLower) && gk < gj) ok = false;
      double v = 0.0;
      if (ok) v = TB ? B[gj + (long long)gk * g.ldb] : B[gk + (long long)gj * g.ldb];
      rb[e] = v;
    }
  };
  auto store_b = [&]() {
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      int j, kk;
      if (TB) {
        j = tid % BN;
        kk = tid / BN + e * (256 / BN);
      } else {
        kk = tid % BK;
        j = tid / BK + e * (256 / BK);
      }
      Bs[kk][j] = rb[e];
    }
  };

  if (kBegin < kEnd) {
    load_a(kBegin);
    load_b(kBegin);
    store_a();
    store_b();
    __syncthreads();
    for (int k0 = kBegin; k0 < kEnd; k0 += BK) {
      const bool more = k0 + BK < kEnd;
      if (more) {
        load_a(k0 + BK);
        load_b(k0 + BK);
      }
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        double fa[FM], fb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = As[kr][wm0 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = Bs[kr][wn0 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i], acc[i][j], 0, 0, 0);
      }
      __syncthreads();
      if (more) {
        store_a();
        store_b();
        __syncthreads();
      }
    }
  }

  // Epilogue.  The f64 16x16x4 accumulator holds D[lane/16 + 4r][lane%16] in
  // register r of lane `lane`; with the operands swapped D = C', so
  // acc[i][j][r] = C(row = m0+wm0+16i+lane%16, col = n0+wn0+16j+lane/16+4r).
  double* C = g.C + (long long)b * g.strideC;
  double* W = (split > 1) ? ws + ((long long)b * split + s) * (long long)g.M * g.N : nullptr;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = m0 + wm0 + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int col = n0 + wn0 + j * 16 + (lane >> 4) + 4 * rr;
        if (row < g.M && col < g.N && (!lower_out || row >= col)) {
          const double v = acc[i][j][rr];
          if (W) {
            W[row + (long long)col * g.M] = g.alpha * v;
          } else {
            double* cp = C + row + (long long)col * g.ldc;
            *cp = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * (*cp);
          }
        }
      }
    }
  }
}

// C = sum_s W[s] + beta*C over the (lower) entries, summed in split order.
__global__ __launch_bounds__(256) void splitk_reduce_kernel(GemmArgs g, int split, const double* ws) {
  const long long MN = (long long)g.M * g.N;
  const int b = blockIdx.y;
  if (g.skip && g.skip[(long long)b * g.skip_stride] != 0) return;
  const bool lower_out = (g.flags & kGemmLowerOut) != 0;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < MN;
       idx += (long long)gridDim.x * blockDim.x) {
    const int row = (int)(idx % g.M);
    const int col = (int)(idx / g.M);
    if (lower_out && row < col) continue;
    const double* w = ws + (long long)b * split * MN + idx;
    double sum = 0.0;
    for (int s = 0; s < split; ++s) sum += w[s * MN];
    double* cp = g.C + (long long)b * g.strideC + row + (long long)col * g.ldc;
    *cp = (g.beta == 0.0) ? sum : sum + g.beta * (*cp);
  }
}

template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmArgs g, int split, double* ws) {
  gemm_f64_body<BM, BN, TA, TB>(g, split, ws);
}

// The Schur Gram product S = Phat' Phat (lower tiles) under its own symbol,
// so profiles attribute the dominant kernel of the Newton step unambiguously.
__global__ __launch_bounds__(256) void schur_gram_f64_kernel(GemmArgs g, int split, double* ws) {
  gemm_f64_body<128, 128, true, false>(g, split, ws);
}

template <int BM, int BN, bool TA, bool TB>
int launch(const GemmArgs& g, hipStream_t stream, int split, double* ws, hipEvent_t after_main = nullptr) {
  const int mt = ceil_div(g.M, BM), nt = ceil_div(g.N, BN);
  const long long tiles = (g.flags & kGemmLowerOut) ? (long long)mt * (mt + 1) / 2 : (long long)mt * nt;
  if (tiles <= 0) return kOk;
  dim3 grid((unsigned)tiles, (unsigned)(g.batch * split));
  if (g.flags & kGemmSchurGram) {
    hipLaunchKernelGGL(schur_gram_f64_kernel, grid, dim3(256), 0, stream, g, split, ws);
  } else {
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, TA, TB>), grid, dim3(256), 0, stream, g, split, ws);
  }
  SDPX_HIP(hipGetLastError());
  if (after_main) SDPX_HIP(hipEventRecord(after_main, stream));
  if (split > 1) {
    const long long MN = (long long)g.M * g.N;
    dim3 rg((unsigned)std::min<long long>(ceil_div(MN, 256), 4096), (unsigned)g.batch);
    hipLaunchKernelGGL(splitk_reduce_kernel, rg, dim3(256), 0, stream, g, split, ws);
    SDPX_HIP(hipGetLastError());
  }
  return kOk;
}

template <int BM, int BN>
int dispatch_trans(const GemmArgs& g, hipStream_t stream, int split, double* ws, bool ta, bool tb) {
  if (!ta && !tb) return launch<BM, BN, false, false>(g, stream, split, ws);
  if (ta && !tb) return launch<BM, BN, true, false>(g, stream, split, ws);
  if (!ta && tb) return launch<BM, BN, false, true>(g, stream, split, ws);
  return launch<BM, BN, true, true>(g, stream, split, ws);
}

}  // namespace

int gemm_f64(const GemmArgs& g0, hipStream_t stream, int split_k, double* workspace, hipEvent_t after_main) {
  if (g0.M <= 0 || g0.N <= 0 || g0.batch <= 0) return kOk;
  if (g0.flags & kGemmSchurGram) {
    // S = Phat' Phat: 128 x 128 lower tiles, op(A) transposed
    if (!(g0.flags & kGemmTransA) || (g0.flags & kGemmTransB) || !(g0.flags & kGemmLowerOut)) {
      set_last_error("gemm_f64: the Schur Gram kernel is TN with lower output");
      return kInvalidArgument;
    }
    GemmArgs g = g0;
    g.flags &= ~(kGemmTransA | kGemmTransB);
    if (split_k < 1) split_k = 1;
    if (split_k > 1 && workspace == nullptr) {
      set_last_error("gemm_f64: split_k > 1 needs a workspace");
      return kInvalidArgument;
    }
    return launch<128, 128, true, false>(g, stream, split_k, workspace, after_main);
  }
  GemmArgs g = g0;
  const bool ta = (g.flags & kGemmTransA) != 0;
  const bool tb = (g.flags & kGemmTransB) != 0;
  g.flags &= ~(kGemmTransA | kGemmTransB);
  if (g.K <= 0) {
    // C = beta*C: run the kernel with an empty K range (accumulators stay zero)
    g.K = 0;
  }
  if (split_k < 1) split_k = 1;
  if (split_k > 1 && workspace == nullptr) {
    set_last_error("gemm_f64: split_k > 1 needs a workspace");
    return kInvalidArgument;
  }
  // Small problems use 64 x 64 tiles so that k x k block products launch enough groups.
  const bool small = (g.M <= 96 && g.N <= 96) || ((long long)g.M * g.N * g.batch < 256LL * 128 * 128 && g.M <= 256 && g.N <= 256);
  if (small) return dispatch_trans<64, 64>(g, stream, split_k, workspace, ta, tb);
  return dispatch_trans<128, 128>(g, stream, split_k, workspace, ta, tb);
}

}  // namespace sdpx
