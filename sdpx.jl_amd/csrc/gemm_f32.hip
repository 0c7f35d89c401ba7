// Reduced-precision Schur Gram products on CDNA4 matrix cores for the
// mixed-precision route: S = Phat' Phat over lower 128 x 128 tiles, Phat
// columns K-contiguous, accumulated in FP32 per K range and written to FP64 S
// (or to FP64 split-K partials reduced in a fixed order).
//
//  * FP32: Phat rounded to FP32, v_mfma_f32_16x16x4_f32.
//  * BF16x3: Phat split into two bf16 planes, Phat ~= H + L with
//    H = bf16(Phat), L = bf16(Phat - H) (about 16 significant bits), and
//    S_ij ~= sum_k (H_ki H_kj + H_ki L_kj + L_ki H_kj) as three
//    v_mfma_f32_16x16x32_bf16 products per K step (the dropped L L term and
//    the splitting leave a relative error of order 2^-16 per product).
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
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

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

// ---- grouped FP32 Gram (batched route) ------------------------------------------
// One work-group per (GEMM, 64 x 64 lower tile) item: 4 waves in 2 x 2, each a
// 32 x 32 quadrant (2 x 2 blocks of v_mfma_f32_16x16x4_f32), BK = 16 slices
// double-buffered through LDS, the next slice prefetched into registers as one
// float4 per thread and operand.  LDS rows (one K value, 64 tile rows) have a
// pitch of 84 floats: the transposing stores (4 K values x 8 rows per 32
// lanes) and the fragment reads (2 K values x 16 rows per 32 lanes) both land
// on distinct banks.  A diagonal tile stages one operand (A == B) and its
// upper-right quadrant wave skips its MFMAs (nothing of it is stored).
constexpr int GBM = 64;
constexpr int GPITCH = GBM + 20;

__global__ __launch_bounds__(256, 2) void gram_grouped_f32_kernel(const GroupGramF32* __restrict__ G,
                                                                  const GroupTile* __restrict__ T) {
  const GroupTile t = T[blockIdx.x];
  const GroupGramF32 g = G[t.g];
  if (g.skip && *g.skip != 0) return;
  __shared__ float As[2][BK][GPITCH], Bs[2][BK][GPITCH];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 32, wn0 = (wave >> 1) * 32;
  const int m0 = t.ti * GBM, n0 = t.tj * GBM;
  const bool diag = t.ti == t.tj;
  const bool idle = diag && wn0 > wm0;  // quadrant above the diagonal of a diagonal tile
  const int M = g.M, K = g.K;
  const float* P = g.P;
  const int lr = tid >> 2, lk = (tid & 3) * 4;
  const long long ra = (long long)min(m0 + lr, M - 1) * g.ldp;
  const long long rb = (long long)min(n0 + lr, M - 1) * g.ldp;
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  float4_t va, vb;
  auto load = [&](int k0) {
    const int kk = k0 + lk;
    const bool in = kk < K;  // K is a multiple of 4: a float4 is wholly in or out
    va = in ? *reinterpret_cast<const float4_t*>(P + ra + kk) : float4_t{0.f, 0.f, 0.f, 0.f};
    if (!diag) vb = in ? *reinterpret_cast<const float4_t*>(P + rb + kk) : float4_t{0.f, 0.f, 0.f, 0.f};
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      As[buf][lk + e][lr] = va[e];
      if (!diag) Bs[buf][lk + e][lr] = vb[e];
    }
  };
  auto compute = [&](int cur) {
    if (idle) return;
    const float(*Bt)[GPITCH] = diag ? As[cur] : Bs[cur];
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      float fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) fa[i] = As[cur][kr][wm0 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) fb[j] = Bt[kr][wn0 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j], fa[i], acc[i][j], 0, 0, 0);
    }
  };
  load(0);
  store(0);
  __syncthreads();
  int cur = 0;
  int k0 = 0;
  for (; k0 + BK < K; k0 += BK) {
    load(k0 + BK);
    compute(cur);
    store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  compute(cur);
  if (idle) return;
  // acc[i][j][r] = C(row = m0 + wm0 + 16 i + lane % 16, col = n0 + wn0 + 16 j + 4 (lane / 16) + r)
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int row = m0 + wm0 + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + 4 * (lane >> 4) + r;
        if (row < M && row >= col) g.C[row + (long long)col * g.ldc] = (double)acc[i][j][r];
      }
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

// ---- BF16x3 ------------------------------------------------------------------
// K per LDS slice (32 or 64) and the slice rows' LDS layout: rows hold KB / 8 16-byte chunks
// of 8 bf16; with kBf16Swizzle < 0 the pitch is KB + 8 bf16 and chunks stay in place (KB = 32
// only: the padded 64-deep tiles exceed 64 KiB of LDS), otherwise the pitch is KB and chunk c of
// row r sits at chunk c ^ ((r >> kBf16Swizzle) & (KB / 8 - 1)).
// Measured at k = 200, m = 4000 (10 launches): 25.2 ms for 32-deep slices (either layout),
// 21.1 ms for 64-deep slices with swizzle 0 (21.3 with 1, 25.4 with 2): the deeper slice
// doubles the MFMA work between the prefetch and its use, which hides more of the load latency.
#ifndef SDPX_BF16_KB
#define SDPX_BF16_KB 64
#endif
#ifndef SDPX_BF16_SWIZZLE
#define SDPX_BF16_SWIZZLE 0
#endif
constexpr int BKB = SDPX_BF16_KB;
constexpr int kBf16Swizzle = SDPX_BF16_SWIZZLE;
static_assert(BKB == 32 || (BKB == 64 && kBf16Swizzle >= 0), "64-deep slices need the swizzled layout");
constexpr int kChunks = BKB / 8;
constexpr int LDB = kBf16Swizzle < 0 ? BKB + 8 : BKB;
__device__ __forceinline__ int bf16_col(int row, int chunk) {
  return 8 * (kBf16Swizzle < 0 ? chunk : (chunk ^ ((row >> (kBf16Swizzle < 0 ? 0 : kBf16Swizzle)) & (kChunks - 1))));
}

// H = bf16(p), L = bf16(p - H) for every entry; rows >= rows of each column are zero
__global__ void split_bf16_kernel(const double* __restrict__ src, long long ld_src, __bf16* __restrict__ H,
                                  __bf16* __restrict__ L, long long ld_dst, long long rows, int cols) {
  const long long n = ld_dst * cols;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const long long r = i % ld_dst, c = i / ld_dst;
    const double p = r < rows ? src[r + c * ld_src] : 0.0;
    const __bf16 h = (__bf16)(float)p;
    H[i] = h;
    L[i] = (__bf16)(float)(p - (double)(float)h);
  }
}

// One 128 x 128 lower tile of S (or of a split-K partial).  256 threads =
// 2 x 2 waves of 64 x 64 (4 x 4 blocks of 16 x 16); the next BKB-deep slice of
// both planes of both operand tiles is prefetched into registers while the
// current one, staged in LDS, feeds 48 MFMAs per wave per 32 of K.  Operand
// layout of v_mfma_f32_16x16x32_bf16: lane l holds rows (l % 16) with K entries
// 8 (l / 16) .. +7 of both operands (one 16-byte LDS read each); the
// accumulator holds D[4 (l / 16) + r][l % 16].  Operands are fed (B, A) so the
// accumulator holds C' and stores are contiguous.  The planes are zero padded
// to a multiple of BKB rows, so no slice needs masking.
__global__ __launch_bounds__(256, 2) void schur_gram_bf16x3_kernel(const __bf16* __restrict__ H,
                                                                   const __bf16* __restrict__ L, long long ldp,
                                                                   int m, int Kp, double* S, long long lds,
                                                                   double beta, int split, double* ws,
                                                                   const int2* tile_order) {
  constexpr int RPT = kChunks / 2;      // slice rows each thread loads per operand plane
  constexpr int ROW_STEP = 256 / kChunks;
  __shared__ __align__(16) __bf16 As[2][BM][LDB];
  __shared__ __align__(16) __bf16 Bs[2][BM][LDB];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;

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

  int kBegin = 0, kEnd = Kp;
  if (split > 1) {
    const int stages = Kp / BKB;
    const int per = (stages + split - 1) / split;
    kBegin = min(Kp, s * per * BKB);
    kEnd = min(Kp, kBegin + per * BKB);
  }

  // thread loads: slice rows lr + ROW_STEP t of each operand tile, K entries kc .. kc + 7, both planes
  const int lr = tid / kChunks, kc = (tid % kChunks) * 8;
  long long ra[RPT], rb[RPT];
  int sc[RPT];
#pragma unroll
  for (int t = 0; t < RPT; ++t) {
    const int r = lr + ROW_STEP * t;
    ra[t] = (long long)min(m0 + r, m - 1) * ldp;
    rb[t] = (long long)min(n0 + r, m - 1) * ldp;
    sc[t] = bf16_col(r, kc / 8);
  }

  float4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  bf16x8_t v[4 * RPT];
  auto load = [&](int k0) {
    const long long o = k0 + kc;
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
      v[4 * t + 0] = *reinterpret_cast<const bf16x8_t*>(H + ra[t] + o);
      v[4 * t + 1] = *reinterpret_cast<const bf16x8_t*>(L + ra[t] + o);
      v[4 * t + 2] = *reinterpret_cast<const bf16x8_t*>(H + rb[t] + o);
      v[4 * t + 3] = *reinterpret_cast<const bf16x8_t*>(L + rb[t] + o);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
      const int r = lr + ROW_STEP * t;
      *reinterpret_cast<bf16x8_t*>(&As[0][r][sc[t]]) = v[4 * t + 0];
      *reinterpret_cast<bf16x8_t*>(&As[1][r][sc[t]]) = v[4 * t + 1];
      *reinterpret_cast<bf16x8_t*>(&Bs[0][r][sc[t]]) = v[4 * t + 2];
      *reinterpret_cast<bf16x8_t*>(&Bs[1][r][sc[t]]) = v[4 * t + 3];
    }
  };
  auto compute = [&](int ks) {
    const int fr = lane & 15, fq = 4 * ks + (lane >> 4);
    bf16x8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = wm0 + 16 * i + fr, c = bf16_col(r, fq);
      ah[i] = *reinterpret_cast<const bf16x8_t*>(&As[0][r][c]);
      al[i] = *reinterpret_cast<const bf16x8_t*>(&As[1][r][c]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = wn0 + 16 * j + fr, c = bf16_col(r, fq);
      bh[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[0][r][c]);
      bl[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[1][r][c]);
    }
    // the three products of one accumulator are issued 16 MFMAs apart, so that
    // consecutive MFMAs never depend on each other (same order per accumulator)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
  };
  if (kBegin < kEnd) {
    load(kBegin);
    for (int k0 = kBegin; k0 < kEnd; k0 += BKB) {
      store();
      __syncthreads();
      if (k0 + BKB < kEnd) load(k0 + BKB);
#pragma unroll
      for (int ks = 0; ks < BKB / 32; ++ks) compute(ks);
      __syncthreads();
    }
  }
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
          const double v2 = (double)acc[i][j][r];
          *cp = (part || beta == 0.0) ? v2 : v2 + beta * (*cp);
        }
      }
  }
}

// The same product on 256 x 256 lower tiles: 512 threads = 2 x 4 waves of
// 128 x 64 (8 x 4 blocks of 16 x 16), 32-deep slices double-buffered through
// LDS (128 KiB: one group per CU), the next slice prefetched into registers
// while the current one feeds 96 MFMAs per wave.  With 16-cycle bf16 MFMAs the
// 128 x 128 kernel above needs about the whole LDS bandwidth of a CU (its
// fragment reads, 1/32 byte per MFMA flop per plane pair, plus the slice
// stores); the 128 x 64 wave tile reads 3/256 and the 256-row slice is stored
// once for eight waves, about 35% less LDS traffic per flop.  Rows hold 4
// chunks of 8 bf16 (64 bytes), chunk c of row r at chunk c ^ ((r >> 1) & 3):
// the eight lanes that share a 128-byte LDS cycle (fragment reads: rows
// r..r+7 of one chunk; slice stores: rows r, r+1 of four chunks) hit eight
// distinct 16-byte slots.  Each accumulator sees its products in the same
// order as in the 128 x 128 kernel (per 32 of K: L'H, H'L, H'H), so the two
// kernels give the same S for the same split.  A diagonal tile stages one
// operand, and the two waves whose blocks lie wholly above the diagonal skip
// their MFMAs.
constexpr int BMW = 256;
constexpr int KSW = 32;
#ifndef SDPX_WIDE_GROUP
#define SDPX_WIDE_GROUP 4
#endif
constexpr int kWideGroup = SDPX_WIDE_GROUP;
__device__ __forceinline__ int bf16w_col(int row, int chunk) { return 8 * (chunk ^ ((row >> 1) & 3)); }

__global__ __launch_bounds__(512, 1) void schur_gram_bf16x3_wide_kernel(const __bf16* __restrict__ H,
                                                                       const __bf16* __restrict__ L, long long ldp,
                                                                       int m, int Kp, double* S, long long lds,
                                                                       double beta, int split, double* ws) {
  // [buffer][operand A / B][plane H / L][row][K]
  __shared__ __align__(16) __bf16 T[2][2][2][BMW][KSW];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 128, wn0 = (wave >> 1) * 64;

  const int total = gridDim.x;
  const int bid = blockIdx.x;
  const int xcd = bid % kNumXcd, slot = bid / kNumXcd;
  const int q = total / kNumXcd, rq = total % kNumXcd;
  const int lin = xcd * q + (xcd < rq ? xcd : rq) + slot;
  const int per_split = total / split;
  const int sp = lin / per_split;
  const int tile = lin % per_split;
  // lower tiles in kWideGroup x kWideGroup super-blocks (row-major over the blocks and inside each):
  // the groups of an XCD that run together share few Phat panels through its L2
  const int nt = (m + BMW - 1) / BMW;
  int ti = 0, tj = 0;
  {
    int rest = tile;
    const int nb = (nt + kWideGroup - 1) / kWideGroup;
    for (int I = 0; I < nb; ++I) {
      const int r0 = I * kWideGroup, r1 = min(nt, r0 + kWideGroup);
      bool found = false;
      for (int J = 0; J <= I; ++J) {
        const int c0 = J * kWideGroup;
        int cnt = 0;  // tiles of block (I, J): rows r0..r1-1, columns c0..min(c0 + G, row + 1) - 1
        for (int r = r0; r < r1; ++r) cnt += min(c0 + kWideGroup, r + 1) - c0;
        if (rest < cnt) {
          for (int r = r0; r < r1; ++r) {
            const int w = min(c0 + kWideGroup, r + 1) - c0;
            if (rest < w) {
              ti = r;
              tj = c0 + rest;
              break;
            }
            rest -= w;
          }
          found = true;
          break;
        }
        rest -= cnt;
      }
      if (found) break;
    }
  }
  const int m0 = ti * BMW, n0 = tj * BMW;
  const bool diag = ti == tj;
  const bool idle = diag && wn0 >= wm0 + 128;

  int kBegin = 0, kEnd = Kp;
  if (split > 1) {
    const int stages = Kp / KSW;
    const int per = (stages + split - 1) / split;
    kBegin = min(Kp, sp * per * KSW);
    kEnd = min(Kp, kBegin + per * KSW);
  }

  // thread loads: slice rows lr and lr + 128 of each operand tile, chunk ch (K entries 8 ch .. 8 ch + 7), both planes
  const int lr = tid >> 2, ch = tid & 3;
  long long ga[2], gb[2];
  int sc[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int r = lr + 128 * t;
    ga[t] = (long long)min(m0 + r, m - 1) * ldp + 8 * ch;
    gb[t] = (long long)min(n0 + r, m - 1) * ldp + 8 * ch;
    sc[t] = bf16w_col(r, ch);
  }
  bf16x8_t v[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      v[4 * t + 0] = *reinterpret_cast<const bf16x8_t*>(H + ga[t] + k0);
      v[4 * t + 1] = *reinterpret_cast<const bf16x8_t*>(L + ga[t] + k0);
      if (!diag) {
        v[4 * t + 2] = *reinterpret_cast<const bf16x8_t*>(H + gb[t] + k0);
        v[4 * t + 3] = *reinterpret_cast<const bf16x8_t*>(L + gb[t] + k0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int r = lr + 128 * t;
      *reinterpret_cast<bf16x8_t*>(&T[buf][0][0][r][sc[t]]) = v[4 * t + 0];
      *reinterpret_cast<bf16x8_t*>(&T[buf][0][1][r][sc[t]]) = v[4 * t + 1];
      if (!diag) {
        *reinterpret_cast<bf16x8_t*>(&T[buf][1][0][r][sc[t]]) = v[4 * t + 2];
        *reinterpret_cast<bf16x8_t*>(&T[buf][1][1][r][sc[t]]) = v[4 * t + 3];
      }
    }
  };
  float4_t acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  const int fr = lane & 15, fq = lane >> 4;
  const int bop = diag ? 0 : 1;
  bf16x8_t bh[4], bl[4];
  auto load_b = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = wn0 + 16 * j + fr, c = bf16w_col(r, fq);
      bh[j] = *reinterpret_cast<const bf16x8_t*>(&T[buf][bop][0][r][c]);
      bl[j] = *reinterpret_cast<const bf16x8_t*>(&T[buf][bop][1][r][c]);
    }
  };
  // row blocks i0 .. i0 + 3 of the wave tile
  auto compute = [&](int buf, int i0) {
#pragma unroll
    for (int i = i0; i < i0 + 4; ++i) {
      const int r = wm0 + 16 * i + fr, c = bf16w_col(r, fq);
      const bf16x8_t ah = *reinterpret_cast<const bf16x8_t*>(&T[buf][0][0][r][c]);
      const bf16x8_t al = *reinterpret_cast<const bf16x8_t*>(&T[buf][0][1][r][c]);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah, acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al, acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah, acc[i][j], 0, 0, 0);
    }
  };
  if (kBegin < kEnd) {
    load(kBegin);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = kBegin; k0 < kEnd; k0 += KSW) {
      const bool next = k0 + KSW < kEnd;
      if (next) load(k0 + KSW);
      if (!idle) {
        load_b(cur);
        compute(cur, 0);
        compute(cur, 4);
      }
      if (next) store(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
  }
  if (idle) return;
  const bool part = split > 1;
  double* out = part ? ws + (long long)sp * m * m : S;
  const long long ldo = part ? (long long)m : lds;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int row = m0 + wm0 + i * 16 + fr;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + 4 * fq + r;
        if (row < m && row >= col) {
          double* cp = out + row + (long long)col * ldo;
          const double v2 = (double)acc[i][j][r];
          *cp = (part || beta == 0.0) ? v2 : v2 + beta * (*cp);
        }
      }
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

int gram_f32_grouped(const GroupGramF32* groups, const GroupTile* tiles, int ntiles, hipStream_t stream) {
  if (ntiles <= 0) return kOk;
  hipLaunchKernelGGL(gram_grouped_f32_kernel, dim3((unsigned)ntiles), dim3(256), 0, stream, groups, tiles);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int split_panel_bf16(const double* src, long long ld_src, __bf16* H, __bf16* L, long long ld_dst, long long rows,
                     int cols, hipStream_t stream) {
  if (rows <= 0 || cols <= 0) return kOk;
  const long long n = ld_dst * cols;
  hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)std::min<long long>(ceil_div(n, 256), 65536)), dim3(256), 0,
                     stream, src, ld_src, H, L, ld_dst, rows, cols);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int gram_bf16x3(const __bf16* H, const __bf16* L, long long ldp, int m, int Kp, double* S, long long lds,
                double beta, int split, double* ws, const int2* tile_order, hipStream_t stream,
                hipEvent_t after_main) {
  if (m <= 0) return kOk;
  if (ldp % BKB != 0 || Kp % BKB != 0 || Kp > ldp || (reinterpret_cast<uintptr_t>(H) & 15) != 0 ||
      (reinterpret_cast<uintptr_t>(L) & 15) != 0) {
    set_last_error("gram_bf16x3: the bf16 planes need 16-byte aligned bases and K, ld multiples of the slice depth");
    return kInvalidArgument;
  }
  if (split < 1) split = 1;
  if (m >= BMW) {
    const int nt = ceil_div(m, BMW);
    const long long tiles = (long long)nt * (nt + 1) / 2;
    hipLaunchKernelGGL(schur_gram_bf16x3_wide_kernel, dim3((unsigned)(tiles * split)), dim3(512), 0, stream, H, L, ldp,
                       m, Kp, S, lds, beta, split, ws);
  } else {
    const int nt = ceil_div(m, BM);
    const long long tiles = (long long)nt * (nt + 1) / 2;
    hipLaunchKernelGGL(schur_gram_bf16x3_kernel, dim3((unsigned)(tiles * split)), dim3(256), 0, stream, H, L, ldp, m,
                       Kp, S, lds, beta, split, ws, tile_order);
  }
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
