// FP32 P-hat chain for the BF16x3 Schur assembly (the reduced-precision bulk
// assembly of kernels/mixed_precision_kkt.jl:972-1111 with FP64 as the target
// precision): P-hat_i = L_X^-1 A_i M_Y formed on v_mfma_f32_16x16x4_f32 from an
// FP32 copy of the constraint panel, and written straight out as the two bf16
// planes H = bf16(P-hat), L = bf16(P-hat - H) the BF16x3 Gram kernel reads.
// The FP32 products add a relative error of order k * 2^-24 to P-hat, below
// the 2^-17 of the two-plane split; the refinement against the matrix-free
// FP64 Schur operator measures what is left.
//
// Two products, as in the FP64 chain (the A_i are symmetric):
//   (0) T = [A_1; ...; A_m] M_Y: rows (i, r) of the stacked panel (row-major
//       view of Av, K-contiguous), op(B) = M_Y lower (K >= column), rows written
//       back per block: T_i(r, c) at T + r + c k + i k^2;
//   (1) P-hat = L_X^-1 [T_1 ... T_m]: op(A) = L_X^-1 lower (K <= row,
//       M-contiguous), op(B) = T (K-contiguous), P-hat_i(r, c) to the planes at
//       r + c k + i ldo.
// 128 x 128 tiles, 256 threads = 2 x 2 waves of 64 x 64 (4 x 4 blocks of
// 16 x 16); 16-deep K slices, double-buffered in LDS with the next slice
// prefetched into registers (one barrier per slice).  LDS layouts keep the
// fragment reads and the staging stores free of bank conflicts:
//   K-contiguous operands [row][k], 16 floats a row, 4-float chunks swizzled;
//   M-contiguous operands [k][row], pitch 144 floats.
// Within a tile each wave trims its K range to its own columns (T) or rows
// (P-hat) of the triangular operand, at slice granularity.
#include "common.hpp"

namespace sdpx {
namespace {

constexpr int PB = 128;        // tile side
constexpr int PKS = 16;        // K slice
constexpr int PMC = PB + 16;   // pitch of [k][row] staging
constexpr int kKcTile = PB * PKS;   // floats per K-contiguous slice buffer
constexpr int kMcTile = PKS * PMC;  // floats per M-contiguous slice buffer

// [row][k] staging, 16 floats a row: the 4-float chunk c of row r sits at chunk
// c ^ ((r >> 2) & 3).  A 16-lane group of staging stores (4 rows x 4 chunks) and
// a fragment read (16 rows, one chunk) both cover all 64 banks once.
__device__ __forceinline__ int kc_at(int row, int kk) {
  return row * PKS + 4 * (((kk >> 2) ^ (row >> 2)) & 3) + (kk & 3);
}

typedef float float4_t __attribute__((ext_vector_type(4)));

// Loads of one 128 x 16 operand slice into registers, 8 floats per thread.
// K-contiguous: element (row, kk) at p[kk + row * ld]; thread t takes rows
// t / 4 and t / 4 + 64, K entries 4 (t % 4) .. +3 (one float4 each).
// M-contiguous: element (row, kk) at p[row + kk * ld]; thread t takes K entries
// t / 32 and t / 32 + 8, rows 4 (t % 32) .. +3.
// Rows >= rows_end and K entries >= k_end load as zero; every bound is a
// multiple of 4 (k % 4 == 0), so a float4 is wholly in or out.
__device__ __forceinline__ void load_kc(float4_t (&r)[2], const float* __restrict__ p, long long ld, long long row0,
                                        long long rows_end, int k0, int k_end, int tid) {
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const long long row = row0 + (tid >> 2) + 64 * e;
    const int kk = k0 + 4 * (tid & 3);
    r[e] = (row < rows_end && kk < k_end) ? *reinterpret_cast<const float4_t*>(p + kk + row * ld)
                                          : float4_t{0.f, 0.f, 0.f, 0.f};
  }
}
__device__ __forceinline__ void store_kc(const float4_t (&r)[2], float* s, int tid) {
#pragma unroll
  for (int e = 0; e < 2; ++e)
    *reinterpret_cast<float4_t*>(s + kc_at((tid >> 2) + 64 * e, 4 * (tid & 3))) = r[e];
}
__device__ __forceinline__ void load_mc(float4_t (&r)[2], const float* __restrict__ p, long long ld, int row0,
                                        int rows_end, int k0, int k_end, int tid) {
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int kk = k0 + (tid >> 5) + 8 * e;
    const int row = row0 + 4 * (tid & 31);
    r[e] = (row < rows_end && kk < k_end) ? *reinterpret_cast<const float4_t*>(p + row + (long long)kk * ld)
                                          : float4_t{0.f, 0.f, 0.f, 0.f};
  }
}
__device__ __forceinline__ void store_mc(const float4_t (&r)[2], float* s, int tid) {
#pragma unroll
  for (int e = 0; e < 2; ++e) *reinterpret_cast<float4_t*>(s + ((tid >> 5) + 8 * e) * PMC + 4 * (tid & 31)) = r[e];
}

// XCD-aware tile index: blocks are dispatched round-robin over the XCDs, so
// each XCD gets a contiguous run of tiles; within it the shorter tile dimension
// varies fastest and neighbouring groups share the panel of the longer one.
__device__ __forceinline__ int xcd_linear(int bid, int total) {
  const int xcd = bid % kNumXcd, slot = bid / kNumXcd;
  const int q = total / kNumXcd, rq = total % kNumXcd;
  return xcd * q + (xcd < rq ? xcd : rq) + slot;
}

// MODE 0: T = stacked(A) M_Y;  MODE 1: P-hat = L_X^-1 T -> bf16 planes.
// A: MODE 0 the FP32 panel Av32 (K-contiguous rows, ld k); MODE 1 L_X^-1 (M-contiguous, ld k).
// B: MODE 0 M_Y (K-contiguous columns, ld k); MODE 1 T (K-contiguous columns, ld k).
// M, N: rows and columns of the product (MODE 0: km x k; MODE 1: k x km).
template <int MODE>
__global__ __launch_bounds__(256, 2) void panel_chain_f32_kernel(const float* __restrict__ A,
                                                                 const float* __restrict__ B, int k, long long M,
                                                                 long long N, float* __restrict__ T,
                                                                 __bf16* __restrict__ H, __bf16* __restrict__ L,
                                                                 long long ldo) {
  constexpr int LA = MODE == 0 ? kKcTile : kMcTile;
  __shared__ __align__(16) float As[2][LA];
  __shared__ __align__(16) float Bs[2][kKcTile];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;
  const long long mt = (M + PB - 1) / PB, nt = (N + PB - 1) / PB;
  const long long tile = xcd_linear(blockIdx.x, gridDim.x);
  long long ti, tj;
  if (MODE == 0) {  // N = k: few column tiles, varied fastest
    tj = tile % nt;
    ti = tile / nt;
  } else {  // M = k: few row tiles, varied fastest
    ti = tile % mt;
    tj = tile / mt;
  }
  const long long m0 = ti * PB, n0 = tj * PB;
  // triangular operands: MODE 0 op(B) = M_Y lower (K >= column), MODE 1 op(A) = L_X^-1 lower (K <= row)
  const int kBegin = MODE == 0 ? (int)n0 : 0;
  const int kEnd = MODE == 0 ? k : (int)(m0 + PB < k ? m0 + PB : k);

  float4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  float4_t ra[2], rb[2];
  auto load = [&](int k0) {
    if (MODE == 0) load_kc(ra, A, k, m0, M, k0, kEnd, tid);
    else load_mc(ra, A, k, (int)m0, (int)M, k0, kEnd, tid);
    load_kc(rb, B, k, n0, N, k0, kEnd, tid);
  };
  auto store = [&](int buf) {
    if (MODE == 0) store_kc(ra, As[buf], tid);
    else store_mc(ra, As[buf], tid);
    store_kc(rb, Bs[buf], tid);
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < PKS / 4; ++ks) {
      const int kk = 4 * ks + (lane >> 4);
      float fa[4], fb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = wm0 + 16 * i + (lane & 15);
        fa[i] = MODE == 0 ? As[buf][kc_at(row, kk)] : As[buf][kk * PMC + row];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) fb[j] = Bs[buf][kc_at(wn0 + 16 * j + (lane & 15), kk)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j], fa[i], acc[i][j], 0, 0, 0);
    }
  };

  if (kBegin < kEnd) {
    load(kBegin);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = kBegin; k0 < kEnd; k0 += PKS) {
      const bool more = k0 + PKS < kEnd;
      if (more) load(k0 + PKS);
      // a wave skips the slices in the zero triangle of its own columns (MODE 0: K < its first
      // column) or rows (MODE 1: K > its last row); they would add exact zeros
      const bool idle = MODE == 0 ? k0 + PKS <= n0 + wn0 : k0 > m0 + wm0 + 63;
      if (!idle) compute(cur);
      if (more) {
        store(cur ^ 1);
        __syncthreads();
        cur ^= 1;
      }
    }
  }

  // acc[i][j][r] = C(row = m0 + wm0 + 16 i + lane % 16, col = n0 + wn0 + 16 j + 4 (lane / 16) + r).
  // The stacked index (MODE 0 rows, MODE 1 columns) is split into (block, index in block) by one
  // division per lane and carried along the tile's offsets.
  const long long k2 = (long long)k * k;
  if (MODE == 0) {
    const long long base = m0 + wm0 + (lane & 15);
    long long ib = base / k;
    int rr = (int)(base - ib * k);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i > 0) {
        rr += 16;
        while (rr >= k) {
          rr -= k;
          ++ib;
        }
      }
      if (base + 16 * i >= M) break;
      float* out = T + rr + ib * k2;
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long long col = n0 + wn0 + 16 * j + 4 * (lane >> 4) + r;
          if (col < N) out[col * k] = acc[i][j][r];
        }
    }
  } else {
    const long long base = n0 + wn0 + 4 * (lane >> 4);
    long long ib = base / k;
    int c = (int)(base - ib * k);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (j > 0 || r > 0) {
          c += (r == 0) ? 13 : 1;  // next column: +1 within a block of four, +16 - 3 to the next block
          while (c >= k) {
            c -= k;
            ++ib;
          }
        }
        const long long col = base + 16 * j + r;
        if (col >= N) continue;
        const long long o0 = (long long)c * k + ib * ldo;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = (int)m0 + wm0 + 16 * i + (lane & 15);
          if (row < M) {
            const float v = acc[i][j][r];
            const __bf16 hv = (__bf16)v;
            H[o0 + row] = hv;
            L[o0 + row] = (__bf16)(v - (float)hv);
          }
        }
      }
  }
}

// rows k2 .. ldo - 1 of every column of both planes are zero (the Gram kernel reads them)
__global__ void zero_plane_rows_kernel(__bf16* H, __bf16* L, long long ldo, long long k2, int m) {
  const long long pad = ldo - k2;
  const long long n = pad * m;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const long long c = e / pad, r = k2 + e % pad;
    H[r + c * ldo] = (__bf16)0.f;
    L[r + c * ldo] = (__bf16)0.f;
  }
}

}  // namespace

int panel_chain_bf16(const float* Av32, const float* MY32, const float* LXinv32, int k, int m, float* T32,
                     __bf16* H, __bf16* L, long long ldo, hipStream_t stream) {
  if (k <= 0 || m <= 0) return kOk;
  if (k % 4 != 0) {
    set_last_error("panel_chain_bf16: the block side must be a multiple of 4");
    return kInvalidArgument;
  }
  const long long k2 = (long long)k * k;
  if (ldo < k2) {
    set_last_error("panel_chain_bf16: plane leading dimension below k^2");
    return kInvalidArgument;
  }
  const long long km = (long long)k * m;
  const long long t0 = ((km + PB - 1) / PB) * ((k + PB - 1) / PB);
  if (t0 > 0x7fffffffLL) {
    set_last_error("panel_chain_bf16: too many tiles");
    return kInvalidArgument;
  }
  hipLaunchKernelGGL(panel_chain_f32_kernel<0>, dim3((unsigned)t0), dim3(256), 0, stream, Av32, MY32, k, km,
                     (long long)k, T32, nullptr, nullptr, 0LL);
  SDPX_HIP(hipGetLastError());
  hipLaunchKernelGGL(panel_chain_f32_kernel<1>, dim3((unsigned)t0), dim3(256), 0, stream, LXinv32, T32, k,
                     (long long)k, km, nullptr, H, L, ldo);
  SDPX_HIP(hipGetLastError());
  if (ldo > k2) {
    const long long n = (ldo - k2) * m;
    hipLaunchKernelGGL(zero_plane_rows_kernel, dim3((unsigned)std::min<long long>(ceil_div(n, 256), 65536)), dim3(256),
                       0, stream, H, L, ldo, k2, m);
    SDPX_HIP(hipGetLastError());
  }
  return kOk;
}

}  // namespace sdpx
