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
// staged through double-buffered LDS in slices of BK = 16: the next slice is
// fetched into registers (branch-free, 16-byte loads when the operand is
// aligned) while the current one feeds the MFMAs, so each slice costs one
// barrier.  The MFMA is fed with (B-fragment, A-fragment) so the accumulator
// holds C transposed: lane%16 indexes a row of C, which makes the
// column-major epilogue stores 128-byte contiguous.
#include "chol_diag.hpp"
#include "common.hpp"

#include <vector>

namespace sdpx {
namespace {

constexpr int BK = 16;     // K slice of the general kernels (the Gram kernel stages 32)
constexpr int PAD = 16;  // [BK][R + 16] layout of operands contiguous along the tile rows
// Operands contiguous along K are staged row-major, [R][BK + 1]: the 17-double
// pitch keeps both the transposing 8-byte stores of the K-contiguous loads and
// the 64-bit fragment reads free of LDS bank conflicts (a [BK][R + 16] layout
// would put the eight lanes that load one row's 16 K values on one bank pair).
// (the K-contiguous pitch is KS + 1 for a K slice of KS)

template <int R, bool KC, int KS = BK>
__device__ __forceinline__ int lds_at(int kk, int r) {
  return KC ? r * (KS + 1) + kk : kk * (R + PAD) + r;
}
template <int R, bool KC, int KS = BK>
constexpr int lds_size() {
  return KC ? R * (KS + 1) : KS * (R + PAD);
}

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef double double2_t __attribute__((ext_vector_type(2)));

// Coordinates (r, kk) inside an R x BK operand tile of element (e, h) owned by
// thread tid.  KC: the operand is contiguous along K in memory; VEC: pairs of
// consecutive elements along the contiguous dimension are fetched together.
template <int R, bool KC, bool VEC, int KS = BK>
__device__ __forceinline__ void tile_coords(int tid, int e, int h, int& r, int& kk) {
  if (KC) {
    if (VEC) {
      kk = (tid % (KS / 2)) * 2 + h;
      r = tid / (KS / 2) + (256 / (KS / 2)) * e;
    } else {
      kk = tid % KS;
      r = tid / KS + (256 / KS) * e;
    }
  } else {
    if (VEC) {
      r = (tid % (R / 2)) * 2 + h;
      kk = tid / (R / 2) + (512 / R) * e;
    } else {
      r = tid % R;
      kk = tid / R + (256 / R) * e;
    }
  }
}

// Loads one R x BK operand slice into registers.  Rows outside the matrix are
// clamped once onto its first or last row (their products only reach C entries that
// the epilogue discards), so full slices need no per-element guard: the main
// K loop is a single basic block of loads, MFMAs and one barrier.  Only the
// final partial slice (K not a multiple of BK) masks along K.
template <int R, bool KC, bool VEC, int KS = BK>
struct TileLoader {
  static constexpr int NE = R * KS / 256;       // doubles per thread
  static constexpr int NL = VEC ? NE / 2 : NE;  // loads per thread
  long long rowoff[NL];                         // offset of (clamped row, kk = element's k)

  __device__ __forceinline__ void init(long long ld, int r0, int rext, int tid) {
    const int rmax = rext - ((VEC && !KC) ? 2 : 1);
#pragma unroll
    for (int e = 0; e < NL; ++e) {
      int r, k;
      tile_coords<R, KC, VEC, KS>(tid, e, 0, r, k);
      const int gr = max(min(r0 + r, rmax), 0);
      rowoff[e] = KC ? (long long)gr * ld + k : (long long)gr + (long long)k * ld;
    }
  }

  __device__ __forceinline__ void load(double* v, const double* __restrict__ P, long long ld, int k0) const {
    const double* Pk = P + (KC ? (long long)k0 : (long long)k0 * ld);
#pragma unroll
    for (int e = 0; e < NL; ++e) {
      if (VEC) {
        const double2_t p = *reinterpret_cast<const double2_t*>(Pk + rowoff[e]);
        v[2 * e] = p[0];
        v[2 * e + 1] = p[1];
      } else {
        v[e] = Pk[rowoff[e]];
      }
    }
  }

  // partial slice: elements with k >= ke read as zero (and are never fetched)
  __device__ __forceinline__ void load_tail(double* v, const double* __restrict__ P, long long ld, int k0, int ke,
                                            int tid) const {
#pragma unroll
    for (int e = 0; e < NL; ++e) {
      int r, k;
      tile_coords<R, KC, VEC, KS>(tid, e, 0, r, k);
      const int gk = k0 + k;
      const long long rowbase = KC ? rowoff[e] - k : rowoff[e] - (long long)k * ld;
      const bool in = gk < ke;
      const long long off = in ? rowoff[e] + (KC ? (long long)k0 : (long long)k0 * ld) : rowbase;
      if (VEC) {
        const double2_t p = *reinterpret_cast<const double2_t*>(P + off);
        v[2 * e] = in ? p[0] : 0.0;
        v[2 * e + 1] = (KC ? gk + 1 < ke : in) ? p[1] : 0.0;
      } else {
        const double x = P[off];
        v[e] = in ? x : 0.0;
      }
    }
  }
};

template <int R, bool KC, bool VEC, int KS = BK>
__device__ __forceinline__ void tile_store(const double* v, double* S, int tid) {
  constexpr int NE = R * KS / 256;
  if (VEC) {
#pragma unroll
    for (int e = 0; e < NE / 2; ++e)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int r, kk;
        tile_coords<R, KC, VEC, KS>(tid, e, h, r, kk);
        S[lds_at<R, KC, KS>(kk, r)] = v[2 * e + h];
      }
  } else {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      int r, kk;
      tile_coords<R, KC, VEC, KS>(tid, e, 0, r, kk);
      S[lds_at<R, KC, KS>(kk, r)] = v[e];
    }
  }
}

// One BM x BN tile of C(m0:, n0:) = alpha * op(A) op(B) (+ beta * C) over
// K in [kBegin, kEnd); only entries with 0 <= row < M, col < N (and
// row >= col when lower_out) are written, at C + row + col * ldc or, with
// brows > 0, at C + row % brows + (row / brows) * bstride + col * ldc.
// LDS (doubles) one tile of the core stages its operands in
template <int BM, int BN, bool TA, bool TB, int KS = BK>
constexpr int core_lds() {
  return 2 * lds_size<BM, TA, KS>() + 2 * lds_size<BN, !TB, KS>();
}

// The tile computation on caller-provided LDS (core_lds<...>() doubles).
// FENCE keeps every MFMA of a slice ahead of the stores of the next one, so
// the prefetch loads have the whole slice to arrive (left to the compiler,
// half the MFMAs are moved past the barrier and the stores wait on the loads
// after 32 of 64; worth 3% on long-K products, 1% worse on short K ranges).
template <int BM, int BN, bool TA, bool TB, bool VEC, int KS = BK, bool FENCE = false>
__device__ __forceinline__ void gemm_core_lds(double* lds, const double* __restrict__ A, long long lda,
                                              const double* __restrict__ B, long long ldb, int M, int N, int kBegin,
                                              int kEnd, int m0, int n0, bool lower_out, double* C, long long ldc,
                                              double alpha, double beta, long long brows, long long bstride,
                                              int skip_lead = 0) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int A_ELEMS = BM * KS / 256;
  constexpr int B_ELEMS = BN * KS / 256;
  constexpr int LA = lds_size<BM, TA, KS>(), LB = lds_size<BN, !TB, KS>();
  double(*As)[LA] = reinterpret_cast<double(*)[LA]>(lds);
  double(*Bs)[LB] = reinterpret_cast<double(*)[LB]>(lds + 2 * LA);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;

  double4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  double ra[A_ELEMS], rb[B_ELEMS];
  // op(A)(i, kk): contiguous along K when TA;  op(B)(kk, j): contiguous along K when !TB
  TileLoader<BM, TA, VEC, KS> la;
  TileLoader<BN, !TB, VEC, KS> lb;
  la.init(lda, m0, M, tid);
  lb.init(ldb, n0, N, tid);
  auto load = [&](int k0) {
    if (k0 + KS <= kEnd) {
      la.load(ra, A, lda, k0);
      lb.load(rb, B, ldb, k0);
    } else {
      la.load_tail(ra, A, lda, k0, kEnd, tid);
      lb.load_tail(rb, B, ldb, k0, kEnd, tid);
    }
  };
  auto store = [&](int buf) {
    tile_store<BM, TA, VEC, KS>(ra, As[buf], tid);
    tile_store<BN, !TB, VEC, KS>(rb, Bs[buf], tid);
  };
  auto compute = [&](int cur) {
#pragma unroll
    for (int ks = 0; ks < KS / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double fa[FM], fb[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) fa[i] = As[cur][lds_at<BM, TA, KS>(kr, wm0 + i * 16 + (lane & 15))];
#pragma unroll
      for (int j = 0; j < FN; ++j) fb[j] = Bs[cur][lds_at<BN, !TB, KS>(kr, wn0 + j * 16 + (lane & 15))];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i], acc[i][j], 0, 0, 0);
    }
  };

  if (kBegin < kEnd) {
    load(kBegin);
    store(0);
    __syncthreads();
    int cur = 0;
    // full slices: prefetch k0 + KS while computing k0 (one barrier per slice)
    const int kFull = kBegin + ((kEnd - kBegin) / KS) * KS;  // end of the full slices
    int k0 = kBegin;
    for (; k0 + 2 * KS <= kFull; k0 += KS) {
      la.load(ra, A, lda, k0 + KS);
      lb.load(rb, B, ldb, k0 + KS);
      compute(cur);
      if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
      store(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
    // remaining one or two slices (the second may be partial)
    if (k0 + KS < kEnd) {
      load(k0 + KS);
      compute(cur);
      store(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
    compute(cur);
  }

  // Epilogue.  The f64 16x16x4 accumulator holds D[lane/16 + 4r][lane%16] in
  // register r of lane `lane`; with the operands swapped D = C', so
  // acc[i][j][r] = C(row = m0+wm0+16i+lane%16, col = n0+wn0+16j+lane/16+4r).
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = m0 + wm0 + i * 16 + (lane & 15);
    const long long roff = brows > 0 ? row % (int)brows + (long long)(row / (int)brows) * bstride : (long long)row;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int col = n0 + wn0 + j * 16 + (lane >> 4) + 4 * rr;
        if (row >= 0 && row < M && col < N && (!lower_out || row >= col) && (row >= skip_lead || col >= skip_lead)) {
          const double v = acc[i][j][rr];
          double* cp = C + roff + (long long)col * ldc;
          *cp = (beta == 0.0) ? alpha * v : alpha * v + beta * (*cp);
        }
      }
    }
  }
}

template <int BM, int BN, bool TA, bool TB, bool VEC, bool FENCE = false>
__device__ __forceinline__ void gemm_core(const double* __restrict__ A, long long lda, const double* __restrict__ B,
                                          long long ldb, int M, int N, int kBegin, int kEnd, int m0, int n0,
                                          bool lower_out, double* C, long long ldc, double alpha, double beta,
                                          long long brows = 0, long long bstride = 0) {
  __shared__ __align__(16) double lds[core_lds<BM, BN, TA, TB>()];
  gemm_core_lds<BM, BN, TA, TB, VEC, BK, FENCE>(lds, A, lda, B, ldb, M, N, kBegin, kEnd, m0, n0, lower_out, C, ldc,
                                                alpha, beta, brows, bstride);
}

// Diagonal 128 x 128 tile of the Gram product S = P' P (m0 == n0, all rows in
// range).  Both operands are the same K slice of P, staged once, and only the
// 36 of the 64 16 x 16 blocks on or below the diagonal are computed, nine per
// wave, so the group takes 9/16 of a full tile's MFMA time.  Blocks are named
// by their 16-row block index (0..7) along each side:
//   waves 0-2 (SQ): 3 x 3 squares, rows r0..r0+2 by columns c0..c0+2 with
//                   (r0, c0) = (5, 0), (5, 3), (2, 0);
//   wave 3:         the 2 x 2 lower triangles {(a,a), (a+1,a), (a+1,a+1)}
//                   at a = 0, 3, 6.
// Each accumulator sees the same operands in the same K order as in the full
// tile, so the written entries are bitwise those of the general path.
template <bool VEC, int KS, bool SQ>
__device__ __forceinline__ void gram_diag_core(double* lds, const double* __restrict__ P, long long ld, int M,
                                               int kBegin, int kEnd, int m0, double* C, long long ldc, double alpha,
                                               double beta, int wave) {
  constexpr int LA = lds_size<128, true, KS>();
  double(*As)[LA] = reinterpret_cast<double(*)[LA]>(lds);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // fragment block indices fb[0..5] and the nine (row, column) pairs into them
  int fbk[6];
  if (SQ) {
    const int r0 = wave == 2 ? 2 : 5, c0 = wave == 1 ? 3 : 0;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      fbk[t] = r0 + t;
      fbk[3 + t] = c0 + t;
    }
  } else {
    fbk[0] = 0, fbk[1] = 1, fbk[2] = 3, fbk[3] = 4, fbk[4] = 6, fbk[5] = 7;
  }
  constexpr int pr_sq[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2}, pc_sq[9] = {3, 4, 5, 3, 4, 5, 3, 4, 5};
  constexpr int pr_tr[9] = {0, 1, 1, 2, 3, 3, 4, 5, 5}, pc_tr[9] = {0, 0, 1, 2, 2, 3, 4, 4, 5};

  double4_t acc[9];
#pragma unroll
  for (int b = 0; b < 9; ++b) acc[b] = double4_t{0.0, 0.0, 0.0, 0.0};

  constexpr int NE = 128 * KS / 256;
  double ra[NE];
  TileLoader<128, true, VEC, KS> la;
  la.init(ld, m0, M, tid);
  auto load = [&](int k0) {
    if (k0 + KS <= kEnd) la.load(ra, P, ld, k0);
    else la.load_tail(ra, P, ld, k0, kEnd, tid);
  };
  auto compute = [&](int cur) {
#pragma unroll
    for (int ks = 0; ks < KS / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double f[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) f[q] = As[cur][lds_at<128, true, KS>(kr, fbk[q] * 16 + (lane & 15))];
#pragma unroll
      for (int b = 0; b < 9; ++b) {
        const int r = SQ ? pr_sq[b] : pr_tr[b], c = SQ ? pc_sq[b] : pc_tr[b];
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(f[c], f[r], acc[b], 0, 0, 0);
      }
    }
  };

  if (kBegin < kEnd) {
    load(kBegin);
    tile_store<128, true, VEC, KS>(ra, As[0], tid);
    __syncthreads();
    int cur = 0;
    const int kFull = kBegin + ((kEnd - kBegin) / KS) * KS;
    int k0 = kBegin;
    for (; k0 + 2 * KS <= kFull; k0 += KS) {
      la.load(ra, P, ld, k0 + KS);
      compute(cur);
      __builtin_amdgcn_sched_barrier(0);
      tile_store<128, true, VEC, KS>(ra, As[cur ^ 1], tid);
      __syncthreads();
      cur ^= 1;
    }
    if (k0 + KS < kEnd) {
      load(k0 + KS);
      compute(cur);
      tile_store<128, true, VEC, KS>(ra, As[cur ^ 1], tid);
      __syncthreads();
      cur ^= 1;
    }
    compute(cur);
  }

#pragma unroll
  for (int b = 0; b < 9; ++b) {
    const int rb = fbk[SQ ? pr_sq[b] : pr_tr[b]], cb = fbk[SQ ? pc_sq[b] : pc_tr[b]];
    const int row = m0 + rb * 16 + (lane & 15);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int col = m0 + cb * 16 + (lane >> 4) + 4 * rr;
      if (row >= col) {
        double* cp = C + row + (long long)col * ldc;
        *cp = (beta == 0.0) ? alpha * acc[b][rr] : alpha * acc[b][rr] + beta * (*cp);
      }
    }
  }
}

// GRAM (the Schur Gram kernel): a partial last tile row with at most 64 rows
// in range runs as flat strips of the same MFMA count per wave, 32 x 256 (two
// tiles in one group; the group of the odd tile exits) or 64 x 128, instead
// of 128-row tiles that compute 2 to 4 times the rows they write.
template <int BM, int BN, bool TA, bool TB, bool VEC, int KS = BK, bool GRAM = false>
__device__ __forceinline__ void gemm_f64_body(const GemmArgs& g, int split, double* ws, double* lds) {
  const int b = blockIdx.y;
  if (g.skip && g.skip[(long long)b * g.skip_stride] != 0) return;

  // XCD-aware remap: blocks are dispatched round-robin over the 8 XCDs, so
  // give each XCD a contiguous run of (split, tile) work items: with split-K
  // an XCD owns whole K ranges, and within a range the tile order keeps the
  // co-resident groups on few A/B panels (shared through that XCD's L2).
  const int total = gridDim.x;
  const int bid = blockIdx.x;
  const int xcd = bid % kNumXcd;
  const int slot = bid / kNumXcd;
  const int q = total / kNumXcd, rq = total % kNumXcd;
  const int lin = xcd * q + (xcd < rq ? xcd : rq) + slot;
  const int per_split = total / split;
  int s = lin / per_split;
  const int tile = lin % per_split;

  int ti, tj;
  const bool lower_out = (g.flags & kGemmLowerOut) != 0;
  if (GRAM && g.item_order) {
    const int2 it = g.item_order[lin];
    s = it.x;
    ti = it.y >> 16;
    tj = it.y & 0xffff;
  } else if (lower_out && g.tile_order) {
    ti = g.tile_order[tile].x;
    tj = g.tile_order[tile].y;
  } else if (lower_out) {
    // linear index over lower tiles, row-major: tile = ti*(ti+1)/2 + tj
    int t = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
    while ((long long)(t + 1) * (t + 2) / 2 <= tile) ++t;
    while ((long long)t * (t + 1) / 2 > tile) --t;
    ti = t;
    tj = tile - t * (t + 1) / 2;
  } else {
    // the shorter tile dimension varies fastest: neighbouring groups share
    // the panel of the longer one through L2
    const int mt = (g.M + g.row_off + BM - 1) / BM;
    const int nt = (g.N + BN - 1) / BN;
    if (mt > nt) {
      tj = tile % nt;
      ti = tile / nt;
    } else {
      ti = tile % mt;
      tj = tile / mt;
    }
  }
  const int m0 = ti * BM - g.row_off;
  const int n0 = tj * BN;

  const double* A = g.A + (long long)b * g.strideA;
  const double* B = g.B + (long long)b * g.strideB;

  int kBegin = 0, kEnd = g.K;
  if (g.flags & kGemmALower) kEnd = min(kEnd, m0 + BM);
  if (g.flags & kGemmBLower) kBegin = max(kBegin, n0);
  if (g.flags & kGemmAUpper) kBegin = max(kBegin, m0);
  kBegin = (kBegin / KS) * KS;
  if (split > 1) {
    const int stages = (kEnd - kBegin + KS - 1) / KS;
    const int per = (stages + split - 1) / split;
    const int sb = kBegin + s * per * KS;
    const int se = min(kEnd, sb + per * KS);
    kBegin = sb;
    kEnd = se;
  }
  // split-K partials go to W (M x M per split, no beta); otherwise straight to C
  const bool part = split > 1;
  double* out = part ? ws + ((long long)b * split + s) * (long long)g.M * g.N : g.C + (long long)b * g.strideC;
  if constexpr (GRAM) {
    const int rows = g.M - m0;  // < BM only in the last tile row
    if (rows <= 32) {
      if (tj & 1) return;  // covered by the strip of tile column tj - 1
      gemm_core_lds<32, 2 * BN, TA, TB, VEC, KS, true>(lds, A, g.lda, B, g.ldb, g.M, g.N, kBegin, kEnd, m0, n0,
                                                       lower_out, out, part ? (long long)g.M : g.ldc, g.alpha,
                                                       part ? 0.0 : g.beta, 0, 0);
      return;
    }
    if (rows <= 64) {
      gemm_core_lds<64, BN, TA, TB, VEC, KS, true>(lds, A, g.lda, B, g.ldb, g.M, g.N, kBegin, kEnd, m0, n0,
                                                   lower_out, out, part ? (long long)g.M : g.ldc, g.alpha,
                                                   part ? 0.0 : g.beta, 0, 0);
      return;
    }
    if (ti == tj && rows >= BM && A == B && g.lda == g.ldb) {
      const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
      const long long ldo = part ? (long long)g.M : g.ldc;
      const double bo = part ? 0.0 : g.beta;
      if (wave < 3) gram_diag_core<VEC, KS, true>(lds, A, g.lda, g.M, kBegin, kEnd, m0, out, ldo, g.alpha, bo, wave);
      else gram_diag_core<VEC, KS, false>(lds, A, g.lda, g.M, kBegin, kEnd, m0, out, ldo, g.alpha, bo, wave);
      return;
    }
  }
  gemm_core_lds<BM, BN, TA, TB, VEC, KS, GRAM>(lds, A, g.lda, B, g.ldb, g.M, g.N, kBegin, kEnd, m0, n0, lower_out, out,
                                         part ? (long long)g.M : g.ldc, g.alpha, part ? 0.0 : g.beta,
                                         part ? 0 : g.block_rows, g.block_stride);
}

// Diagonal 64 x 64 tile of a grouped Gram product C = P' P (lower output,
// m0 == n0): the 10 of 16 blocks of 16 x 16 on or below the diagonal, from one
// staged operand slice.  Blocks are named by 16-row block index (0..3); wave W
// reads the fragments of blocks fb[] and computes the (row, column) pairs
// (fb[pr[b]], fb[pc[b]]):
//   wave 0: (2,0) (3,0) (3,1)    wave 1: (2,1) (2,2) (3,2)
//   wave 2: (1,0) (1,1)          wave 3: (0,0) (3,3)
// so the group takes 3/4 of a full tile's MFMA time.  Each accumulator sees the
// same operands in the same K order as in the full tile.  Rows at or beyond M
// (a partial tile) are computed on clamped rows and not written.
template <int W>
struct DiagPlan64;
template <>
struct DiagPlan64<0> {
  static constexpr int nf = 4, np = 3;
  static constexpr int fb[4] = {2, 3, 0, 1}, pr[3] = {0, 1, 1}, pc[3] = {2, 2, 3};
};
template <>
struct DiagPlan64<1> {
  static constexpr int nf = 3, np = 3;
  static constexpr int fb[3] = {1, 2, 3}, pr[3] = {1, 1, 2}, pc[3] = {0, 1, 1};
};
template <>
struct DiagPlan64<2> {
  static constexpr int nf = 2, np = 2;
  static constexpr int fb[2] = {0, 1}, pr[2] = {1, 1}, pc[2] = {0, 1};
};
template <>
struct DiagPlan64<3> {
  static constexpr int nf = 2, np = 2;
  static constexpr int fb[2] = {0, 3}, pr[2] = {0, 1}, pc[2] = {0, 1};
};

template <int W, bool VEC>
__device__ __forceinline__ void grouped_diag_core(double* lds, const double* __restrict__ P, long long ld, int M,
                                                  int kBegin, int kEnd, int m0, double* C, long long ldc,
                                                  double alpha, double beta) {
  using Plan = DiagPlan64<W>;
  constexpr int LA = lds_size<64, true, BK>();
  double(*As)[LA] = reinterpret_cast<double(*)[LA]>(lds);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  double4_t acc[Plan::np];
#pragma unroll
  for (int b = 0; b < Plan::np; ++b) acc[b] = double4_t{0.0, 0.0, 0.0, 0.0};
  constexpr int NE = 64 * BK / 256;
  double ra[NE];
  TileLoader<64, true, VEC, BK> la;
  la.init(ld, m0, M, tid);
  auto load = [&](int k0) {
    if (k0 + BK <= kEnd) la.load(ra, P, ld, k0);
    else la.load_tail(ra, P, ld, k0, kEnd, tid);
  };
  auto compute = [&](int cur) {
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double f[Plan::nf];
#pragma unroll
      for (int q = 0; q < Plan::nf; ++q) f[q] = As[cur][lds_at<64, true, BK>(kr, Plan::fb[q] * 16 + (lane & 15))];
#pragma unroll
      for (int b = 0; b < Plan::np; ++b)
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(f[Plan::pc[b]], f[Plan::pr[b]], acc[b], 0, 0, 0);
    }
  };
  if (kBegin < kEnd) {
    load(kBegin);
    tile_store<64, true, VEC, BK>(ra, As[0], tid);
    __syncthreads();
    int cur = 0;
    const int kFull = kBegin + ((kEnd - kBegin) / BK) * BK;
    int k0 = kBegin;
    for (; k0 + 2 * BK <= kFull; k0 += BK) {
      la.load(ra, P, ld, k0 + BK);
      compute(cur);
      tile_store<64, true, VEC, BK>(ra, As[cur ^ 1], tid);
      __syncthreads();
      cur ^= 1;
    }
    if (k0 + BK < kEnd) {
      load(k0 + BK);
      compute(cur);
      tile_store<64, true, VEC, BK>(ra, As[cur ^ 1], tid);
      __syncthreads();
      cur ^= 1;
    }
    compute(cur);
  }
#pragma unroll
  for (int b = 0; b < Plan::np; ++b) {
    const int row = m0 + Plan::fb[Plan::pr[b]] * 16 + (lane & 15);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int col = m0 + Plan::fb[Plan::pc[b]] * 16 + (lane >> 4) + 4 * rr;
      if (row < M && row >= col) {
        double* cp = C + row + (long long)col * ldc;
        *cp = (beta == 0.0) ? alpha * acc[b][rr] : alpha * acc[b][rr] + beta * (*cp);
      }
    }
  }
}

// ---- grouped launches: one work-group per (GEMM, tile) item ------------------
template <int BM, int BN, bool TA, bool TB, bool VEC>
__global__ __launch_bounds__(256, 2) void gemm_grouped_kernel(const GroupGemm* __restrict__ G,
                                                             const GroupTile* __restrict__ T) {
  const GroupTile t = T[blockIdx.x];
  const GroupGemm g = G[t.g];
  if (g.skip && *g.skip != 0) return;
  const int m0 = t.ti * BM, n0 = t.tj * BN;
  int kBegin = 0, kEnd = g.K;
  if (g.flags & kGemmALower) kEnd = min(kEnd, m0 + BM);
  if (g.flags & kGemmBLower) kBegin = max(kBegin, n0);
  if (g.flags & kGemmAUpper) kBegin = max(kBegin, m0);
  kBegin = (kBegin / BK) * BK;
  __shared__ __align__(16) double lds[core_lds<BM, BN, TA, TB>()];
  if constexpr (TA && !TB && BM == 64 && BN == 64) {
    // diagonal tiles of the batched Schur Gram products
    if ((g.flags & kGemmLowerOut) && t.ti == t.tj && g.A == g.B && g.lda == g.ldb) {
      static_assert(lds_size<64, true, BK>() * 2 <= core_lds<BM, BN, TA, TB>(), "LDS");
      switch (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) {
        case 0: grouped_diag_core<0, VEC>(lds, g.A, g.lda, g.M, kBegin, kEnd, m0, g.C, g.ldc, g.alpha, g.beta); break;
        case 1: grouped_diag_core<1, VEC>(lds, g.A, g.lda, g.M, kBegin, kEnd, m0, g.C, g.ldc, g.alpha, g.beta); break;
        case 2: grouped_diag_core<2, VEC>(lds, g.A, g.lda, g.M, kBegin, kEnd, m0, g.C, g.ldc, g.alpha, g.beta); break;
        default: grouped_diag_core<3, VEC>(lds, g.A, g.lda, g.M, kBegin, kEnd, m0, g.C, g.ldc, g.alpha, g.beta);
      }
      return;
    }
  }
  // (the slice fence of the Gram kernel measured 0.7% slower on the batched
  // Schur Gram products here: 64 x 64 tiles issue 16 MFMAs per slice)
  gemm_core_lds<BM, BN, TA, TB, VEC>(lds, g.A, g.lda, g.B, g.ldb, g.M, g.N, kBegin, kEnd, m0, n0,
                                     (g.flags & kGemmLowerOut) != 0, g.C, g.ldc, g.alpha, g.beta, 0, 0);
}

template <int BM, int BN, bool TA, bool TB, bool VEC>
__global__ __launch_bounds__(256, 2) void gemm_f64_kernel(GemmArgs g, int split, double* ws) {
  __shared__ __align__(16) double lds[core_lds<BM, BN, TA, TB>()];
  gemm_f64_body<BM, BN, TA, TB, VEC>(g, split, ws, lds);
}

// Gram kernel configuration: K slice and resident work-groups per CU (a 32-deep
// slice with one resident group per CU, 135 KB of LDS, measured 55 ms against
// 42-44 ms: the second resident group hides more latency than the halved
// barrier count saves)
constexpr int kGramKs = BK;
constexpr int kGramWgPerCu = 2;

// The Schur Gram product S = Phat' Phat (lower tiles) under its own symbol,
// so profiles attribute the dominant kernel of the Newton step unambiguously.
constexpr int kGramLds = std::max({core_lds<128, 128, true, false, kGramKs>(), core_lds<32, 256, true, false, kGramKs>(),
                                   core_lds<64, 128, true, false, kGramKs>()});
static_assert(kGramLds * 8 * kGramWgPerCu <= 160 * 1024, "two Gram groups per CU");
template <bool VEC>
__global__ __launch_bounds__(256, kGramWgPerCu) void schur_gram_f64_kernel(GemmArgs g, int split, double* ws) {
  __shared__ __align__(16) double lds[kGramLds];
  gemm_f64_body<128, 128, true, false, VEC, kGramKs, true>(g, split, ws, lds);
}

// Trailing update of a blocked Cholesky, A22 -= A21 A21' over the lower tiles
// of A22 (one work-group per tile, batch entries along y), fused with the
// factor of the next diagonal block: work-group 0 of each entry (dispatched
// first) updates only that jbn x jbn block (a K = jb product) and factors it
// at once, while the tile groups update everything else (the tile holding the
// block leaves it out), so the factor's latency overlaps the update.
template <int BT, bool VEC>
__global__ __launch_bounds__(256, 2) void trailing_diag_kernel(GemmArgs t, int jn, int jbn, double* dinv,
                                                               long long dinv_stride, int* info) {
  constexpr int kCore = core_lds<BT, BT, false, true>();
  __shared__ __align__(16) double lds[kCore > chol::kDiagLds ? kCore : chol::kDiagLds];
  static_assert(core_lds<64, 64, false, true>() <= (kCore > chol::kDiagLds ? kCore : chol::kDiagLds), "LDS");
  const int b = blockIdx.y;
  if (info[b] != 0) return;
  double* C = t.C + (long long)b * t.strideC;
  const double* A = t.A + (long long)b * t.strideA;
  const double* B = t.B + (long long)b * t.strideB;
  if (blockIdx.x == 0) {
    gemm_core_lds<64, 64, false, true, VEC>(lds, A, t.lda, B, t.ldb, jbn, jbn, 0, t.K, 0, 0, true, C, t.ldc, t.alpha,
                                            t.beta, 0, 0);
    __syncthreads();  // the block's update is written by this work-group; LDS is reused
    chol::potrf_diag_body(C, t.ldc, jn, jbn,
                          dinv + (long long)b * dinv_stride + (long long)(jn / kCholNb) * kCholNb * kCholNb, info + b,
                          lds);
    return;
  }
  // lower tiles, row-major: tile = ti (ti + 1) / 2 + tj
  const int tile = blockIdx.x - 1;
  int ti = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
  while (ti * (ti + 1) / 2 > tile) --ti;
  const int tj = tile - ti * (ti + 1) / 2;
  gemm_core_lds<BT, BT, false, true, VEC>(lds, A, t.lda, B, t.ldb, t.M, t.N, 0, t.K, ti * BT, tj * BT, true, C, t.ldc,
                                          t.alpha, t.beta, 0, 0, jbn);
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

template <int BM, int BN, bool TA, bool TB, bool VEC>
int launch(const GemmArgs& g0, hipStream_t stream, int split, double* ws, hipEvent_t after_main) {
  GemmArgs g = g0;
  const int mt = ceil_div(g.M, BM), nt = ceil_div(g.N, BN);
  g.row_off = (g.flags & kGemmRaggedFirst) ? mt * BM - g.M : 0;
  const long long tiles = (g.flags & kGemmLowerOut) ? (long long)mt * (mt + 1) / 2 : (long long)mt * nt;
  if (tiles <= 0) return kOk;
  dim3 grid((unsigned)(tiles * split), (unsigned)g.batch);
  if (g.flags & kGemmSchurGram) {
    hipLaunchKernelGGL((schur_gram_f64_kernel<VEC>), grid, dim3(256), 0, stream, g, split, ws);
  } else {
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, TA, TB, VEC>), grid, dim3(256), 0, stream, g, split, ws);
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

template <int BM, int BN, bool VEC>
int dispatch_trans(const GemmArgs& g, hipStream_t stream, int split, double* ws, bool ta, bool tb, hipEvent_t ev) {
  if (!ta && !tb) return launch<BM, BN, false, false, VEC>(g, stream, split, ws, ev);
  if (ta && !tb) return launch<BM, BN, true, false, VEC>(g, stream, split, ws, ev);
  if (!ta && tb) return launch<BM, BN, false, true, VEC>(g, stream, split, ws, ev);
  return launch<BM, BN, true, true, VEC>(g, stream, split, ws, ev);
}

// 16-byte operand fetches need the contiguous dimension, leading dimension
// and batch strides even and 16-byte aligned bases.
bool vec_ok(const GemmArgs& g, bool ta, bool tb) {
  auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  const bool a_ok = aligned(g.A) && g.lda % 2 == 0 && (ta ? g.K % 2 == 0 : g.M % 2 == 0) &&
                    (g.batch == 1 || g.strideA % 2 == 0);
  const bool b_ok = aligned(g.B) && g.ldb % 2 == 0 && (!tb ? g.K % 2 == 0 : g.N % 2 == 0) &&
                    (g.batch == 1 || g.strideB % 2 == 0);
  return a_ok && b_ok;
}

}  // namespace

int gemm_f64(const GemmArgs& g0, hipStream_t stream, int split_k, double* workspace, hipEvent_t after_main) {
  if (g0.M <= 0 || g0.N <= 0 || g0.batch <= 0) return kOk;
  GemmArgs g = g0;
  const bool ta = (g.flags & kGemmTransA) != 0;
  const bool tb = (g.flags & kGemmTransB) != 0;
  g.flags &= ~(kGemmTransA | kGemmTransB);
  if (g.K < 0) g.K = 0;  // K = 0: C = beta*C (accumulators stay zero)
  if (split_k < 1) split_k = 1;
  if (split_k > 1 && workspace == nullptr) {
    set_last_error("gemm_f64: split_k > 1 needs a workspace");
    return kInvalidArgument;
  }
  if ((g.block_rows > 0 || (g.flags & kGemmRaggedFirst)) && (split_k > 1 || (g.flags & kGemmLowerOut))) {
    set_last_error("gemm_f64: blocked output and ragged-first tiles need split_k = 1 and a full output");
    return kInvalidArgument;
  }
  const bool vec = vec_ok(g, ta, tb);
  if (g.flags & kGemmSchurGram) {
    // S = Phat' Phat: 128 x 128 lower tiles, op(A) transposed
    if (!ta || tb || !(g.flags & kGemmLowerOut)) {
      set_last_error("gemm_f64: the Schur Gram kernel is TN with lower output");
      return kInvalidArgument;
    }
    return vec ? launch<128, 128, true, false, true>(g, stream, split_k, workspace, after_main)
               : launch<128, 128, true, false, false>(g, stream, split_k, workspace, after_main);
  }
  // Rectangular tiles on request (chosen by the plan-time autotuner for the
  // P-hat panel products; instantiated for the NN and TN forms those use).
  if ((g.flags & (kGemmTallTiles | kGemmWideTiles)) && !tb) {
    const bool tall = (g.flags & kGemmTallTiles) != 0;
    if (ta) {
      if (tall)
        return vec ? launch<128, 64, true, false, true>(g, stream, split_k, workspace, after_main)
                   : launch<128, 64, true, false, false>(g, stream, split_k, workspace, after_main);
      return vec ? launch<64, 128, true, false, true>(g, stream, split_k, workspace, after_main)
                 : launch<64, 128, true, false, false>(g, stream, split_k, workspace, after_main);
    }
    if (tall)
      return vec ? launch<128, 64, false, false, true>(g, stream, split_k, workspace, after_main)
                 : launch<128, 64, false, false, false>(g, stream, split_k, workspace, after_main);
    return vec ? launch<64, 128, false, false, true>(g, stream, split_k, workspace, after_main)
               : launch<64, 128, false, false, false>(g, stream, split_k, workspace, after_main);
  }
  // Small problems use 64 x 64 tiles so that k x k block products launch enough groups.
  // (and so does a product with a side of at most 64, which would leave half of a 128-wide tile idle)
  // A full-output product that would launch fewer 128 x 128 groups than half
  // the CUs (the k x k block products) takes 64 x 64 tiles too: four times the
  // groups for the same K loop (every entry's K order is unchanged).
  const long long tiles128 = (long long)ceil_div(g.M, 128) * ceil_div(g.N, 128) * g.batch;
  const bool small = (g.flags & kGemmSmallTiles) || (g.M <= 96 && g.N <= 96) || g.M <= 64 || g.N <= 64 ||
                     ((long long)g.M * g.N * g.batch < 256LL * 128 * 128 && g.M <= 256 && g.N <= 256) ||
                     (!(g.flags & kGemmLowerOut) && tiles128 < 128);
  if (small) {
    return vec ? dispatch_trans<64, 64, true>(g, stream, split_k, workspace, ta, tb, after_main)
               : dispatch_trans<64, 64, false>(g, stream, split_k, workspace, ta, tb, after_main);
  }
  return vec ? dispatch_trans<128, 128, true>(g, stream, split_k, workspace, ta, tb, after_main)
             : dispatch_trans<128, 128, false>(g, stream, split_k, workspace, ta, tb, after_main);
}

}  // namespace sdpx

namespace sdpx {

int gemm_f64_grouped(const GroupGemm* groups, const GroupTile* tiles, int ntiles, bool ta, bool tb, bool vec,
                     hipStream_t stream) {
  if (ntiles <= 0) return kOk;
  const dim3 grid((unsigned)ntiles), block(256);
  if (ta && !tb) {
    if (vec)
      hipLaunchKernelGGL((gemm_grouped_kernel<64, 64, true, false, true>), grid, block, 0, stream, groups, tiles);
    else
      hipLaunchKernelGGL((gemm_grouped_kernel<64, 64, true, false, false>), grid, block, 0, stream, groups, tiles);
  } else if (!ta && tb) {
    if (vec)
      hipLaunchKernelGGL((gemm_grouped_kernel<64, 64, false, true, true>), grid, block, 0, stream, groups, tiles);
    else
      hipLaunchKernelGGL((gemm_grouped_kernel<64, 64, false, true, false>), grid, block, 0, stream, groups, tiles);
  } else if (!ta && !tb) {
    if (vec)
      hipLaunchKernelGGL((gemm_grouped_kernel<64, 64, false, false, true>), grid, block, 0, stream, groups, tiles);
    else
      hipLaunchKernelGGL((gemm_grouped_kernel<64, 64, false, false, false>), grid, block, 0, stream, groups, tiles);
  } else {
    set_last_error("gemm_f64_grouped: TT products are not instantiated");
    return kInvalidArgument;
  }
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int potrf_trailing_diag(const GemmArgs& t, int jn, int jbn, double* dinv, long long dinv_stride, int* info,
                        hipStream_t stream) {
  if (t.M <= 0 || t.batch <= 0) return kOk;
  if (t.M != t.N || (t.flags & ~(kGemmTransB | kGemmLowerOut)) != 0 || !(t.flags & kGemmTransB) ||
      !(t.flags & kGemmLowerOut) || jbn < 1 || jbn > kCholNb || jbn > t.M) {
    set_last_error("potrf_trailing_diag: expects the NT lower-output trailing update of a blocked Cholesky");
    return kInvalidArgument;
  }
  // the tile side gemm_f64 picks for the same product
  const bool small = (t.M <= 96) || ((long long)t.M * t.N * t.batch < 256LL * 128 * 128 && t.M <= 256);
  const bool vec = vec_ok(t, false, true);
  const int BT = small ? 64 : 128;
  const int mt = ceil_div(t.M, BT);
  const dim3 grid((unsigned)((long long)mt * (mt + 1) / 2 + 1), (unsigned)t.batch);  // + the diagonal group
  if (small) {
    if (vec)
      hipLaunchKernelGGL((trailing_diag_kernel<64, true>), grid, dim3(256), 0, stream, t, jn, jbn, dinv, dinv_stride, info);
    else
      hipLaunchKernelGGL((trailing_diag_kernel<64, false>), grid, dim3(256), 0, stream, t, jn, jbn, dinv, dinv_stride,
                         info);
  } else {
    if (vec)
      hipLaunchKernelGGL((trailing_diag_kernel<128, true>), grid, dim3(256), 0, stream, t, jn, jbn, dinv, dinv_stride,
                         info);
    else
      hipLaunchKernelGGL((trailing_diag_kernel<128, false>), grid, dim3(256), 0, stream, t, jn, jbn, dinv, dinv_stride,
                         info);
  }
  SDPX_HIP(hipGetLastError());
  return kOk;
}

void lower_tile_order(int nt, int group, int2* out) {
  int n = 0;
  for (int si = 0; si < nt; si += group)
    for (int sj = 0; sj <= si; sj += group)
      for (int ti = si; ti < std::min(nt, si + group); ++ti)
        for (int tj = sj; tj < std::min(sj + group, ti + 1); ++tj) out[n++] = make_int2(ti, tj);
}

void gram_item_order(int m, int group, int split, int2* out) {
  const int nt = ceil_div(m, 128);
  const int per_split = nt * (nt + 1) / 2;
  std::vector<int2> order((size_t)per_split);
  lower_tile_order(nt, group, order.data());
  // length class of a tile in the Gram kernel: 0 full, 1 diagonal (9/16),
  // 2 strip of a short last tile row (1/2), 3 exits at once
  auto cls = [&](int2 t) {
    const int rows = m - t.x * 128;
    if (rows <= 32) return (t.y & 1) ? 3 : 2;
    if (rows <= 64) return 2;
    return (t.x == t.y && rows >= 128) ? 1 : 0;
  };
  const int total = per_split * split;
  const int q = total / kNumXcd, rq = total % kNumXcd;
  for (int x = 0; x < kNumXcd; ++x) {
    const int lo = x * q + std::min(x, rq), n = q + (x < rq ? 1 : 0);
    int w = lo;
    for (int c = 0; c < 4; ++c)
      for (int lin = lo; lin < lo + n; ++lin) {
        const int2 t = order[lin % per_split];
        if (cls(t) == c) out[w++] = make_int2(lin / per_split, t.x << 16 | t.y);
      }
  }
}

}  // namespace sdpx
