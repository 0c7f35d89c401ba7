// Shared definitions for the SDPX MI355X (gfx950) Newton-step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <string>

namespace sdpx {

// Status codes returned by every C-ABI entry point (mirrored in include/sdpx_mi355x.h).
enum Status : int {
  kOk = 0,
  kNotPositiveDefinite = 1,
  kNonFinite = 2,
  kInvalidArgument = -1,
  kHipError = -2,
  kOutOfMemory = -3,
};

void set_last_error(const std::string& msg);
const char* last_error();

#define SDPX_HIP(expr)                                                          \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::sdpx::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
      return ::sdpx::kHipError;                                                 \
    }                                                                           \
  } while (0)

#define SDPX_CHECK(expr)          \
  do {                            \
    int _s = (expr);              \
    if (_s != ::sdpx::kOk) return _s; \
  } while (0)

constexpr int kWave = 64;
constexpr int kNumXcd = 8;  // MI355X: 256 CUs in 8 XCDs, blocks dispatched round-robin

inline int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// FP64 GEMM on MFMA (gemm_f64.hip)
//   C(i,j) = alpha * sum_kk op(A)(i,kk) op(B)(kk,j) + beta * C(i,j)
// column-major; op(A)(i,kk) = TA ? A[kk + i*lda] : A[i + kk*lda]
//               op(B)(kk,j) = TB ? B[j + kk*ldb] : B[kk + j*ldb]
// ---------------------------------------------------------------------------
enum GemmFlags : unsigned {
  kGemmLowerOut = 1u << 0,  // only C(i,j) with i >= j is computed / written (M == N)
  // Triangular hints only shorten the K range of a tile (to BK granularity);
  // the operand must hold explicit zeros in its other triangle.
  kGemmALower = 1u << 1,    // op(A) is lower triangular: kk <= i
  kGemmBLower = 1u << 2,    // op(B) is lower triangular: kk >= j
  kGemmAUpper = 1u << 3,    // op(A) is upper triangular: kk >= i
  // Row tiles start at -(mt*BM - M) so that the partial tile is the first one
  // (with kGemmALower it then has the shortest K range); no lower output.
  kGemmRaggedFirst = 1u << 4,
  kGemmSmallTiles = 1u << 5,  // force 64 x 64 tiles (more resident groups for short K ranges)
  kGemmTallTiles = 1u << 6,   // 128 x 64 tiles (NN and TN products only)
  kGemmWideTiles = 1u << 7,   // 64 x 128 tiles (NN and TN products only)
  kGemmTransA = 1u << 8,    // op(A) = A'
  kGemmTransB = 1u << 9,    // op(B) = B'
  kGemmSchurGram = 1u << 10,  // route to the dedicated Schur Gram kernel (TN, 128 x 128, lower output)
};

struct GemmArgs {
  int M, N, K;
  double alpha, beta;
  const double* A;
  long long lda, strideA;
  const double* B;
  long long ldb, strideB;
  double* C;
  long long ldc, strideC;
  int batch;
  unsigned flags;
  const int* skip;        // optional per-batch flag: skip the batch entry when *skip != 0
  long long skip_stride;  // stride between per-batch skip flags
  const int2* tile_order; // optional (ti, tj) visiting order of the lower tiles (kGemmLowerOut)
  // optional Schur Gram work list (kGemmSchurGram), one entry per work item in
  // the XCD-remapped launch order: x = split index, y = ti << 16 | tj
  const int2* item_order;
  // Optional blocked output: when block_rows > 0, C(row, col) is stored at
  // C + row % block_rows + (row / block_rows) * block_stride + col * ldc
  // (the rows of a vertically stacked product go back to separate blocks).
  long long block_rows, block_stride;
  int row_off;            // set by the launcher for kGemmRaggedFirst
};

// Grouped launches: many independent GEMMs of different shapes in one
// launch, one work-group per (GEMM, output tile) item.  Flags as above
// (LowerOut and the triangular K-range hints; no split-K).
struct GroupGemm {
  const double* A;
  const double* B;
  double* C;
  long long lda, ldb, ldc;
  int M, N, K;
  unsigned flags;
  double alpha, beta;
  const int* skip;  // optional: the GEMM is skipped when *skip != 0
};
struct GroupTile {
  int g, ti, tj, pad;
};
// 64 x 64 tiles; ta / tb select op(A) = A' / op(B) = B'; vec: every GEMM meets
// the 16-byte-load conditions (even leading dimensions, 16-byte aligned bases,
// even M (ta=0) or K (ta=1) for A and even K (tb=0) or N (tb=1) for B).
int gemm_f64_grouped(const GroupGemm* groups, const GroupTile* tiles, int ntiles, bool ta, bool tb, bool vec,
                     hipStream_t stream);

// Lower-tile visiting order in square super-blocks of `group` tile rows and
// columns (row-major inside), so that co-resident work-groups share panels.
void lower_tile_order(int nt, int group, int2* out);
// Work list of the FP64 Schur Gram launch for an m x m lower output on 128-row
// tiles, `group` super-blocks and `split` K ranges (nt(nt+1)/2 * split entries,
// GemmArgs::item_order layout).  Each XCD's contiguous run of items lists its
// full tiles first and the shorter items (diagonal tiles, strips of a short last
// tile row, groups that exit) last, so that the co-resident groups of an XCD,
// all of one length, start and finish together and share Phat panels in L2.
void gram_item_order(int m, int group, int split, int2* out);

// Launches the GEMM on `stream`. `split_k` > 1 requires `workspace` of
// split_k * M * N * batch doubles; partial sums are reduced in a fixed order
// (deterministic). Returns kOk / kHipError.
int gemm_f64(const GemmArgs& g, hipStream_t stream, int split_k = 1, double* workspace = nullptr,
             hipEvent_t after_main = nullptr);

// ---------------------------------------------------------------------------
// FP32 Schur Gram product (gemm_f32.hip): S (lower, FP64) = P' P + beta S for
// an FP32 panel P (m columns of K floats, leading dimension ldp % 4 == 0,
// 16-byte aligned); split-K partials in ws (split * m * m doubles).
int gram_f32(const float* P, long long ldp, int m, int K, double* S, long long lds, double beta, int split,
             double* ws, const int2* tile_order, hipStream_t stream, hipEvent_t after_main = nullptr);
// BF16x3 Schur Gram product (gemm_f32.hip): S (lower, FP64) = P' P + beta S
// with P ~= H + L given as two bf16 planes (ld ldp, Kp rows, both multiples of
// kBf16Align, padding rows zero), three bf16 MFMA products H'H + H'L + L'H per step.
constexpr int kBf16Align = 64;  // the deepest K slice of the kernel divides it
int gram_bf16x3(const __bf16* H, const __bf16* L, long long ldp, int m, int Kp, double* S, long long lds, double beta,
                int split, double* ws, const int2* tile_order, hipStream_t stream, hipEvent_t after_main = nullptr);
// H = bf16(src), L = bf16(src - H) (ld ld_dst, rows >= `rows` zeroed), rows x cols
int split_panel_bf16(const double* src, long long ld_src, __bf16* H, __bf16* L, long long ld_dst, long long rows,
                     int cols, hipStream_t stream);
// Grouped FP32 Schur Gram products (gemm_f32.hip): for each listed GEMM g,
// C_g (lower, FP64, leading dim ldc) = P_g' P_g over 64 x 64 lower tiles (the
// GroupTile list: g, ti, tj), P_g an FP32 panel of M columns of K floats
// (K and ldp multiples of 4, rows K..ldp-1 unread, 16-byte aligned base).
// A GEMM whose *skip != 0 is skipped.
struct GroupGramF32 {
  const float* P;
  double* C;
  long long ldp, ldc;
  int M, K;
  const int* skip;
};
int gram_f32_grouped(const GroupGramF32* groups, const GroupTile* tiles, int ntiles, hipStream_t stream);
// FP32 P-hat chain for the BF16x3 assembly (panel_f32.hip): P-hat_i = L_X^-1 A_i M_Y from FP32
// operands (Av32: k^2 x m, MY32 / LXinv32: k x k lower with explicit zeros above, all ld k),
// T32 (k^2 x m floats) scratch, written as bf16 planes H = bf16(P-hat), L = bf16(P-hat - H)
// (ld ldo >= k^2, rows k^2 .. ldo - 1 zeroed); k % 4 == 0.
int panel_chain_bf16(const float* Av32, const float* MY32, const float* LXinv32, int k, int m, float* T32,
                     __bf16* H, __bf16* L, long long ldo, hipStream_t stream);
// dst (ld_dst >= rows, padding rows zeroed) = (float) src (ld_src), rows x cols
int convert_panel_f32(const double* src, long long ld_src, float* dst, long long ld_dst, long long rows, int cols,
                      hipStream_t stream);

// ---------------------------------------------------------------------------
// Dense Cholesky and triangular solves (chol_f64.hip)
// ---------------------------------------------------------------------------
constexpr int kCholNb = 64;

// Batched in-place lower Cholesky of `batch` n x n matrices (column-major,
// leading dim lda, stride between matrices `stride`). The inverse of every
// diagonal nb x nb factor block is written to dinv (batch * ceil(n/nb) * nb*nb
// doubles, stride dinv_stride per matrix). info[b] (device) is set to the
// 1-based column of the first non-positive / non-finite pivot, 0 on success.
int potrf_f64(double* A, long long lda, long long stride, int n, int batch, double* dinv,
              long long dinv_stride, int* info, hipStream_t stream);

// One matrix of a grouped Cholesky: n x n lower at A (leading dim lda), its
// diagonal-block inverses (ceil(n/nb) * nb*nb doubles) and its info flag.
struct CholDesc {
  double* A;
  long long lda;
  int n;
  double* dinv;
  int* info;
};
// Diagonal-block step j of the grouped Cholesky for matrices d[list[0..count)]
// (device arrays); matrices with *info != 0 or n <= j are left alone.  The
// panel and trailing updates of the step are grouped GEMMs built by the caller.
int potrf_diag_grouped(const CholDesc* d, const int* list, int count, int j, hipStream_t stream);
// Trailing update t (A22 -= A21 A21', NT, lower output, batch entries skipped
// when info != 0) fused with the factor of the next diagonal block: jbn x jbn
// at column jn of each matrix (the top-left block of t's output), whose
// inverse goes to dinv (block jn / nb of each entry's dinv_stride slice).
int potrf_trailing_diag(const GemmArgs& t, int jn, int jbn, double* dinv, long long dinv_stride, int* info,
                        hipStream_t stream);

// Solves L x = b (trans=0) or L' x = b (trans=1) in place for one right-hand
// side, using the diagonal-block inverses written by potrf_f64.
int trsv_f64(const double* L, long long lda, int n, const double* dinv, double* x, int trans,
             hipStream_t stream);
// The same solve with one launch per diagonal block (block inverse and the
// update of the other rows fused); y is an n-vector of scratch.
int trsv_f64_steps(const double* L, long long lda, int n, const double* dinv, double* x, double* y, int trans,
                   hipStream_t stream);

// FP32 Cholesky (chol_f32.hip): A (FP32, n x n, ld lda, lower) factored in place, the 64 x 64
// diagonal blocks in FP64 (their inverses into dinv, ceil(n/64) * 64 * 64 doubles; blk is a
// 64 x 64 FP64 scratch block, dinv32 a 64 x 64 FP32 one), panels and trailing updates on FP32
// MFMA; *info (device) = 1-based column of the first bad pivot, 0 on success.
int potrf_f32(float* A, long long lda, int n, double* blk, double* dinv, float* dinv32, int* info,
              hipStream_t stream);
// dst (FP32, ld ldf) = lower(src) (FP64, ld n) with diagonal += shift max(|src_ii|, 1)
int lower_to_f32(const double* src, int n, double shift, float* dst, long long ldf, hipStream_t stream);
// dst (FP64, ld n) = lower(src) (FP32, ld ldf)
int lower_to_f64(const float* src, long long ldf, int n, double* dst, hipStream_t stream);

// Inverse of a lower-triangular n x n factor (Linv written to out, ld n),
// blocked on the diagonal-block inverses written by potrf_f64.
int trtri_lower_f64(const double* L, long long lda, int n, const double* dinv, double* out,
                    double* work, hipStream_t stream);

}  // namespace sdpx
