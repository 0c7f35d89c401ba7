// Factor of one nb x nb diagonal block in registers (device code shared by the
// Cholesky kernels of chol_f64.hip and the fused trailing-update kernel of
// gemm_f64.hip).
#pragma once

#include "common.hpp"

namespace sdpx {
namespace chol {

constexpr int NB = kCholNb;

// Factors the jb x jb diagonal block at a (column j of its matrix), writes L
// over its lower triangle and the block's inverse to di (NB x NB, zero
// padded); a bad pivot sets *info = j + c + 1.
//
// Four wavefronts in lock step, state in registers (loops unrolled over
// compile-time indices; the block is padded to the identity beyond jb):
//  * waves 0, 1 (lane r = row r of the block, 32 columns each): right-looking
//    elimination on the unscaled columns, a_rg -= (a_rc / p) a_gc for g > c;
//  * waves 2, 3 (lane q = column q of M = L_u^-1 where A = L_u D L_u', 32 rows
//    each): the same row operations applied to the identity,
//    M_gq -= (M_cq / p) a_gc for g > c.
// Step c: the wave holding column c publishes it, the wave holding row c of M
// publishes that row (LDS, double buffered: one barrier per step); every wave
// reads its 32 entries of the column as broadcasts in one batch and updates.
// Column c of L = L_u D^1/2 and row c of L^-1 = D^-1/2 M are final at step c
// and leave the registers there.
constexpr int kDiagThreads = 4 * NB;
constexpr int kDiagHalf = NB / 2;
static_assert(NB == 64, "one lane per row / column of the diagonal block");

// The wave holding entries [H0, H0 + kDiagHalf) (columns of A for the
// elimination waves, rows of M for the inverse waves); `elim` is wave-uniform.
// Every loop is fully unrolled, so the conditions on c fold at compile time.
// Returns the failing step (first non-positive pivot), or -1.
template <int H0>
__device__ __forceinline__ int diag_wave(double* a, long long lda, int jb, bool elim, int lane, double (*col)[NB],
                                         double (*mrow)[NB], double (*s)[NB + 2]) {
  const bool live = lane < jb;
  double v[kDiagHalf];
#pragma unroll
  for (int i = 0; i < kDiagHalf; ++i) {
    const int g = H0 + i;
    v[i] = (elim && live && g <= lane) ? a[lane + (long long)g * lda] : (g == lane ? 1.0 : 0.0);
  }
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    const bool mine = c >= H0 && c < H0 + kDiagHalf;
    double* cc = col[c & 1];
    double* mr = mrow[c & 1];
    if (mine) {
      if (elim) cc[lane] = v[c - H0];
      else mr[lane] = v[c - H0];
    }
    __syncthreads();
    const double p = cc[c];
    const double own = elim ? cc[lane] : mr[lane];  // a_rc (elimination) or M_cq (inverse)
    // use `own` here so that its LDS read goes out with the pivot's (one round
    // trip) instead of after the pivot test: 10% of the block's latency
    asm volatile("" ::"v"(own));
    if (!(p > 0.0 && p <= 1.7976931348623157e308)) return c;  // not positive or not finite; uniform
    const double f = own / p;
    if (H0 + kDiagHalf - 1 > c) {
      const int lo = c + 1 > H0 ? c + 1 : H0;
      double cv[kDiagHalf];
#pragma unroll
      for (int g = lo & ~1; g < H0 + kDiagHalf; g += 2) {
        const double2 t = reinterpret_cast<const double2*>(cc)[g / 2];
        cv[g - H0] = t.x;
        cv[g + 1 - H0] = t.y;
      }
      __builtin_amdgcn_sched_group_barrier(0x100, kDiagHalf / 2, 0);
#pragma unroll
      for (int g = lo; g < H0 + kDiagHalf; ++g) v[g - H0] -= f * cv[g - H0];
    }
    // column c of A and row c of M are final: column c of L = L_u D^1/2 and
    // row c of L^-1 = D^-1/2 M leave the registers here
    if (mine) {
      const double d = sqrt(p);
      if (elim) {
        if (live && lane >= c && c < jb) a[lane + (long long)c * lda] = lane == c ? d : own / d;
      } else {
        s[lane][c] = own / d;  // (L^-1)_cq, transposed for the coalesced store of di
      }
    }
  }
  return -1;
}

// LDS (doubles, 16-byte aligned) the diagonal-block body works in.
constexpr int kDiagLds = 4 * NB + NB * (NB + 2);

__device__ __forceinline__ void potrf_diag_body(double* a, long long lda, int j, int jb, double* di, int* info,
                                                double* lds) {
  double(*col)[NB] = reinterpret_cast<double(*)[NB]>(lds);
  double(*mrow)[NB] = reinterpret_cast<double(*)[NB]>(lds + 2 * NB);
  double(*s)[NB + 2] = reinterpret_cast<double(*)[NB + 2]>(lds + 4 * NB);  // s[q][c] = (L^-1)_cq
  const int wave = threadIdx.x / NB;
  const int lane = threadIdx.x & (NB - 1);
  const bool elim = wave < 2;
  const int bad = (wave & 1) ? diag_wave<kDiagHalf>(a, lda, jb, elim, lane, col, mrow, s)
                             : diag_wave<0>(a, lda, jb, elim, lane, col, mrow, s);
  if (bad >= 0) {
    if (threadIdx.x == 0) *info = j + bad + 1;
    return;
  }
  __syncthreads();
  for (int c = wave; c < NB; c += kDiagThreads / NB)
    di[lane + (long long)c * NB] = (lane < jb && c < jb) ? s[c][lane] : 0.0;
}

}  // namespace chol
}  // namespace sdpx
