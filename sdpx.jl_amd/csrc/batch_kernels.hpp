// Device side of the batched route: many independent single-block SDPs
// (block side k <= 64, m <= 2048 constraints each) resident in HBM.  Every
// kernel works on a compact list of member ids (one work-group per listed
// member, 256 threads) or on a precomputed work list; k x k blocks live in LDS.
#pragma once

#include "common.hpp"

namespace sdpx {
namespace batch {

constexpr int KMAX = 64;      // largest PSD block of a member
constexpr int MMAX = 2048;    // most constraints of a member
constexpr int LP = KMAX + 1;  // LDS pitch of a k x k block (column-major, conflict-light)
constexpr int NT = 256;       // threads per work-group

// per-member scalar slots (double), written by kernels, read by the host
enum Slot : int {
  S_PRES = 0,   // ||P||_inf
  S_DRES,       // ||d||_inf
  S_COMP,       // <X, Y>
  S_COMPT,      // <X + a dX, Y + b dY>
  S_LXMIN, S_LXMAX, S_LYMIN, S_LYMAX, S_SFMIN, S_SFMAX,  // |diag| extrema of L_X, L_Y, L_S
  S_RHO,        // ||r - S dx||_inf
  S_RNORM,      // ||r||_inf
  S_POBJ, S_DOBJ,
  S_XMAX, S_XBMAX, S_YBMAX,  // |x|_inf, |X|_inf, |Y|_inf after an update
  S_TX, S_TY,   // line-search results
  S_DACC,       // ||d||_inf after the accepted-step recurrence
  S_PSOL, S_DSOL,  // recomputed residuals of the returned point
  NSLOT = 24
};
// per-member integer slots
// I_SKIP: the member's FP64 Schur assembly is skipped; I_SKIP32: its FP32 assembly (mixed
// precision: the panel pass writes Phat in FP32 and the FP32 grouped Gram product reads it)
enum ISlot : int { I_INFOX = 0, I_INFOY, I_INFOS, I_SKIP, I_SKIP32, NISLOT = 5 };
// per-member parameters (double), written by the host
enum Param : int {
  P_MU = 0, P_A, P_B, P_PF, P_DF, P_GAMMA, P_MINSTEP, P_RATIO, P_SIGMA, P_MUG, P_MODE, P_BETA, P_REG, P_SCALE,
  P_AFF,  // 1: residual R without the mu I term (affine predictor)
  NPARAM = 16
};

struct Member {
  int k, m, nblk, kp;      // kp: k rounded up to 16
  long long oAv;           // Av: k^2 x m, column i = vec(A_i)
  long long oPh;           // Phat: ldP x m
  long long ldP;           // k^2 rounded up to even
  long long ldPf;          // k^2 rounded up to a multiple of 4: the FP32 Phat (floats at the same offset)
  long long oK;            // offset of this member's k x k blocks
  long long oM;            // offset of its m-vectors
  long long oS;            // offset of its m x m matrices
  long long oD;            // offset of its diagonal-block inverses (nblk * 64 * 64)
  long long oG;            // offset of its per-chunk Av x partial sums (ceil(m / kColChunk) * k^2)
};

constexpr int kColChunk = 64;  // constraints per work-group in the panel passes

// Device buffers of a batch (all offsets per Member).
struct Arrays {
  const Member* mem;
  double *Av, *Ph;
  double *C, *X, *Y, *LX, *LY, *LXinv, *Xinv, *P, *R, *dX, *dY, *W1, *W2, *bX, *bY, *G;  // k x k blocks
  double *c, *x, *dx, *dxb, *d, *v, *r, *rho, *delta, *bx, *ckx;                       // m-vectors
  double *ckX, *ckY;  // differentiation checkpoint (k x k blocks)
  double *S, *Sf, *dinv;
  double* gpart;   // Av x partial sums
  int max_chunks;  // most column chunks of any member
  int max_m;       // most constraints of any member (dynamic LDS of the residual pass)
  // Sparse copy of the constraints when they were given as COO (null otherwise):
  // constraint i of member p owns entries [sp_ptr[oM + i], sp_ptr[oM + i + 1]), each a
  // position r + c k of the k x k block (both mirror positions of an off-diagonal
  // entry, duplicates summed, positions unique within a constraint) and its value.
  const long long* sp_ptr;
  const int* sp_idx;
  const double* sp_val;
  // the same entries ordered by position: position e of member p owns entries
  // [sp_pptr[oK + e], sp_pptr[oK + e + 1]), each a constraint index and its value,
  // in increasing constraint order
  const long long* sp_pptr;
  const int* sp_pcon;
  const double* sp_pval;
  double* sc;   // B x NSLOT
  int* si;      // B x NISLOT
  double* par;  // B x NPARAM
};

// ---- launchers (batch_kernels.hip) ------------------------------------------
// mode 0: P = Av x - X - C, S_PRES;  1: dX = Av dx + P;  2: W1 = Av x - X - C, S_PSOL;  3: dX = Av dx
int launch_gemv(const Arrays& a, const int* list, int n, int mode, hipStream_t st);
// mode 0: d = c - Av' vec(Y), S_DRES;  1: r = -(d + Av' vec(W2));  2: v = c - Av' vec(Y), S_DSOL;
// 3: rho = r - Av' vec(W2), S_RHO = ||rho||_inf, S_RNORM = ||r||_inf
int launch_gemvt(const Arrays& a, const int* list, int n, int mode, hipStream_t st);
// R = -X Y (+ P_MU I unless P_AFF); L_X, L_Y, their |diag| extrema and info; L_X^-1, X^-1
int launch_block_factor(const Arrays& a, const int* list, int n, hipStream_t st);
// Phat_i = L_X^-1 A_i L_Y for the (member, first constraint) items of `work`: in FP64 (ldP) for
// members with I_SKIP == 0, in FP32 (ldPf, padding rows zeroed) for members with I_SKIP32 == 0
int launch_panel(const Arrays& a, const int2* work, const int* off, int chunk, hipStream_t st);
// Sf = lower(S) + P_REG * max(|S_ii|, 1) on the diagonal; info S reset
int launch_shift_copy(const Arrays& a, const int* list, int n, hipStream_t st);
// |diag| extrema of the Schur factor
int launch_factor_extrema(const Arrays& a, const int* list, int n, hipStream_t st);
// W2 = X^-1 dX Y (the matrix-free FP64 Schur operator: (S dx)_i = <A_i, X^-1 (sum_j dx_j A_j) Y>)
int launch_exact_block(const Arrays& a, const int* list, int n, hipStream_t st);
// W2 = X^-1 (P Y - R)
int launch_rhs_block(const Arrays& a, const int* list, int n, hipStream_t st);
// mode 0: dx = S^-1 r;  1: delta = S^-1 rho   (two triangular solves on Sf)
int launch_solve(const Arrays& a, const int* list, int n, int mode, hipStream_t st);
// rho = r - S dx, S_RHO, S_RNORM
int launch_residual(const Arrays& a, const int* list, int n, hipStream_t st);
// mode 0: dxb = dx;  1: dx += delta;  2: dx = dxb
int launch_vec_op(const Arrays& a, const int* list, int n, int mode, hipStream_t st);
// dY = sym(X^-1 (R - dX Y))
int launch_dir_block(const Arrays& a, const int* list, int n, hipStream_t st);
// S_COMP = <X,Y>, S_COMPT = <X + P_A dX, Y + P_B dY> (a negative P_A / P_B takes S_TX / S_TY)
int launch_comp(const Arrays& a, const int* list, int n, hipStream_t st);
// R = -X Y - dX dY + target I (P_MODE 0 fixed (P_MU, P_BETA), 1 local (P_SIGMA), 2 global (P_SIGMA, P_MUG))
int launch_corrector_block(const Arrays& a, const int* list, int n, hipStream_t st);
// fraction to boundary: S_TX = P_PF * bX (if bX < 1), S_TY likewise
// Size class of a member's line-search kernel (a bit of the launch mask): k <= 2, <= 16, <= 32, <= 48, <= 64.
inline unsigned step_class(int k) { return k <= 2 ? 1u : k <= 16 ? 2u : k <= 32 ? 4u : k <= 48 ? 8u : 16u; }
int launch_ftb(const Arrays& a, const int* list, int n, unsigned classes, hipStream_t st);
// backtracking with P_GAMMA, P_MINSTEP, P_RATIO: S_TX, S_TY
int launch_backtrack(const Arrays& a, const int* list, int n, unsigned classes, hipStream_t st);
// X += P_A dX, Y += P_B dY, x += P_A dx; S_COMP, S_XMAX, S_XBMAX, S_YBMAX
int launch_update(const Arrays& a, const int* list, int n, hipStream_t st);
// d = (1 - P_B) d - P_B rho, S_DACC
int launch_accepted(const Arrays& a, const int* list, int n, hipStream_t st);
// S_POBJ = c'x, S_DOBJ = <C, Y>
int launch_objectives(const Arrays& a, const int* list, int n, hipStream_t st);
// which 1: X *= P_SCALE, 2: Y *= P_SCALE
int launch_scale(const Arrays& a, const int* list, int n, int which, hipStream_t st);
// dir 0: (bx, bX, bY) = (x, X, Y);  1: the reverse;  2: (ckx, ckX, ckY) = (x, X, Y);  3: the reverse
int launch_best(const Arrays& a, const int* list, int n, int dir, hipStream_t st);
// adjoint: W = sym(X^-1 G Y) into W1 / W2 and d = -v (v holds h)
int launch_adj_w(const Arrays& a, const int* list, int n, hipStream_t st);
// adjoint: W2 = -(sym(X^-1 A_lambda Y) + W1), A_lambda in dX
int launch_adj_final(const Arrays& a, const int* list, int n, hipStream_t st);
// si[p * NISLOT + slot] = value for the listed members
int launch_set_int(const Arrays& a, const int* list, int n, int slot, int value, hipStream_t st);
// symmetric A_j densified from upper-triangle COO into Av (members' panels zeroed first)
int launch_coo_densify(const Arrays& a, const long long* cptr, const int* crow, const int* ccol, const double* cval,
                       const long long* cons_member, int ncons, hipStream_t st);

}  // namespace batch
}  // namespace sdpx
