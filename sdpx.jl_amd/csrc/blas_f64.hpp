// Declarations of the bandwidth-bound FP64 kernels (blas_f64.hip).
#pragma once
#include "common.hpp"

namespace sdpx {

// out = sum_i (x_i + tx dx_i)(y_i + ty dy_i); part needs 512 doubles
int dot_f64(const double* x, const double* dx, double tx, const double* y, const double* dy, double ty, long long n,
            double* part, double* out, hipStream_t stream);
// out = max_i |x_i| (NaN if any entry is NaN); part needs 512 doubles
int absmax_f64(const double* x, long long n, double* part, double* out, hipStream_t stream);
// out[2b], out[2b+1] = min, max |diag| of factor b
int diag_extrema_f64(const double* L, long long ld, long long stride, int n, int batch, double* out,
                     hipStream_t stream);
// out = alpha * Av x + b1 y1 + b2 y2 (Av: K2 x m column-major); part holds up to part_cap doubles
int avx_f64(const double* Av, long long K2, int m, const double* x, double alpha, double* out, double b1,
            const double* y1, double b2, const double* y2, double* part, long long part_cap, hipStream_t stream);
// v_i = beta v_i + alpha <Av[:, i], M>
int avt_f64(const double* Av, long long K2, int m, const double* M, double alpha, double beta, double* v,
            hipStream_t stream);
// z = a x + b y (y may be null: z = a x)
int axpbyz_f64(long long n, double a, const double* x, double b, const double* y, double* z, hipStream_t stream);
int fill_f64(long long n, double v, double* z, hipStream_t stream);
// A_ii += v   (or vdev_scale * *vdev when vdev is non-null)
int add_diag_f64(double* A, int n, long long ld, double v, const double* vdev, double vdev_scale, hipStream_t stream);
int symmetrize_f64(double* A, int n, hipStream_t stream);
// strict upper triangle of the n x n matrix A := 0
int zero_upper_f64(double* A, int n, long long ld, hipStream_t stream);
// dst lower = src lower, diagonal += shift * max(|src_ii|, 1)
int copy_lower_shift_f64(const double* src, double* dst, int n, double shift, hipStream_t stream);
// y = alpha S x + beta y0 from the lower triangle of S; part needs ceil(n/64) * n doubles
int symv_lower_f64(const double* S, int n, const double* x, double alpha, double beta, const double* y0, double* y,
                   double* part, hipStream_t stream);

}  // namespace sdpx
