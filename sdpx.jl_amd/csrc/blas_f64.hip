// Bandwidth-bound FP64 kernels of the Newton step: the constraint-panel
// GEMVs (P = sum x_i A_i, v_i = <A_i, M>), deterministic two-stage
// reductions (dot products, sup-norms, Cholesky-diagonal extrema) and the
// elementwise block updates.  Every reduction sums in a fixed order, so two
// runs on the same inputs agree bitwise.
#include "blas_f64.hpp"

namespace sdpx {
namespace {

constexpr int RB = 256;       // threads per reduction block
constexpr int RMAXBLK = 512;  // stage-1 blocks (partials)

__device__ inline double nan_max(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  return a > b ? a : b;
}

template <typename Op>
__device__ inline double block_reduce(double v, Op op, double* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = op(r, sh[w]);
    sh[0] = r;
  }
  __syncthreads();
  const double r = sh[0];
  __syncthreads();
  return r;
}

struct SumOp {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct MaxOp {
  __device__ double operator()(double a, double b) const { return nan_max(a, b); }
};
struct MinOp {
  __device__ double operator()(double a, double b) const {
    if (a != a || b != b) return __builtin_nan("");
    return a < b ? a : b;
  }
};

// sum_i (x_i + tx dx_i) * (y_i + ty dy_i); dx / dy may be null
__global__ __launch_bounds__(RB) void dot_stage1(const double* x, const double* dx, double tx, const double* y,
                                                  const double* dy, double ty, long long n, double* part) {
  __shared__ double sh[RB / 64];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB) {
    const double u = dx ? x[i] + tx * dx[i] : x[i];
    const double w = dy ? y[i] + ty * dy[i] : y[i];
    acc += u * w;
  }
  acc = block_reduce(acc, SumOp{}, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(RB) void absmax_stage1(const double* x, long long n, double* part) {
  __shared__ double sh[RB / 64];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB)
    acc = nan_max(acc, fabs(x[i]));
  acc = block_reduce(acc, MaxOp{}, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <typename Op>
__global__ __launch_bounds__(RB) void reduce_stage2(const double* part, int nparts, double init, double* out) {
  __shared__ double sh[RB / 64];
  double acc = init;
  for (int i = threadIdx.x; i < nparts; i += RB) acc = Op{}(acc, part[i]);
  acc = block_reduce(acc, Op{}, sh);
  if (threadIdx.x == 0) *out = acc;
}

// min and max of |diag(L)| of `batch` n x n factors -> out[2*b], out[2*b+1]
__global__ __launch_bounds__(RB) void diag_extrema_kernel(const double* L, long long ld, long long stride, int n,
                                                           double* out) {
  __shared__ double sh[RB / 64];
  const int b = blockIdx.x;
  const double* a = L + (long long)b * stride;
  double mn = __builtin_inf(), mx = 0.0;
  for (int i = threadIdx.x; i < n; i += RB) {
    const double v = fabs(a[i + (long long)i * ld]);
    mn = MinOp{}(mn, v);
    mx = MaxOp{}(mx, v);
  }
  mn = block_reduce(mn, MinOp{}, sh);
  mx = block_reduce(mx, MaxOp{}, sh);
  if (threadIdx.x == 0) {
    out[2 * b] = mn;
    out[2 * b + 1] = mx;
  }
}

// ---- constraint panel GEMV: out = alpha * Av x + b1 * y1 + b2 * y2 --------
__global__ __launch_bounds__(256) void avx_partial_kernel(const double* __restrict__ Av, long long K2, int m,
                                                          const double* __restrict__ x, int cols, double* part) {
  __shared__ double xs[256];
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * cols;
  const int c1 = min(m, c0 + cols);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int cb = c0; cb < c1; cb += 256) {
    const int nc = min(256, c1 - cb);
    __syncthreads();
    if ((int)threadIdx.x < nc) xs[threadIdx.x] = x[cb + threadIdx.x];
    __syncthreads();
    if (r < K2) {
      const double* p = Av + r + (long long)cb * K2;
      int c = 0;
      for (; c + 4 <= nc; c += 4) {
        a0 += p[(long long)c * K2] * xs[c];
        a1 += p[(long long)(c + 1) * K2] * xs[c + 1];
        a2 += p[(long long)(c + 2) * K2] * xs[c + 2];
        a3 += p[(long long)(c + 3) * K2] * xs[c + 3];
      }
      for (; c < nc; ++c) a0 += p[(long long)c * K2] * xs[c];
    }
  }
  if (r < K2) part[(long long)blockIdx.y * K2 + r] = (a0 + a1) + (a2 + a3);
}

// Same partial sums with two rows per thread (16-byte fetches): needs an even
// K2 and a 16-byte aligned Av.
__global__ __launch_bounds__(256) void avx_partial2_kernel(const double* __restrict__ Av, long long K2, int m,
                                                           const double* __restrict__ x, int cols, double* part) {
  __shared__ double xs[256];
  const long long K2h = K2 / 2;
  const long long r2 = (long long)blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * cols;
  const int c1 = min(m, c0 + cols);
  double s0[2] = {0.0, 0.0}, s1[2] = {0.0, 0.0}, s2[2] = {0.0, 0.0}, s3[2] = {0.0, 0.0};
  for (int cb = c0; cb < c1; cb += 256) {
    const int nc = min(256, c1 - cb);
    __syncthreads();
    if ((int)threadIdx.x < nc) xs[threadIdx.x] = x[cb + threadIdx.x];
    __syncthreads();
    if (r2 < K2h) {
      const double2* p = reinterpret_cast<const double2*>(Av + (long long)cb * K2) + r2;
      int c = 0;
      for (; c + 4 <= nc; c += 4) {
        const double2 v0 = p[(long long)c * K2h], v1 = p[(long long)(c + 1) * K2h];
        const double2 v2 = p[(long long)(c + 2) * K2h], v3 = p[(long long)(c + 3) * K2h];
        s0[0] += v0.x * xs[c];
        s0[1] += v0.y * xs[c];
        s1[0] += v1.x * xs[c + 1];
        s1[1] += v1.y * xs[c + 1];
        s2[0] += v2.x * xs[c + 2];
        s2[1] += v2.y * xs[c + 2];
        s3[0] += v3.x * xs[c + 3];
        s3[1] += v3.y * xs[c + 3];
      }
      for (; c < nc; ++c) {
        const double2 v0 = p[(long long)c * K2h];
        s0[0] += v0.x * xs[c];
        s0[1] += v0.y * xs[c];
      }
    }
  }
  if (r2 < K2h) {
    double* o = part + (long long)blockIdx.y * K2 + 2 * r2;
    o[0] = (s0[0] + s1[0]) + (s2[0] + s3[0]);
    o[1] = (s0[1] + s1[1]) + (s2[1] + s3[1]);
  }
}

__global__ __launch_bounds__(256) void avx_combine_kernel(const double* part, int splits, long long K2, double alpha,
                                                          double* out, double b1, const double* y1, double b2,
                                                          const double* y2) {
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= K2) return;
  double s = 0.0;
  for (int i = 0; i < splits; ++i) s += part[(long long)i * K2 + r];
  double v = alpha * s;
  if (y1) v += b1 * y1[r];
  if (y2) v += b2 * y2[r];
  out[r] = v;
}

// ---- constraint panel transposed GEMV: v_i = beta v_i + alpha <A_i, M> ------
__global__ __launch_bounds__(256) void avt_kernel(const double* __restrict__ Av, long long K2, int m,
                                                  const double* __restrict__ M, double alpha, double beta, double* v) {
  __shared__ double sh[4];
  const int i = blockIdx.x;
  const double* col = Av + (long long)i * K2;
  double a0 = 0.0, a1 = 0.0;
  long long r = threadIdx.x;
  for (; r + 256 < K2; r += 512) {
    a0 += col[r] * M[r];
    a1 += col[r + 256] * M[r + 256];
  }
  if (r < K2) a0 += col[r] * M[r];
  double acc = block_reduce(a0 + a1, SumOp{}, sh);
  if (threadIdx.x == 0) v[i] = (beta == 0.0) ? alpha * acc : beta * v[i] + alpha * acc;
}

// Four constraints per work-group: each 16-byte fetch of M serves four 16-byte
// fetches of the panel (M is re-read m/4 times from L2 instead of m times).
// Needs an even K2 and 16-byte aligned Av and M.
constexpr int kAvtCols = 4;
__global__ __launch_bounds__(256) void avt4_kernel(const double* __restrict__ Av, long long K2, int m,
                                                   const double* __restrict__ M, double alpha, double beta, double* v) {
  __shared__ double sh[4];
  const int i0 = blockIdx.x * kAvtCols;
  const int nc = min(kAvtCols, m - i0);
  const long long K2h = K2 / 2;
  const double2* Mv = reinterpret_cast<const double2*>(M);
  const double2* cv[kAvtCols];
#pragma unroll
  for (int c = 0; c < kAvtCols; ++c)
    cv[c] = reinterpret_cast<const double2*>(Av + (long long)(i0 + (c < nc ? c : 0)) * K2);
  double acc[kAvtCols] = {};
#pragma unroll 2
  for (long long r = threadIdx.x; r < K2h; r += 256) {
    const double2 w = Mv[r];
#pragma unroll
    for (int c = 0; c < kAvtCols; ++c) {
      const double2 a = cv[c][r];
      acc[c] += a.x * w.x + a.y * w.y;
    }
  }
#pragma unroll
  for (int c = 0; c < kAvtCols; ++c) {
    const double t = block_reduce(acc[c], SumOp{}, sh);
    if (threadIdx.x == 0 && c < nc) v[i0 + c] = (beta == 0.0) ? alpha * t : beta * v[i0 + c] + alpha * t;
  }
}

// ---- elementwise --------------------------------------------------------
__global__ void axpbyz_kernel(long long n, double a, const double* x, double b, const double* y, double* z) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    z[i] = (y ? a * x[i] + b * y[i] : a * x[i]);
}

__global__ void fill_kernel(long long n, double v, double* z) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) z[i] = v;
}

__global__ void add_diag_kernel(double* A, int n, long long ld, double v, const double* vdev, double vdev_scale) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) A[i + (long long)i * ld] += (vdev ? vdev_scale * (*vdev) : v);
}

__global__ void symmetrize_kernel(double* A, int n) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r <= c) return;  // strictly lower thread owns the pair
  const double avg = (A[r + (long long)c * n] + A[c + (long long)r * n]) / 2.0;
  A[r + (long long)c * n] = avg;
  A[c + (long long)r * n] = avg;
}

__global__ void zero_upper_kernel(double* A, int n, long long ld) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r < c) A[r + (long long)c * ld] = 0.0;
}

// dst lower = src lower (+ shift * max(|src_ii|, 1) on the diagonal)
__global__ void copy_lower_shift_kernel(const double* src, double* dst, int n, double shift) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r < c) return;
  double v = src[idx];
  if (r == c && shift != 0.0) v += shift * fmax(fabs(v), 1.0);
  dst[idx] = v;
}

// y = alpha * S x + beta * y using only the lower triangle of S (n x n, ld n).
// Tile (bi, bj), bi >= bj, of TS rows: its row contribution goes to part[bj][rows of bi]
// and, off the diagonal, its transposed contribution to part[bi][rows of bj];
// every (slot, block) pair is written exactly once and summed in slot order.
constexpr int TS = 64;
__global__ __launch_bounds__(256) void symv_tile_kernel(const double* __restrict__ S, int n,
                                                        const double* __restrict__ x, double* part) {
  const int nb = (n + TS - 1) / TS;
  const int tile = blockIdx.x;
  int bi = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= tile) ++bi;
  while (bi * (bi + 1) / 2 > tile) --bi;
  const int bj = tile - bi * (bi + 1) / 2;
  __shared__ double t[TS][TS + 1];
  __shared__ double xi[TS], xj[TS];
  const int r0 = bi * TS, c0 = bj * TS;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < TS * TS; idx += 256) {
    const int rr = idx % TS, cc = idx / TS;
    const int gr = r0 + rr, gc = c0 + cc;
    double v = 0.0;
    if (gr < n && gc < n && gr >= gc) v = S[gr + (long long)gc * n];
    t[rr][cc] = v;
  }
  if (tid < TS) {
    xi[tid] = (r0 + tid < n) ? x[r0 + tid] : 0.0;
    xj[tid] = (c0 + tid < n) ? x[c0 + tid] : 0.0;
  }
  __syncthreads();
  if (tid < TS) {
    // row contribution: (T xj)[tid]
    double acc = 0.0;
    if (bi != bj) {
      for (int cc = 0; cc < TS; ++cc) acc += t[tid][cc] * xj[cc];
    } else {
      // diagonal tile: mirror the strict lower part
      for (int cc = 0; cc < TS; ++cc) acc += (tid >= cc ? t[tid][cc] : t[cc][tid]) * xj[cc];
    }
    if (r0 + tid < n) part[(long long)bj * n + r0 + tid] = acc;
  } else if (tid < 2 * TS && bi != bj) {
    // transposed contribution: (T' xi)[c]
    const int cc = tid - TS;
    double acc = 0.0;
    for (int rr = 0; rr < TS; ++rr) acc += t[rr][cc] * xi[rr];
    if (c0 + cc < n) part[(long long)bi * n + c0 + cc] = acc;
  }
  (void)nb;
}

__global__ void symv_combine_kernel(const double* part, int n, int nb, double alpha, double beta, const double* y0,
                                    double* y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int slot = 0; slot < nb; ++slot) s += part[(long long)slot * n + i];
  y[i] = (y0 ? beta * y0[i] : 0.0) + alpha * s;
}

}  // namespace

int dot_f64(const double* x, const double* dx, double tx, const double* y, const double* dy, double ty, long long n,
            double* part, double* out, hipStream_t stream) {
  const int blocks = std::max(1, std::min(RMAXBLK, ceil_div(n, RB * 4)));
  hipLaunchKernelGGL(dot_stage1, dim3(blocks), dim3(RB), 0, stream, x, dx, tx, y, dy, ty, n, part);
  hipLaunchKernelGGL(reduce_stage2<SumOp>, dim3(1), dim3(RB), 0, stream, part, blocks, 0.0, out);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int absmax_f64(const double* x, long long n, double* part, double* out, hipStream_t stream) {
  if (n <= 0) {
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(256), 0, stream, 1LL, 0.0, out);
    SDPX_HIP(hipGetLastError());
    return kOk;
  }
  const int blocks = std::max(1, std::min(RMAXBLK, ceil_div(n, RB * 4)));
  hipLaunchKernelGGL(absmax_stage1, dim3(blocks), dim3(RB), 0, stream, x, n, part);
  hipLaunchKernelGGL(reduce_stage2<MaxOp>, dim3(1), dim3(RB), 0, stream, part, blocks, 0.0, out);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int diag_extrema_f64(const double* L, long long ld, long long stride, int n, int batch, double* out,
                     hipStream_t stream) {
  hipLaunchKernelGGL(diag_extrema_kernel, dim3(batch), dim3(RB), 0, stream, L, ld, stride, n, out);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int avx_f64(const double* Av, long long K2, int m, const double* x, double alpha, double* out, double b1,
            const double* y1, double b2, const double* y2, double* part, long long part_cap, hipStream_t stream) {
  const int rb = ceil_div(K2, 256);
  const bool vec = K2 % 2 == 0 && (reinterpret_cast<uintptr_t>(Av) & 15u) == 0;
  const int pb = vec ? ceil_div(K2 / 2, 256) : rb;  // work-groups along the rows
  int splits = std::max(1, std::min(m, ceil_div(2048, pb)));
  while ((long long)splits * K2 > part_cap && splits > 1) splits = (splits + 1) / 2;
  const int cols = ceil_div(m, splits);
  splits = ceil_div(m, cols);
  if (m > 0) {
    if (vec)
      hipLaunchKernelGGL(avx_partial2_kernel, dim3(pb, splits), dim3(256), 0, stream, Av, K2, m, x, cols, part);
    else
      hipLaunchKernelGGL(avx_partial_kernel, dim3(rb, splits), dim3(256), 0, stream, Av, K2, m, x, cols, part);
  } else {
    splits = 0;
  }
  hipLaunchKernelGGL(avx_combine_kernel, dim3(rb), dim3(256), 0, stream, part, splits, K2, alpha, out, b1, y1, b2,
                     y2);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int avt_f64(const double* Av, long long K2, int m, const double* M, double alpha, double beta, double* v,
            hipStream_t stream) {
  if (m <= 0) return kOk;
  const bool vec = K2 % 2 == 0 && (reinterpret_cast<uintptr_t>(Av) & 15u) == 0 &&
                   (reinterpret_cast<uintptr_t>(M) & 15u) == 0;
  if (vec)
    hipLaunchKernelGGL(avt4_kernel, dim3(ceil_div(m, kAvtCols)), dim3(256), 0, stream, Av, K2, m, M, alpha, beta, v);
  else
    hipLaunchKernelGGL(avt_kernel, dim3(m), dim3(256), 0, stream, Av, K2, m, M, alpha, beta, v);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int axpbyz_f64(long long n, double a, const double* x, double b, const double* y, double* z, hipStream_t stream) {
  if (n <= 0) return kOk;
  const int blocks = std::min(4096, ceil_div(n, 256));
  hipLaunchKernelGGL(axpbyz_kernel, dim3(blocks), dim3(256), 0, stream, n, a, x, b, y, z);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int fill_f64(long long n, double v, double* z, hipStream_t stream) {
  if (n <= 0) return kOk;
  const int blocks = std::min(4096, ceil_div(n, 256));
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, stream, n, v, z);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int add_diag_f64(double* A, int n, long long ld, double v, const double* vdev, double vdev_scale, hipStream_t stream) {
  hipLaunchKernelGGL(add_diag_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, A, n, ld, v, vdev, vdev_scale);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int symmetrize_f64(double* A, int n, hipStream_t stream) {
  hipLaunchKernelGGL(symmetrize_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, stream, A, n);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int zero_upper_f64(double* A, int n, long long ld, hipStream_t stream) {
  hipLaunchKernelGGL(zero_upper_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, stream, A, n, ld);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int copy_lower_shift_f64(const double* src, double* dst, int n, double shift, hipStream_t stream) {
  hipLaunchKernelGGL(copy_lower_shift_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, stream, src, dst,
                     n, shift);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

int symv_lower_f64(const double* S, int n, const double* x, double alpha, double beta, const double* y0, double* y,
                   double* part, hipStream_t stream) {
  const int nb = ceil_div(n, TS);
  const int tiles = nb * (nb + 1) / 2;
  hipLaunchKernelGGL(symv_tile_kernel, dim3(tiles), dim3(256), 0, stream, S, n, x, part);
  hipLaunchKernelGGL(symv_combine_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, part, n, nb, alpha, beta, y0,
                     y);
  SDPX_HIP(hipGetLastError());
  return kOk;
}

}  // namespace sdpx
