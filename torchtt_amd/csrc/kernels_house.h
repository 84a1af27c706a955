// Householder QR steps on a column-major matrix in device memory (launch.h form), shared by the
// global residual's R-only sweep (residual.hip) and the QR of more columns than a workgroup's LDS
// holds (linalg_large.hip). Golub & Van Loan, Matrix Computations, 4th ed., Alg. 5.2.1: step j
// forms the reflector H_j = I - tau v v^T with v = x - alpha e_j, x = column j from row j down,
// alpha = -sign(x_j) ||x||, tau = 2 / (v^T v) = 1 / (||x||^2 - alpha x_j).
#pragma once
#include "ttamen_internal.h"
#include "kernels_blas1.h"

namespace ttamen {

// Sum over the 256 threads of a block in a fixed order; every thread gets the result.
__device__ __forceinline__ double hh_block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Householder step j on M (rows x n, column-major, leading dimension rows); block c works on
// column j + c. Writes row j of R (k x n row-major): zeros left of the diagonal, R[j][j] = alpha,
// R[j][c] = (H_j M)[j][c].
struct KHouseStep {
  struct Args { double* M; int64_t rows; int n; int j; double* R; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    __shared__ double sh[4];
    const int j = a.j, c = j + (int)bk.x;
    const int tid = threadIdx.x;
    const double* xj = a.M + (int64_t)j * a.rows;
    double p = 0.0;
    for (int64_t i = j + tid; i < a.rows; i += 256) p += xj[i] * xj[i];
    const double nrm2 = hh_block_sum(p, sh);
    const double x0 = xj[j];
    const double nrm = sqrt(nrm2);
    const double alpha = (x0 >= 0.0) ? -nrm : nrm;  // H x = alpha e_j, v = x - alpha e_j
    if (c == j) {  // row j of R left of the diagonal is zero; the next core's GEMM reads all of R
      for (int i = tid; i < j; i += 256) a.R[(int64_t)j * a.n + i] = 0.0;
      if (tid == 0) a.R[(int64_t)j * a.n + j] = (nrm2 > 0.0) ? alpha : 0.0;
      return;
    }
    double* xc = a.M + (int64_t)c * a.rows;
    if (!(nrm2 > 0.0)) {  // column j is zero below the diagonal: H = I
      if (tid == 0) a.R[(int64_t)j * a.n + c] = xc[j];
      return;
    }
    const double tau = 1.0 / (nrm2 - alpha * x0);  // 2 / (v^T v)
    double q = 0.0;
    for (int64_t i = j + tid; i < a.rows; i += 256) q += ((i == j) ? x0 - alpha : xj[i]) * xc[i];
    const double f = tau * hh_block_sum(q, sh);
    for (int64_t i = j + tid; i < a.rows; i += 256) {
      const double nv = xc[i] - f * ((i == j) ? x0 - alpha : xj[i]);
      xc[i] = nv;
      if (i == j) a.R[(int64_t)j * a.n + c] = nv;
    }
  }
};

// Backward accumulation of Q = H_0 H_1 ... H_{k-1} [I; 0] (Golub & Van Loan, Sec. 5.1.6): step j
// applies H_j to columns j .. k-1 of Y (rows x k, column-major), block c working on column j + c.
// The reflector is rebuilt from M after the KHouseStep sweep: column j of M still holds x (rows j
// and below; later steps never touch it), and alpha, tau follow from the same reduction, so H_j is
// the one the forward sweep applied. Columns left of j are zero from row j down, which H_j keeps.
struct KHouseQ {
  struct Args { const double* M; int64_t rows; int j; double* Y; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    __shared__ double sh[4];
    const int j = a.j, c = j + (int)bk.x;
    const int tid = threadIdx.x;
    const double* xj = a.M + (int64_t)j * a.rows;
    double p = 0.0;
    for (int64_t i = j + tid; i < a.rows; i += 256) p += xj[i] * xj[i];
    const double nrm2 = hh_block_sum(p, sh);
    if (!(nrm2 > 0.0)) return;  // H_j = I
    const double x0 = xj[j];
    const double nrm = sqrt(nrm2);
    const double alpha = (x0 >= 0.0) ? -nrm : nrm;
    const double tau = 1.0 / (nrm2 - alpha * x0);
    double* yc = a.Y + (int64_t)c * a.rows;
    double q = 0.0;
    for (int64_t i = j + tid; i < a.rows; i += 256) q += ((i == j) ? x0 - alpha : xj[i]) * yc[i];
    const double f = tau * hh_block_sum(q, sh);
    for (int64_t i = j + tid; i < a.rows; i += 256) yc[i] -= f * ((i == j) ? x0 - alpha : xj[i]);
  }
};

}  // namespace ttamen
