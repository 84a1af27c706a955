// Dense direct solver for the small AMEn local systems (order < max_full, 500 by
// default), replacing at::linalg_solve in the reference's direct branch
// (cpp/amen_solve.h:319-330).
//
// Blocked right-looking LU with partial pivoting on the augmented matrix [B | rhs]
// (n x (n+1), row-major), panel width 32:
//   panel kernel   one workgroup factors the n-j0 x 32 panel resident in LDS (pivot
//                  search, row swaps, rank-1 updates), writes it back with the pivots;
//   swap+trsm      one thread per trailing column (rhs included) applies the panel's
//                  row swaps and the unit-lower triangular solve with L11 (in LDS);
//   update         A22 -= L21 U12 on the matrix cores (the FP64 MFMA GEMM);
// then a blocked back substitution, one small kernel per block of 32 unknowns.
// The trailing matrix is read and written once per panel instead of once per column,
// and the update spreads over the whole chip.
#include "ttamen_internal.h"
#include <algorithm>

namespace ttamen {

constexpr int kLuNb = 32;

struct KLuPanel {
  struct Args { double* Aug; int n, ld, j0, jb; int* piv; };
  static constexpr int kThreads = 1024;
  static __device__ void run(const Args& a, const Blk&);
};
__device__ void KLuPanel::run(const Args& a, const Blk&) {
  double* Aug = a.Aug; const int n = a.n, ld = a.ld, j0 = a.j0, jb = a.jb; int* piv = a.piv;
  extern __shared__ double P[];  // column-major: P[c * rows + i]
  __shared__ double shv[16];
  __shared__ int shi[16];
  __shared__ int s_p;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int rows = n - j0;
  const int tot = rows * jb;
  for (int t0 = tid; t0 < tot; t0 += 8 * nt) {  // 8 independent loads in flight per thread
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u * nt;
      const int c = t / rows, i = t - c * rows;
      v[u] = (t < tot) ? Aug[(int64_t)(j0 + i) * ld + j0 + c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (t0 + u * nt < tot) P[t0 + u * nt] = v[u];
  }
  __syncthreads();
  for (int jj = 0; jj < jb; ++jj) {
    // pivot: max |P[i][jj]|, i >= jj, smallest index on ties
    double v = -1.0;
    int idx = rows;
    for (int i = jj + tid; i < rows; i += nt) {
      double a = fabs(P[jj * rows + i]);
      if (a > v) { v = a; idx = i; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double ov = __shfl_xor(v, o, 64);
      int oi = __shfl_xor(idx, o, 64);
      if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    if ((tid & 63) == 0) { shv[tid >> 6] = v; shi[tid >> 6] = idx; }
    __syncthreads();
    if (tid < 16) {  // the waves' candidates: (largest value, then smallest index) in any order
      double bv = (tid < (nt >> 6)) ? shv[tid] : -1.0;
      int bi = (tid < (nt >> 6)) ? shi[tid] : rows;
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 16);
        const int oi = __shfl_xor(bi, o, 16);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      if (tid == 0) {
        if (bi >= rows) bi = jj;
        s_p = bi;
        piv[j0 + jj] = j0 + bi;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != jj) {
      for (int c = tid; c < jb; c += nt) {
        double t = P[c * rows + jj];
        P[c * rows + jj] = P[c * rows + p];
        P[c * rows + p] = t;
      }
    }
    __syncthreads();
    const double d = P[jj * rows + jj];
    for (int i = jj + 1 + tid; i < rows; i += nt) P[jj * rows + i] /= d;
    __syncthreads();
    // rank-1 update of the panel's trailing part: lanes along the rows (conflict-free), waves
    // over the columns. A lane takes 8 rows at a time: their multipliers once for all of the
    // wave's columns, and each column's 8 entries with their loads in flight together.
    for (int i0 = jj + 1 + (tid & 63); i0 < rows; i0 += 64 * 8) {
      double l[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) l[k] = (i0 + 64 * k < rows) ? P[jj * rows + i0 + 64 * k] : 0.0;
      for (int c = jj + 1 + (tid >> 6); c < jb; c += nt >> 6) {
        const double u = P[c * rows + jj];
        double x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = (i0 + 64 * k < rows) ? P[c * rows + i0 + 64 * k] : 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (i0 + 64 * k < rows) {
            x[k] -= l[k] * u;
            P[c * rows + i0 + 64 * k] = x[k];
          }
      }
    }
    __syncthreads();
  }
  for (int t = tid; t < rows * jb; t += nt) {
    int c = t / rows, i = t - c * rows;
    Aug[(int64_t)(j0 + i) * ld + j0 + c] = P[t];
  }
}

// Trailing columns c in [j0+jb, n]: apply the panel's row swaps, then U12 = L11^-1 A12.
struct KLuSwapTrsm {
  struct Args { double* Aug; int n, ld, j0, jb; const int* piv; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk);
};
__device__ void KLuSwapTrsm::run(const Args& args, const Blk& bk) {
  double* Aug = args.Aug; const int n = args.n, ld = args.ld, j0 = args.j0, jb = args.jb;
  const int* piv = args.piv;
  __shared__ double L[kLuNb][kLuNb];
  for (int t = threadIdx.x; t < jb * jb; t += blockDim.x) {
    int i = t / jb, c = t - i * jb;
    L[i][c] = Aug[(int64_t)(j0 + i) * ld + j0 + c];
  }
  __syncthreads();
  const int c = j0 + jb + bk.x * blockDim.x + threadIdx.x;
  if (c > n) return;
  double a[kLuNb];
#pragma unroll
  for (int i = 0; i < kLuNb; ++i) {
    if (i < jb) {
      const int r = j0 + i, p = piv[r];
      if (p != r) {
        double t = Aug[(int64_t)r * ld + c];
        Aug[(int64_t)r * ld + c] = Aug[(int64_t)p * ld + c];
        Aug[(int64_t)p * ld + c] = t;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kLuNb; ++i) a[i] = (i < jb) ? Aug[(int64_t)(j0 + i) * ld + c] : 0.0;
#pragma unroll
  for (int i = 0; i < kLuNb; ++i) {
    if (i < jb) {
      double v = a[i];
#pragma unroll
      for (int t = 0; t < kLuNb; ++t)
        if (t < i) v -= L[i][t] * a[t];
      a[i] = v;
    }
  }
#pragma unroll
  for (int i = 0; i < kLuNb; ++i)
    if (i < jb) Aug[(int64_t)(j0 + i) * ld + c] = a[i];
}

// Back substitution for rows [j0, j0+jb): x_blk = U_blk^-1 y_blk, then y[0:j0] -= U[0:j0, blk] x_blk.
struct KLuBacksolve {
  struct Args { double* Aug; int n, ld, j0, jb; double* x; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk);
};
__device__ void KLuBacksolve::run(const Args& a, const Blk& bk) {
  double* Aug = a.Aug; const int n = a.n, ld = a.ld, j0 = a.j0, jb = a.jb; double* x = a.x;
  __shared__ double U[kLuNb][kLuNb];
  __shared__ double xb[kLuNb];
  __shared__ double yb[kLuNb];
  for (int t = threadIdx.x; t < jb * jb; t += blockDim.x) {
    int i = t / jb, c = t - i * jb;
    U[i][c] = Aug[(int64_t)(j0 + i) * ld + j0 + c];
  }
  for (int i = threadIdx.x; i < jb; i += blockDim.x) yb[i] = Aug[(int64_t)(j0 + i) * ld + n];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = jb - 1; i >= 0; --i) {
      double v = yb[i];
      for (int t = i + 1; t < jb; ++t) v -= U[i][t] * xb[t];
      xb[i] = v / U[i][i];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < jb; i += blockDim.x) x[j0 + i] = xb[i];
  for (int r = bk.x * blockDim.x + threadIdx.x; r < j0; r += bk.gx * blockDim.x) {
    const double* row = Aug + (int64_t)r * ld;
    double u[kLuNb];  // the row's block entries loaded together, applied in order
#pragma unroll
    for (int t = 0; t < kLuNb; ++t) u[t] = (t < jb) ? row[j0 + t] : 0.0;
    double v = row[n];
#pragma unroll
    for (int t = 0; t < kLuNb; ++t)
      if (t < jb) v -= u[t] * xb[t];
    Aug[(int64_t)r * ld + n] = v;
  }
}

// Largest order whose panel (n x 1 column at least) fits in LDS.
int lu_max_order() { return 160 * 1024 / 8 - 512; }

void lu_solve_aug(double* Aug, int n, double* x, Arena& ws, hipStream_t st) {
  if (n <= 0) return;
  TT_CHECK(n <= lu_max_order(), ERR_UNSUPPORTED, "dense local solve: order too large (lower max_full)");
  // panel width: 32 columns, fewer when the n x width panel would not fit in LDS
  const int nb = std::max(1, std::min(kLuNb, (160 * 1024 / 8 - 512) / n));
  size_t mk = ws.mark();
  int* piv = ws.alloc<int>(n);
  const int ld = n + 1;
  for (int j0 = 0; j0 < n; j0 += nb) {
    const int jb = std::min(nb, n - j0);
    if (!ws.measuring()) {
      size_t lds = (size_t)(n - j0) * jb * sizeof(double);
      launch<KLuPanel>(dim3(1), dim3(1024), lds, st, {Aug, n, ld, j0, jb, piv});
      TT_LAUNCH_CHECK();
      const int ncols = n + 1 - (j0 + jb);
      if (ncols > 0) {
        launch<KLuSwapTrsm>(dim3((ncols + 255) / 256), dim3(256), 0, st, {Aug, n, ld, j0, jb, piv});
        TT_LAUNCH_CHECK();
      }
    }
    const int rem = n - (j0 + jb);
    if (rem > 0) {
      GemmDesc g;  // A22 (and the rhs column) -= L21 U12
      g.M = rem; g.N = rem + 1; g.K = jb;
      g.a.p = Aug + (int64_t)(j0 + jb) * ld + j0; g.a.r = ld; g.a.c = 1;
      g.b.p = Aug + (int64_t)j0 * ld + j0 + jb; g.b.r = ld; g.b.c = 1;
      g.c.p = Aug + (int64_t)(j0 + jb) * ld + j0 + jb; g.c.r = ld; g.c.c = 1;
      g.alpha = -1.0; g.beta = 1.0;
      gemm_f64(g, ws, st);
    }
  }
  if (!ws.measuring()) {
    const int last = ((n - 1) / nb) * nb;
    for (int j0 = last; j0 >= 0; j0 -= nb) {
      const int jb = std::min(nb, n - j0);
      const int blocks = std::max(1, std::min(64, (j0 + 255) / 256));
      launch<KLuBacksolve>(dim3(blocks), dim3(256), 0, st, {Aug, n, ld, j0, jb, x});
      TT_LAUNCH_CHECK();
    }
  }
  ws.release(mk);
}

void lu_init() {
  static int done_dev = -1;
  int dev = 0;
  TT_HIP(hipGetDevice(&dev));
  if (done_dev == dev) return;
  int dev_max = 0;
  TT_HIP(hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  for (const void* fn : {(const void*)k_one<KLuPanel>, (const void*)k_grp<KLuPanel>}) {
    hipFuncAttributes fa;
    TT_HIP(hipFuncGetAttributes(&fa, fn));
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            dev_max - (int)fa.sharedSizeBytes) != hipSuccess)
      (void)hipGetLastError();
  }
  done_dev = dev;
}

}  // namespace ttamen
