// Local preconditioners of the AMEn iterative local solve (reference: AMENsolveMV::setter /
// apply_prec, cpp/matvecs.h:153-194; Python twin torchtt/solvers.py:93-136). Both are right
// preconditioners: GMRES runs on B P and the local solution is P y.
//   'c' (central, prec 1): for every pair (l, L) of interface indices the n x n block
//       M_{lL}[m][n] = sum_{s,S} PhiL[l,s,l] A[s,m,n,S] PhiR[L,S,L]  is inverted;
//       (P x)[l,:,L] = M_{lL}^{-1} x[l,:,L].
//   'r' (right/block, prec 2): for every l the (n r2) x (n r2) block
//       M_l[(m,L)][(n,R)] = sum_{s,S} PhiL[l,s,l] A[s,m,n,S] PhiR[L,S,R]  is inverted;
//       (P x)[l,:,:] = M_l^{-1} x[l,:,:].
// Every block is built and inverted by one workgroup in LDS (in-place Gauss-Jordan with
// partial pivoting); blocks of order above kPrecMaxOrder (140) are not supported. The apply is
// a strided batched GEMM on the matrix cores.
#include "ttamen_internal.h"
#include <algorithm>

namespace ttamen {

constexpr int kPrecMaxOrder = 140;  // N x N FP64 block in LDS

struct KPrecBuild {
  struct Args {
    int mode;  // 1: 'c', 2: 'r'
    const double* PhiL; const double* A; const double* PhiR; double* J;
    int r1, s, n, S, r2;
  };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk);
};

__device__ void KPrecBuild::run(const Args& a, const Blk& bk) {
  extern __shared__ double W[];  // N x N row-major, inverted in place
  __shared__ double s_piv;
  __shared__ int s_p;
  __shared__ double shv[4];
  __shared__ int shi[4];
  __shared__ int perm[kPrecMaxOrder];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = a.n, s = a.s, S = a.S, r1 = a.r1, r2 = a.r2;
  int N, l, L = 0;
  if (a.mode == 1) {
    N = n;
    l = (int)(bk.x / r2);
    L = (int)(bk.x % r2);
  } else {
    N = n * r2;
    l = (int)bk.x;
  }
  // build
  for (int t = tid; t < N * N; t += nt) {
    const int row = t / N, col = t - row * N;
    int m, mm, Lr, Rc;
    if (a.mode == 1) {
      m = row; mm = col; Lr = L; Rc = L;
    } else {
      m = row / r2; Lr = row - m * r2;
      mm = col / r2; Rc = col - mm * r2;
    }
    double acc = 0.0;
    for (int si = 0; si < s; ++si) {
      const double pl = a.PhiL[((int64_t)l * s + si) * r1 + l];
      if (pl == 0.0) continue;
      for (int Si = 0; Si < S; ++Si)
        acc += pl * a.A[(((int64_t)si * n + m) * n + mm) * S + Si] *
               a.PhiR[((int64_t)Lr * S + Si) * r2 + Rc];
    }
    W[t] = acc;
  }
  __syncthreads();
  // in-place Gauss-Jordan with partial (row) pivoting; the row swaps are undone on the
  // columns of the inverse at the end
  for (int j = 0; j < N; ++j) {
    double v = -1.0;
    int idx = N;
    for (int i = j + tid; i < N; i += nt) {
      const double x = fabs(W[i * N + j]);
      if (x > v) { v = x; idx = i; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(v, o, 64);
      const int oi = __shfl_xor(idx, o, 64);
      if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    if ((tid & 63) == 0) { shv[tid >> 6] = v; shi[tid >> 6] = idx; }
    __syncthreads();
    if (tid == 0) {
      double bv = shv[0];
      int bi = shi[0];
      for (int w = 1; w < (nt >> 6); ++w)
        if (shv[w] > bv || (shv[w] == bv && shi[w] < bi)) { bv = shv[w]; bi = shi[w]; }
      if (bi >= N) bi = j;
      s_p = bi;
      perm[j] = bi;
    }
    __syncthreads();
    const int p = s_p;
    if (p != j)
      for (int c = tid; c < N; c += nt) {
        const double t = W[j * N + c];
        W[j * N + c] = W[p * N + c];
        W[p * N + c] = t;
      }
    __syncthreads();
    if (tid == 0) {
      s_piv = W[j * N + j];
      W[j * N + j] = 1.0;
    }
    __syncthreads();
    const double inv = 1.0 / s_piv;  // a singular block gives inf/NaN, as an inverse would
    for (int c = tid; c < N; c += nt) W[j * N + c] *= inv;
    __syncthreads();
    // eliminate column j from the other rows: f_i = W[i][j] (read before any write)
    for (int t = tid; t < N * N; t += nt) {
      const int i = t / N, c = t - i * N;
      if (i == j || c == j) continue;
      W[i * N + c] -= W[i * N + j] * W[j * N + c];
    }
    __syncthreads();
    for (int i = tid; i < N; i += nt)
      if (i != j) W[i * N + j] = -W[i * N + j] * W[j * N + j];
    __syncthreads();
  }
  for (int j = N - 1; j >= 0; --j) {
    const int p = perm[j];
    if (p != j)
      for (int i = tid; i < N; i += nt) {
        const double t = W[i * N + j];
        W[i * N + j] = W[i * N + p];
        W[i * N + p] = t;
      }
    __syncthreads();
  }
  double* J = a.J + (size_t)bk.x * N * N;
  for (int t = tid; t < N * N; t += nt) J[t] = W[t];
}

int prec_max_order() { return kPrecMaxOrder; }

size_t prec_elems(int mode, int r1, int n, int r2) {
  if (mode == 1) return (size_t)r1 * r2 * n * n;
  if (mode == 2) return (size_t)r1 * (size_t)(n * r2) * (n * r2);
  return 0;
}

void prec_build(int mode, const double* PhiL, const double* A, const double* PhiR, double* J,
                int r1, int s, int n, int S, int r2, hipStream_t st) {
  const int N = (mode == 1) ? n : n * r2;
  TT_CHECK(mode == 1 || mode == 2, ERR_UNSUPPORTED, "preconditioner must be 'c' or 'r'");
  TT_CHECK(N <= prec_max_order(), ERR_UNSUPPORTED,
           std::string("preconditioner block of order ") + std::to_string(N) + " exceeds " +
               std::to_string(prec_max_order()));
  const int blocks = (mode == 1) ? r1 * r2 : r1;
  launch<KPrecBuild>(dim3(blocks), dim3(256), (size_t)N * N * sizeof(double), st,
                     {mode, PhiL, A, PhiR, J, r1, s, n, S, r2});
  TT_LAUNCH_CHECK();
}

// y = P x for x, y of shape r1 x n x r2 (y must not alias x).
void prec_apply(int mode, const double* J, const double* x, double* y, int r1, int n, int r2,
                Arena& ws, hipStream_t st, const int* skip) {
  GemmDesc g;
  g.skip = skip;
  if (mode == 1) {
    // batch (l, L): y[l,:,L] = J[l][L] (n x n) x[l,:,L]
    g.M = n; g.N = 1; g.K = n; g.B1 = r1; g.B2 = r2;
    g.a.p = J; g.a.r = n; g.a.c = 1; g.a.b1 = (int64_t)r2 * n * n; g.a.b2 = (int64_t)n * n;
    g.b.p = x; g.b.r = r2; g.b.c = 0; g.b.b1 = (int64_t)n * r2; g.b.b2 = 1;
    g.c.p = y; g.c.r = r2; g.c.c = 0; g.c.b1 = (int64_t)n * r2; g.c.b2 = 1;
  } else {
    // batch l: y[l,:,:] = J[l] ((n r2) x (n r2)) x[l,:,:]
    const int64_t N = (int64_t)n * r2;
    g.M = (int)N; g.N = 1; g.K = (int)N; g.B1 = r1;
    g.a.p = J; g.a.r = N; g.a.c = 1; g.a.b1 = N * N;
    g.b.p = x; g.b.r = 1; g.b.c = 0; g.b.b1 = N;
    g.c.p = y; g.c.r = 1; g.c.c = 0; g.c.b1 = N;
  }
  gemm_f64(g, ws, st);
}

void prec_init() {
  static int done_dev = -1;
  int dev = 0;
  TT_HIP(hipGetDevice(&dev));
  if (done_dev == dev) return;
  const int need = kPrecMaxOrder * kPrecMaxOrder * (int)sizeof(double);
  if (allow_lds<KPrecBuild>(need) != hipSuccess) (void)hipGetLastError();
  done_dev = dev;
}

}  // namespace ttamen
