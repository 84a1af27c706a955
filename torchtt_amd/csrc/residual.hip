// Global residual of a TT solve, ||A x - b|| and ||b||, in FP64 on the GPU (reference: the check
// every caller of amen_solve makes with TT algebra, torchtt/_tt_base.py:780-846 (A @ x), :545 (-),
// :990-1052 (norm)).
//
// The TT of r = A x - b has ranks rA rx + rb: its cores are the block-diagonal concatenation of the
// cores of A x and of b (first core [Ax_0 | -b_0], last core the sum of the two blocks). Its norm
// is taken by a left-to-right orthogonalisation sweep without forming those cores:
//   cur_0 = R_{-1} C_0 with R_{-1} = [1 | -1];  cur_k = R_{k-1} C_k, rows (q, m), columns the right
//   ranks of both blocks;  R_k = the R factor of cur_k;  ||r|| = ||cur_{d-1}[:,Ax] + cur_{d-1}[:,b]||.
// R_{k-1} C_k is two strided GEMMs for the A x block (x first, then the operator core) and one for
// the b block, written straight into cur_k stored column-major.
//
// The R factor is Householder's (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 5.2.1), one
// launch per column j: every workgroup forms the reflector of column j from the same reduction
// (identical in every workgroup), and workgroup c applies it to column j + c. Householder is
// backward stable whatever the conditioning of cur_k, which is what the norm of a residual needs:
// the two blocks cancel to ~1e-8 of their size, and a Gram-matrix (Cholesky) route would lose that
// to the squaring. Only R is kept (Q is never needed).
#include "ttamen_internal.h"
#include "kernels_blas1.h"
#include "kernels_house.h"
#include "../../include/ttamen.h"

#include <cmath>
#include <vector>

namespace ttamen {

// v[i] = M[i] + M[rows + i]: the sum of the last core's two blocks (columns 0 and 1).
struct KAddCols {
  struct Args { const double* M; int64_t rows; double* v; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    for (int64_t i = bk.x * (int64_t)256 + threadIdx.x; i < a.rows; i += (int64_t)bk.gx * 256)
      a.v[i] = a.M[i] + a.M[a.rows + i];
  }
};

// R factor (k x n row-major, k = min(rows, n)) of M (rows x n column-major, overwritten).
static int64_t house_r(double* M, int64_t rows, int n, double* R, hipStream_t st) {
  const int64_t k = std::min<int64_t>(rows, n);
  for (int j = 0; j < k; ++j) {
    launch<KHouseStep>(dim3(n - j), dim3(256), 0, st, {M, rows, n, j, R});
    TT_LAUNCH_CHECK();
  }
  return k;
}

struct ResidualPlan {
  int d;
  std::vector<int64_t> N, rA, rx, rb;
  bool with_ax;
  size_t cur_elems = 0, u_elems = 0, apm_elems = 0, r_elems = 0;
};

static ResidualPlan residual_plan(int d, const int64_t* N, const int64_t* rA, const int64_t* rx, const int64_t* rb,
                                  bool with_ax) {
  ResidualPlan p;
  p.d = d;
  p.with_ax = with_ax;
  p.N.assign(N, N + d);
  p.rb.assign(rb, rb + d + 1);
  if (with_ax) {
    p.rA.assign(rA, rA + d + 1);
    p.rx.assign(rx, rx + d + 1);
  } else {
    p.rA.assign(d + 1, 0);
    p.rx.assign(d + 1, 0);
  }
  // left width of core k (columns of R_{k-1}): the two blocks' ranks, 2 at the first core
  int64_t q = 1;
  for (int k = 0; k < d; ++k) {
    const int64_t L = p.rA[k] * p.rx[k] + p.rb[k];
    const int64_t Lr = p.rA[k + 1] * p.rx[k + 1] + p.rb[k + 1];
    const int64_t rows = q * p.N[k];
    p.cur_elems = std::max<size_t>(p.cur_elems, (size_t)(rows * Lr));
    p.r_elems = std::max<size_t>(p.r_elems, (size_t)(std::min(rows, Lr) * Lr + L));
    if (with_ax) {
      p.u_elems = std::max<size_t>(p.u_elems, (size_t)(q * p.rA[k] * p.N[k] * p.rx[k + 1]));
      p.apm_elems = std::max<size_t>(p.apm_elems, (size_t)(p.rA[k] * p.N[k] * p.N[k] * p.rA[k + 1]));
    }
    q = std::min(rows, Lr);
  }
  return p;
}

static size_t residual_ws_bytes(const ResidualPlan& p) {
  Arena a;
  a.alloc<double>(p.cur_elems);
  a.alloc<double>(p.r_elems);
  a.alloc<double>(p.r_elems);
  a.alloc<double>(std::max<size_t>(p.u_elems, 1));
  a.alloc<double>(std::max<size_t>(p.apm_elems, 1));
  a.alloc<double>(kRedBlocks * 64 + 64);
  a.alloc<unsigned>(64);
  return a.peak() + (64ull << 20);  // split-K partials of the GEMMs
}

// ||A x - b|| (with_ax) or ||b|| into *out (device slot).
static void residual_norm(const ResidualPlan& p, const double* const* A, const double* const* x,
                          const double* const* b, double* out, Arena& ws, hipStream_t st) {
  const int d = p.d;
  double* cur = ws.alloc<double>(p.cur_elems);
  double* Rp = ws.alloc<double>(p.r_elems);  // R_{k-1}, row-major q x L
  double* Rn = ws.alloc<double>(p.r_elems);
  double* u = ws.alloc<double>(std::max<size_t>(p.u_elems, 1));
  double* Apm = ws.alloc<double>(std::max<size_t>(p.apm_elems, 1));
  RedScratch rs;
  rs.partial = ws.alloc<double>(kRedBlocks * 64 + 64);
  rs.counter = ws.alloc<unsigned>(64);
  zero_words(rs.counter, 64, st);
  // R_{-1} = [1 | -1] (with the A x block) or [1] (b alone)
  int64_t q = 1;
  if (p.with_ax) {
    fill(Rp, 1, 1.0, st);
    fill(Rp + 1, 1, -1.0, st);
  } else {
    fill(Rp, 1, 1.0, st);
  }
  for (int k = 0; k < d; ++k) {
    const int64_t n = p.N[k], ra = p.rA[k], ra1 = p.rA[k + 1], rx = p.rx[k], rx1 = p.rx[k + 1];
    const int64_t rb = p.rb[k], rb1 = p.rb[k + 1];
    const int64_t L = ra * rx + rb, Lax1 = ra1 * rx1, Lr = Lax1 + rb1;
    const int64_t rows = q * n;
    if (p.with_ax) {
      {  // u[q][a][n][i'] = sum_i R[q][a rx + i] x[i][n][i']
        GemmDesc g;
        g.M = (int)q; g.N = (int)(n * rx1); g.K = (int)rx; g.B2 = (int)ra;
        g.a.p = Rp; g.a.r = L; g.a.c = 1; g.a.b2 = rx;
        g.b.p = x[k]; g.b.r = n * rx1; g.b.c = 1;
        g.c.p = u; g.c.r = ra * n * rx1; g.c.c = 1; g.c.b2 = n * rx1;
        gemm_f64(g, ws, st);
      }
      permute_core_A(A[k], Apm, (int)ra, (int)n, (int)ra1, st);
      {  // cur[(q, m)][(a', i')] = sum_(a, n) Apm[(m, a')][(a, n)] u[q][(a, n)][i']
        GemmDesc g;
        g.M = (int)n; g.N = (int)rx1; g.K = (int)(ra * n); g.B1 = (int)q; g.B2 = (int)ra1;
        g.a.p = Apm; g.a.r = ra1 * ra * n; g.a.c = 1; g.a.b2 = ra * n;
        g.b.p = u; g.b.r = rx1; g.b.c = 1; g.b.b1 = ra * n * rx1;
        g.c.p = cur; g.c.r = 1; g.c.c = rows; g.c.b1 = n; g.c.b2 = rx1 * rows;
        gemm_f64(g, ws, st);
      }
    }
    {  // cur[(q, m)][Lax1 + b'] = sum_b R[q][ra rx + b] b[b][m][b']
      GemmDesc g;
      g.M = (int)q; g.N = (int)n; g.K = (int)rb; g.B2 = (int)rb1;
      g.a.p = Rp + ra * rx; g.a.r = L; g.a.c = 1;
      g.b.p = b[k]; g.b.r = n * rb1; g.b.c = rb1; g.b.b2 = 1;
      g.c.p = cur + Lax1 * rows; g.c.r = n; g.c.c = 1; g.c.b2 = rows;
      gemm_f64(g, ws, st);
    }
    if (k == d - 1) break;
    q = house_r(cur, rows, (int)Lr, Rn, st);
    std::swap(Rp, Rn);
  }
  const int64_t rows = q * p.N[d - 1];
  if (p.with_ax) {
    // u is free now and holds rows elements: at the last core it was q rA n rx' >= q n of them
    double* v = u;
    TT_CHECK((size_t)rows <= p.u_elems, ERR_INVALID, "residual: workspace layout");
    launch<KAddCols>(dim3(grid_1d(rows)), dim3(256), 0, st, {cur, rows, v});
    TT_LAUNCH_CHECK();
    norm2(v, rows, out, rs, st);
  } else {
    norm2(cur, rows, out, rs, st);
  }
}

}  // namespace ttamen

using namespace ttamen;

extern "C" {

size_t ttamen_tt_residual_norm_workspace_size(int d, const int64_t* N, const int64_t* rA, const int64_t* rx,
                                              const int64_t* rb) {
  size_t out = 0;
  const int rc = abi_call([&] {
    TT_CHECK(d >= 1 && N && rA && rx && rb, ERR_INVALID, "null argument");
    out = std::max(residual_ws_bytes(residual_plan(d, N, rA, rx, rb, true)),
                   residual_ws_bytes(residual_plan(d, N, rA, rx, rb, false)));
  });
  return rc == 0 ? out : 0;
}

int ttamen_tt_residual_norm(int d, const int64_t* N, const int64_t* rA, const int64_t* rx, const int64_t* rb,
                            const double* const* A_cores, const double* const* x_cores,
                            const double* const* b_cores, void* workspace, size_t workspace_bytes,
                            double* out2, void* stream) {
  return abi_call([&] {
    TT_CHECK(d >= 1 && N && rA && rx && rb && A_cores && x_cores && b_cores && workspace && out2, ERR_INVALID,
             "null argument");
    for (int k = 0; k <= d; ++k)
      TT_CHECK(rA[k] >= 1 && rx[k] >= 1 && rb[k] >= 1, ERR_INVALID, "ranks must be positive");
    TT_CHECK(rA[0] == 1 && rA[d] == 1 && rx[0] == 1 && rx[d] == 1 && rb[0] == 1 && rb[d] == 1, ERR_INVALID,
             "boundary ranks must be 1");
    hipStream_t st = (hipStream_t)stream;
    const ResidualPlan pr = residual_plan(d, N, rA, rx, rb, true);
    const ResidualPlan pb = residual_plan(d, N, rA, rx, rb, false);
    Arena ws(workspace, workspace_bytes);
    double* slots = ws.alloc<double>(2);
    {
      const size_t mk = ws.mark();
      residual_norm(pr, A_cores, x_cores, b_cores, slots, ws, st);
      ws.release(mk);
    }
    residual_norm(pb, A_cores, x_cores, b_cores, slots + 1, ws, st);
    TT_HIP(hipMemcpyAsync(out2, slots, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    TT_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
