// QR and SVD of matrices with more columns than a workgroup's LDS holds as a square FP64 block
// (kMaxSmall = 140): the rank range above the LDS-resident Cholesky-QR and Jacobi kernels of
// linalg.hip, reached when a solve's rank capacity is raised (ttamen_rank_capacity) up to
// kMaxRank. Both keep their matrices in device memory and launch one kernel per step:
//
//   qr_large    Householder QR (kernels_house.h): n steps of KHouseStep build R and leave the
//               reflectors in the working copy, n steps of KHouseQ accumulate Q backwards. Q is
//               orthonormal whatever the rank of A (a zero column gives H_j = I and Q keeps e_j's
//               image), so no completion is needed.
//   svd_large   one-sided Jacobi on the columns of the k x k factor (Hestenes; the rotation of
//               Rutishauser, and the round-robin order and threshold k u of the LDS kernel): one
//               launch per round of k/2 disjoint column pairs, a workgroup per pair; the number of
//               rotations of a sweep is counted on the device and read once per sweep. Then the
//               column norms give the singular values, zero ones get an orthonormal completion
//               (classical Gram-Schmidt twice against the live columns), and the columns are
//               ranked on the device (descending, stable by column).
#include "ttamen_internal.h"
#include "kernels_blas1.h"
#include "kernels_house.h"
#include "gmres.h"
#include "cohort.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace ttamen {

constexpr double kLargeUnitRoundoff = 1.1102230246251565e-16;  // 2^-53
constexpr int kLargeJacobiMaxSweeps = 48;

// Round-robin tournament on K (even) players: round t in [0, K-2], pair i in [0, K/2).
__device__ __forceinline__ void rr_pair_large(int t, int i, int K, int& p, int& q) {
  if (i == 0) {
    p = t;
    q = K - 1;
  } else {
    p = (t + i) % (K - 1);
    q = (t - i + (K - 1)) % (K - 1);
  }
  if (p > q) { const int tmp = p; p = q; q = tmp; }
}

// One Jacobi round: workgroup i rotates the column pair i of round t of W and V (k x k, column-major)
// when |w_p . w_q| > tol ||w_p|| ||w_q||; *rot counts the rotations.
struct KJacobiRound {
  struct Args { double* W; double* V; int k, K, t; double tol; int* rot; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    __shared__ double sh[3][4];
    int p, q;
    rr_pair_large(a.t, (int)bk.x, a.K, p, q);
    if (q >= a.k) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    double* wp = a.W + (int64_t)p * a.k;
    double* wq = a.W + (int64_t)q * a.k;
    double al = 0.0, be = 0.0, ga = 0.0;
    for (int i = tid; i < a.k; i += 256) {
      const double x = wp[i], y = wq[i];
      al += x * x;
      be += y * y;
      ga += x * y;
    }
    al = wave_sum(al);
    be = wave_sum(be);
    ga = wave_sum(ga);
    if (lane == 0) {
      sh[0][w] = al;
      sh[1][w] = be;
      sh[2][w] = ga;
    }
    __syncthreads();
    al = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    be = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    ga = (sh[2][0] + sh[2][1]) + (sh[2][2] + sh[2][3]);
    if (!(al > 0.0 && be > 0.0 && fabs(ga) > a.tol * sqrt(al) * sqrt(be))) return;
    const double zeta = (be - al) / (2.0 * ga);
    double t;
    if (fabs(zeta) > 1e150) t = 0.5 / zeta;
    else t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
    double* vp = a.V + (int64_t)p * a.k;
    double* vq = a.V + (int64_t)q * a.k;
    for (int i = tid; i < a.k; i += 256) {
      const double x = wp[i], y = wq[i];
      wp[i] = c * x - s * y;
      wq[i] = s * x + c * y;
      const double u = vp[i], v = vq[i];
      vp[i] = c * u - s * v;
      vq[i] = s * u + c * v;
    }
    if (tid == 0) atomicAdd(a.rot, 1);
  }
};

// sig[j] = ||W[:, j]||; W[:, j] /= sig[j] (zero when sig[j] <= tiny). One workgroup per column.
struct KColNormalize {
  struct Args { double* W; int64_t rows; double* sig; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    __shared__ double sh[4];
    double* col = a.W + (int64_t)bk.x * a.rows;
    double p = 0.0;
    for (int64_t i = threadIdx.x; i < a.rows; i += 256) p += col[i] * col[i];
    const double s = sqrt(hh_block_sum(p, sh));
    const double inv = (s > 1e-300) ? 1.0 / s : 0.0;
    for (int64_t i = threadIdx.x; i < a.rows; i += 256) col[i] *= inv;
    if (threadIdx.x == 0) a.sig[bk.x] = s;
  }
};

// h[c] = W[:, c] . W[:, j] for the live columns c != j (sig[c] > tiny or c < j), else 0.
struct KLiveDots {
  struct Args { const double* W; int64_t rows; int j; const double* sig; double* h; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    __shared__ double sh[4];
    const int c = (int)bk.x;
    const bool live = c != a.j && (a.sig[c] > 1e-300 || c < a.j);
    if (!live) {
      if (threadIdx.x == 0) a.h[c] = 0.0;
      return;
    }
    const double* wc = a.W + (int64_t)c * a.rows;
    const double* wj = a.W + (int64_t)a.j * a.rows;
    double p = 0.0;
    for (int64_t i = threadIdx.x; i < a.rows; i += 256) p += wc[i] * wj[i];
    const double s = hh_block_sum(p, sh);
    if (threadIdx.x == 0) a.h[c] = s;
  }
};

// W[i, j] -= sum_c h[c] W[i, c] (c in column order).
struct KLiveUpdate {
  struct Args { double* W; int64_t rows; int ncols; int j; const double* h; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    for (int64_t i = bk.x * (int64_t)256 + threadIdx.x; i < a.rows; i += (int64_t)bk.gx * 256) {
      double v = a.W[(int64_t)a.j * a.rows + i];
      for (int c = 0; c < a.ncols; ++c) v -= a.h[c] * a.W[(int64_t)c * a.rows + i];
      a.W[(int64_t)a.j * a.rows + i] = v;
    }
  }
};

// W[:, j] = e_p
struct KUnitCol {
  struct Args { double* W; int64_t rows; int j; int p; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    for (int64_t i = bk.x * (int64_t)256 + threadIdx.x; i < a.rows; i += (int64_t)bk.gx * 256)
      a.W[(int64_t)a.j * a.rows + i] = (i == a.p) ? 1.0 : 0.0;
  }
};

// Descending rank of every column (ties by column index), the sorted values, and the gathered
// outputs: Ur (k x k row-major, column r = r-th left vector), V (k rows, leading dimension ldv).
struct KRankGather {
  struct Args {
    const double* sig; const double* W; const double* Vc; int k; double* s_out; double* Ur; double* V;
    int64_t ldv;
  };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    const int j = (int)bk.x;
    __shared__ int rank;
    if (threadIdx.x == 0) {
      const double sj = a.sig[j];
      int r = 0;
      for (int i = 0; i < a.k; ++i) {
        const double si = a.sig[i];
        r += (si > sj) || (si == sj && i < j);
      }
      rank = r;
      a.s_out[r] = sj;
    }
    __syncthreads();
    const int r = rank;
    for (int i = threadIdx.x; i < a.k; i += 256) {
      a.Ur[(int64_t)i * a.k + r] = a.W[(int64_t)j * a.k + i];
      a.V[(int64_t)i * a.ldv + r] = a.Vc[(int64_t)j * a.k + i];
    }
  }
};

static thread_local int tl_rank_capacity = kMaxSmall;

int rank_capacity() { return tl_rank_capacity; }

int set_rank_capacity(int cap) {
  const int prev = tl_rank_capacity;
  if (cap > 0) tl_rank_capacity = std::min(std::max(cap, kMaxSmall), kMaxRank);
  return prev;
}

int qr_large(const double* A, int64_t sr, int64_t sc, int64_t m, int64_t n, double* Q, double* R, Arena& ws,
             hipStream_t st) {
  TT_CHECK(m >= n && n >= 1, ERR_INVALID, "qr_large: needs m >= n >= 1");
  TT_CHECK(n <= kMaxRank, ERR_UNSUPPORTED,
           "qr: more than " + std::to_string(kMaxRank) + " columns is not supported");
  size_t mk = ws.mark();
  double* M = ws.alloc<double>(m * n);
  double* Y = ws.alloc<double>(m * n);
  if (!ws.measuring()) {
    copy2d(A, sc, sr, M, m, n, m, st);  // M[j m + i] = A(i, j), column-major
    for (int j = 0; j < n; ++j) {
      launch<KHouseStep>(dim3((unsigned)(n - j)), dim3(256), 0, st, {M, m, (int)n, j, R});
      TT_LAUNCH_CHECK();
    }
    eye(Y, n, m, st);  // column-major [I_n; 0]
    for (int j = (int)n - 1; j >= 0; --j) {
      launch<KHouseQ>(dim3((unsigned)(n - j)), dim3(256), 0, st, {M, m, j, Y});
      TT_LAUNCH_CHECK();
    }
    copy2d(Y, 1, m, Q, n, m, n, st);  // Q row-major, leading dimension n
  }
  ws.release(mk);
  return (int)n;
}

void svd_large(const double* Mx, int64_t sr, int64_t sc, int k, double* s, double* Ur, double* V, int64_t ldv,
               Arena& ws, hipStream_t st) {
  TT_CHECK(k >= 1 && k <= kMaxRank, ERR_UNSUPPORTED,
           "svd: rank above " + std::to_string(kMaxRank) + " is not supported");
  const int K = (k % 2) ? k + 1 : k;
  size_t mk = ws.mark();
  double* W = ws.alloc<double>((size_t)k * k);
  double* Vc = ws.alloc<double>((size_t)k * k);
  double* sig = ws.alloc<double>(k);
  double* h = ws.alloc<double>(k);
  double* slot = ws.alloc<double>(2);
  int* rot = ws.alloc<int>(2);
  if (ws.measuring()) {
    ws.release(mk);
    return;
  }
  copy2d(Mx, sc, sr, W, k, k, k, st);  // W[j k + i] = M(i, j)
  eye(Vc, k, k, st);
  // the pinned mailbox of this solve (a cohort member's own, or the thread's): 4096 doubles
  double* host = member_host_sync().host_dbl();
  const double tol = (double)k * kLargeUnitRoundoff;
  for (int sw = 0; sw < kLargeJacobiMaxSweeps && k > 1; ++sw) {
    zero_words(rot, 1, st);
    for (int t = 0; t < K - 1; ++t) {
      launch<KJacobiRound>(dim3((unsigned)(K / 2)), dim3(256), 0, st, {W, Vc, k, K, t, tol, rot});
      TT_LAUNCH_CHECK();
    }
    read_to_host(host, rot, sizeof(int), st);
    if (*reinterpret_cast<int*>(host) == 0) break;
  }
  launch<KColNormalize>(dim3((unsigned)k), dim3(256), 0, st, {W, k, sig});
  TT_LAUNCH_CHECK();
  // Orthonormal completion of the columns whose singular value is zero (rare): the unit vector of a
  // probe row, orthogonalised twice against the live columns, kept once its norm exceeds 1/2.
  read_to_host(host, sig, sizeof(double) * k, st);
  std::vector<int> dead;
  for (int j = 0; j < k; ++j)
    if (!(host[j] > 1e-300)) dead.push_back(j);
  int probe = 0;
  for (int j : dead) {
    for (int tries = 0; tries < 4 * k + 8; ++tries) {
      const int p = (probe++ * 37 + j * 101) % k;
      launch<KUnitCol>(dim3(grid_1d(k)), dim3(256), 0, st, {W, k, j, p});
      TT_LAUNCH_CHECK();
      for (int pass = 0; pass < 2; ++pass) {
        launch<KLiveDots>(dim3((unsigned)k), dim3(256), 0, st, {W, k, j, sig, h});
        launch<KLiveUpdate>(dim3(grid_1d(k)), dim3(256), 0, st, {W, k, k, j, h});
        TT_LAUNCH_CHECK();
      }
      launch<KColNormalize>(dim3(1), dim3(256), 0, st, {W + (int64_t)j * k, k, slot});
      TT_LAUNCH_CHECK();
      read_to_host(host, slot, sizeof(double), st);
      if (host[0] > 0.5) break;
    }
  }
  launch<KRankGather>(dim3((unsigned)k), dim3(256), 0, st, {sig, W, Vc, k, s, Ur, V, ldv});
  TT_LAUNCH_CHECK();
  ws.release(mk);
}

}  // namespace ttamen
