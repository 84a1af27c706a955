// AMEn approximation of TT products: y ~= A b for a TT-matrix A and a TT vector b (reference
// torchtt/_amen.py: amen_mv :55-101, _amen_mm_python :126-413 on vector cores), and C ~= A B for
// two TT-matrices through the same sweep on the operator A (x) I_N (amen_mm :104-123; the
// reference keeps the cores of C as [r, m, n, r'], which is the vector form with mode (m, n)).
//
// The sweep has no linear solve: at each core the new solution is the projection of A b onto
// the current interfaces, truncated by the singular-value rule rank_chop at
// ||w|| eps / d^1.5 (d^0.5 in the last sweep), then enriched with kickrank directions of the
// residual proxy z, as in the reference. Every contraction is a library routine:
//   _local_AB(Phi_L, Phi_R, A, b)   -> local_matvec (rectangular operator core)
//   _compute_phi_{fwd,bck}_AB       -> phi_{fwd,bck}_A_rect with x on the left, b on the right
//   _compute_phi_{fwd,bck}_x        -> phi_{fwd,bck}_rhs
// The scalar bookkeeping (nrmsc, normb, normx) runs on the host from device norms; the
// reference's normA stays 1 (its updates are commented out, _amen.py:201-205, :355-359).
#include "ttamen_internal.h"
#include "kernels_blas1.h"
#include "../../include/ttamen.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

namespace ttamen {

void linalg_init();

namespace {

// D[a][(m,n')][(k,n)][A] = A[a][m][k][A] if n == n' (the operator A (x) I_N of amen_mm).
struct KKronEye {
  struct Args { const double* A; double* D; int ra, M, K, N, rA; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    const int64_t MN = (int64_t)p.M * p.N, KN = (int64_t)p.K * p.N;
    const int64_t total = (int64_t)p.ra * MN * KN * p.rA;
    for (int64_t t = bk.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)bk.gx * blockDim.x) {
      int64_t q = t;
      const int64_t Ai = q % p.rA; q /= p.rA;
      const int64_t col = q % KN; q /= KN;
      const int64_t row = q % MN; const int64_t a = q / MN;
      const int64_t m = row / p.N, n1 = row - m * p.N, k = col / p.N, n2 = col - k * p.N;
      p.D[t] = (n1 == n2) ? p.A[((a * p.M + m) * p.K + k) * p.rA + Ai] : 0.0;
    }
  }
};

struct ProductPlan {
  int d = 0;
  std::vector<int64_t> M, K, rA, rb, caps, Rb;
  int64_t rz = 0, maxR = 1, maxM = 1, maxK = 1, maxrA = 1, maxrb = 1, maxcore = 1;
};

ProductPlan make_product_plan(int d, const int64_t* M, const int64_t* K, const int64_t* rA,
                              const int64_t* rb, const int64_t* rx0, const int64_t* caps,
                              int64_t kickrank, int64_t kick2) {
  TT_CHECK(d >= 1 && M && K && rA && rb && caps, ERR_INVALID, "null argument");
  ProductPlan p;
  p.d = d;
  p.M.assign(M, M + d);
  p.K.assign(K, K + d);
  p.rA.assign(rA, rA + d + 1);
  p.rb.assign(rb, rb + d + 1);
  p.caps.assign(caps, caps + d + 1);
  for (int k = 0; k < d; ++k) TT_CHECK(M[k] >= 1 && K[k] >= 1, ERR_INVALID, "mode sizes must be positive");
  for (int k = 0; k <= d; ++k) TT_CHECK(rA[k] >= 1 && rb[k] >= 1, ERR_INVALID, "ranks must be positive");
  TT_CHECK(rA[0] == 1 && rA[d] == 1 && rb[0] == 1 && rb[d] == 1, ERR_INVALID, "boundary ranks must be 1");
  for (int k = 1; k < d; ++k) TT_CHECK(caps[k] >= 1, ERR_INVALID, "rank caps must be positive");
  TT_CHECK(kickrank >= 0 && kick2 >= 0, ERR_INVALID, "invalid enrichment sizes");
  p.rz = kickrank + kick2;
  std::vector<int64_t> right(d + 1, 1);
  for (int k = d - 1; k >= 0; --k) right[k] = std::min<int64_t>(right[k + 1] * M[k], int64_t(1) << 40);
  p.Rb.assign(d + 1, 1);
  for (int k = 0; k < d; ++k) {
    int64_t b = std::min(std::min<int64_t>(p.Rb[k] * M[k], int64_t(1) << 40),
                         std::min(caps[k + 1] + p.rz, right[k + 1] + p.rz));
    b = std::min<int64_t>(b, rank_capacity());  // a sweep that needs more stops with a status (solve again at
                                                //   a larger capacity: the result does not depend on it)
    if (k + 1 == d) b = 1;
    if (rx0) b = std::max(b, rx0[k + 1]);
    p.Rb[k + 1] = std::max<int64_t>(b, 1);
  }
  if (rx0) p.Rb[0] = std::max<int64_t>(1, rx0[0]);
  for (int k = 0; k < d; ++k) {
    p.maxcore = std::max(p.maxcore, p.Rb[k] * M[k] * p.Rb[k + 1]);
    p.maxM = std::max(p.maxM, M[k]);
    p.maxK = std::max(p.maxK, K[k]);
  }
  for (int k = 0; k <= d; ++k) {
    p.maxR = std::max(p.maxR, p.Rb[k]);
    p.maxrA = std::max(p.maxrA, rA[k]);
    p.maxrb = std::max(p.maxrb, rb[k]);
  }
  TT_CHECK(p.maxR <= kMaxRank, ERR_UNSUPPORTED,
           "initial-guess rank " + std::to_string(p.maxR) + " exceeds " + std::to_string(kMaxRank));
  return p;
}

struct ProductOptions {
  int64_t nswp = 22, kickrank = 4, kick2 = 0;
  double eps = 1e-10;
  uint64_t seed = 0;
  int verbose = 0;
};

class ProductSweep {
 public:
  ProductSweep(const ProductPlan& p, const ProductOptions& o, Arena& ws, hipStream_t st)
      : P(p), O(o), ws_(ws), st_(st), d(p.d) {}

  void allocate(const double* const* A, const double* const* b, const double* const* x0,
                const int64_t* rx0, const double* const* z0) {
    A_.assign(A, A + d);
    b_.assign(b, b + d);
    x0_ = x0;
    rx0_ = rx0;
    z0_ = z0;
    const int64_t rz = std::max<int64_t>(P.rz, 1);
    Apm_.resize(d);
    x_.resize(d);
    z_.resize(d);
    for (int k = 0; k < d; ++k) {
      Apm_[k] = ws_.alloc<double>(P.rA[k] * P.M[k] * P.K[k] * P.rA[k + 1]);
      x_[k] = ws_.alloc<double>(P.maxcore);
      z_[k] = ws_.alloc<double>(rz * P.M[k] * rz);
    }
    spare_ = ws_.alloc<double>(P.maxcore);
    Phis_.resize(d + 1);
    Phisr_.resize(d + 1);
    Phiz_.resize(d + 1);
    Phizr_.resize(d + 1);
    for (int k = 0; k <= d; ++k) {
      Phis_[k] = ws_.alloc<double>(P.Rb[k] * P.Rb[k]);
      Phisr_[k] = ws_.alloc<double>(P.Rb[k] * P.rA[k] * P.rb[k]);
      Phiz_[k] = ws_.alloc<double>(rz * P.Rb[k]);
      Phizr_[k] = ws_.alloc<double>(rz * P.rA[k] * P.rb[k]);
    }
    slot_ = ws_.alloc<double>(8);
    rs_.partial = ws_.alloc<double>(kRedBlocks * 64 + 64);
    rs_.counter = ws_.alloc<unsigned>(64);
  }

  void run() {
    if (ws_.measuring()) return;
    zero_words(rs_.counter, 64, st_);
    for (int k = 0; k < d; ++k)
      permute_core_A_rect(A_[k], Apm_[k], (int)P.rA[k], (int)P.M[k], (int)P.K[k], (int)P.rA[k + 1], st_);
    rx_.assign(d + 1, 1);
    if (x0_) {
      for (int k = 0; k <= d; ++k) rx_[k] = rx0_[k];
      for (int k = 0; k < d; ++k) copy(x0_[k], x_[k], rx_[k] * P.M[k] * rx_[k + 1], st_);
    } else {
      for (int k = 0; k < d; ++k) fill(x_[k], P.M[k], 0.0, st_);  // zero cores of rank 1 (:134-137)
    }
    rz_.assign(d + 1, 1);
    for (int k = 1; k < d; ++k) rz_[k] = P.rz;
    if (P.rz > 0) {
      TT_CHECK(z0_, ERR_INVALID, "residual-proxy cores are required when kickrank + kick2 > 0");
      for (int k = 0; k < d; ++k) copy(z0_[k], z_[k], rz_[k] * P.M[k] * rz_[k + 1], st_);
      rl_orthogonal_z();
    }
    for (int k : {0, d}) {
      fill(Phis_[k], 1, 1.0, st_);
      fill(Phisr_[k], 1, 1.0, st_);
      fill(Phiz_[k], 1, 1.0, st_);
      fill(Phizr_[k], 1, 1.0, st_);
    }
    normb_.assign(std::max(d - 1, 0), 1.0);
    normx_.assign(std::max(d - 1, 0), 1.0);
    nrmsc_ = 1.0;
    bool last = false;
    int swp = 0;
    for (swp = 0; swp < O.nswp; ++swp) {
      for (int k = d - 1; k > 0; --k) rl_step(k, swp, last);
      double max_dx = 0.0;
      for (int k = 0; k < d; ++k) lr_step(k, last, max_dx);
      if (O.verbose) {
        printf("Sweep %d%s: rank [", swp + 1, last ? " (last)" : "");
        for (auto r : rx_) printf(" %lld", (long long)r);
        printf(" ] max dx %g\n", max_dx);
        fflush(stdout);
      }
      stats_.sweeps = swp + 1;
      if (last) break;
      if (max_dx < O.eps) last = true;  // a NaN max dx (0/0 for a zero product) never stops
    }
    // rescale by the geometric mean of the interface norms (:400-403)
    if (d > 1) {
      double lg = 0.0;
      for (double v : normx_) lg += std::log(v);
      const double nx = std::exp(lg / d);
      for (int k = 0; k < d; ++k) axpby(nx, x_[k], 0.0, x_[k], rx_[k] * P.M[k] * rx_[k + 1], st_);
    }
    for (auto r : rx_) stats_.max_rank = std::max<int64_t>(stats_.max_rank, r);
  }

  void write_out(double* out, size_t out_elems, int64_t* rx_out, int64_t* offsets) {
    size_t off = 0;
    for (int k = 0; k < d; ++k) {
      const int64_t n = rx_[k] * P.M[k] * rx_[k + 1];
      TT_CHECK(off + (size_t)n <= out_elems, ERR_WORKSPACE, "output buffer too small");
      copy(x_[k], out + off, n, st_);
      offsets[k] = (int64_t)off;
      off += (size_t)n;
    }
    for (int k = 0; k <= d; ++k) rx_out[k] = rx_[k];
    TT_HIP(hipStreamSynchronize(st_));
  }

  ttamen_stats stats_{};

 private:
  const ProductPlan& P;
  const ProductOptions& O;
  Arena& ws_;
  hipStream_t st_;
  int d;
  std::vector<const double*> A_, b_;
  const double* const* x0_ = nullptr;
  const int64_t* rx0_ = nullptr;
  const double* const* z0_ = nullptr;
  std::vector<double*> Apm_, x_, z_, Phis_, Phisr_, Phiz_, Phizr_;
  double* spare_ = nullptr;
  double* slot_ = nullptr;
  RedScratch rs_;
  std::vector<int64_t> rx_, rz_;
  std::vector<double> normb_, normx_;
  double nrmsc_ = 1.0;
  uint64_t rand_calls_ = 0;

  double* host() {
    static thread_local double* h = nullptr;
    if (!h) TT_HIP(hipHostMalloc((void**)&h, sizeof(double) * 8192, hipHostMallocDefault));
    return h;
  }
  double read1(const double* dev) {
    read_to_host(host(), dev, sizeof(double), st_);
    return host()[0];
  }
  // ||x||, with 0 replaced by 1 when guard (the reference's `norm if norm > 0 else 1.0`)
  double norm(const double* x, int64_t n, bool guard) {
    norm2(x, n, slot_, rs_, st_, guard);
    return read1(slot_);
  }
  void scale(double* x, int64_t n, double a) { axpby(a, x, 0.0, x, n, st_); }

  // w[l][m][L] = sum PhiL[l,a,b] A_k[a,m,q,A] b_k[b,q,B] PhiR[L,A,B]   (_local_AB)
  void local_AB(int k, const double* PhiL, int l, const double* PhiR, int L, double* w) {
    LocalOp op;
    op.PhiL = PhiL; op.Apm = Apm_[k]; op.PhiR = PhiR;
    op.l1 = l; op.s = (int)P.rA[k]; op.r1 = (int)P.rb[k]; op.n = (int)P.K[k]; op.n_out = (int)P.M[k];
    op.S = (int)P.rA[k + 1]; op.l2 = L; op.r2 = (int)P.rb[k + 1];
    local_matvec(op, b_[k], w, 1, ws_, st_);
  }
  // out[z][p][Z] = sum L[z,r] c[r,p,R] Rm[Z,R]   (einsum 'zr,rmnR,ZR->zmnZ')
  void proj(const double* Lm, int z, const double* c, int r, int64_t p, int R, const double* Rm, int Z,
            double* out) {
    size_t mk = ws_.mark();
    double* t = ws_.alloc<double>((size_t)z * p * R);
    GemmDesc g;
    g.M = z; g.N = (int)(p * R); g.K = r;
    g.a = rm(Lm, r); g.b = rm(c, p * R); g.c = orm(t, p * R);
    gemm_f64(g, ws_, st_);
    GemmDesc h;
    h.M = (int)(z * p); h.N = Z; h.K = R;
    h.a = rm(t, R); h.b = rmT(Rm, R); h.c = orm(out, Z);
    gemm_f64(h, ws_, st_);
    ws_.release(mk);
  }
  void add_randn_cols(double* buf, int64_t rows, int64_t ld, int64_t c0, int k) {
    size_t mk = ws_.mark();
    const int64_t nc = ld - c0;
    double* t = ws_.alloc<double>(rows * nc);
    launch<KRandn>(dim3(grid_1d(rows * nc)), dim3(256), 0, st_,
                   {t, rows * nc, O.seed, (uint64_t)(1000003 * (++rand_calls_) + k)});
    TT_LAUNCH_CHECK();
    copy2d(t, nc, 1, buf + c0, ld, rows, nc, st_);
    ws_.release(mk);
  }

  void rl_orthogonal_z() {
    for (int i = d - 1; i > 0; --i) {
      size_t mk = ws_.mark();
      const int64_t r0 = rz_[i], m = P.M[i] * rz_[i + 1];
      const int64_t kq = std::min(m, r0);
      double* Q = ws_.alloc<double>(m * kq);
      double* R = ws_.alloc<double>(kq * r0);
      qr(z_[i], 1, m, m, r0, Q, R, ws_, st_);
      double* zn = ws_.alloc<double>(kq * m);
      copy2d(Q, 1, kq, zn, m, kq, m, st_);
      copy(zn, z_[i], kq * m, st_);
      const int64_t rp = rz_[i - 1], np = P.M[i - 1];
      double* t = ws_.alloc<double>(rp * np * kq);
      GemmDesc g;
      g.M = (int)(rp * np); g.N = (int)kq; g.K = (int)r0;
      g.a = rm(z_[i - 1], r0); g.b = rmT(R, r0); g.c = orm(t, kq);
      gemm_f64(g, ws_, st_);
      copy(t, z_[i - 1], rp * np * kq, st_);
      rz_[i] = kq;
      ws_.release(mk);
    }
  }

  // right-to-left step at core k (_amen.py:199-262)
  void rl_step(int k, int swp, bool last) {
    const int64_t Mk = P.M[k];
    if (!last && P.rz > 0) {
      size_t mk = ws_.mark();
      const int64_t rzk = rz_[k], rzk1 = rz_[k + 1];
      const int64_t cols = Mk * rzk1;
      const double* czp;
      int64_t sr, sc, ncz;
      if (swp > 0) {
        double* czAB = ws_.alloc<double>(rzk * cols);
        double* czx = ws_.alloc<double>(rzk * cols);
        local_AB(k, Phizr_[k], (int)rzk, Phizr_[k + 1], (int)rzk1, czAB);
        proj(Phiz_[k], (int)rzk, x_[k], (int)rx_[k], Mk, (int)rx_[k + 1], Phiz_[k + 1], (int)rzk1, czx);
        axpby(-1.0, czx, nrmsc_, czAB, rzk * cols, st_);  // czAB * nrmsc - czx
        const int64_t kc = std::min(rzk, cols);
        double* U = ws_.alloc<double>(rzk * kc);
        double* sv = ws_.alloc<double>(kc);
        double* V = ws_.alloc<double>(cols * kc);
        svd(czAB, cols, 1, rzk, cols, U, sv, V, ws_, st_);
        const int64_t kk = std::min<int64_t>(O.kickrank, kc);
        ncz = kk + ((k < d - 1) ? O.kick2 : 0);
        double* buf = ws_.alloc<double>(cols * std::max<int64_t>(ncz, 1));
        if (kk > 0) copy2d(V, kc, 1, buf, ncz, cols, kk, st_);
        if (ncz > kk) add_randn_cols(buf, cols, ncz, kk, k);
        czp = buf; sr = ncz; sc = 1;
      } else {
        czp = z_[k]; sr = 1; sc = cols; ncz = rzk;  // z_k reshaped (rz_k, M rz_k+1), transposed
      }
      if (ncz > 0) {
        const int64_t kq = std::min(cols, ncz);
        double* Qz = ws_.alloc<double>(cols * kq);
        double* Rz = ws_.alloc<double>(kq * ncz);
        qr(czp, sr, sc, cols, ncz, Qz, Rz, ws_, st_);
        copy2d(Qz, 1, kq, z_[k], cols, kq, cols, st_);  // z_k = Qz^T (kq, M, rz_k+1)
        rz_[k] = kq;
      }
      ws_.release(mk);
    }
    if (swp > 0) nrmsc_ = nrmsc_ * 1.0 * normx_[k - 1] / normb_[k - 1];
    // QR of x_k^T, R merged into x_{k-1}
    {
      size_t mk = ws_.mark();
      const int64_t r1 = rx_[k], m = Mk * rx_[k + 1];
      const int64_t kq = std::min(m, r1);
      double* Q = ws_.alloc<double>(m * kq);
      double* R = ws_.alloc<double>(kq * r1);
      qr(x_[k], 1, m, m, r1, Q, R, ws_, st_);
      const int64_t r0 = rx_[k - 1], np = P.M[k - 1];
      GemmDesc g;
      g.M = (int)(r0 * np); g.N = (int)kq; g.K = (int)r1;
      g.a = rm(x_[k - 1], r1); g.b = rmT(R, r1); g.c = orm(spare_, kq);
      gemm_f64(g, ws_, st_);
      std::swap(x_[k - 1], spare_);
      const double cn = norm(x_[k - 1], r0 * np * kq, true);
      scale(x_[k - 1], r0 * np * kq, 1.0 / cn);
      normx_[k - 1] *= cn;
      copy2d(Q, 1, kq, x_[k], m, kq, m, st_);  // x_k = Q^T
      rx_[k] = kq;
      ws_.release(mk);
    }
    const int64_t r1 = rx_[k], r2 = rx_[k + 1];
    phi_bck_rhs(Phis_[k + 1], x_[k], x_[k], Phis_[k], (int)r1, (int)r1, (int)Mk, (int)r2, (int)r2, ws_, st_);
    phi_bck_A_rect(Phisr_[k + 1], x_[k], A_[k], b_[k], Phisr_[k], (int)r1, (int)P.rA[k], (int)P.rb[k],
                   (int)Mk, (int)P.K[k], (int)P.rA[k + 1], (int)r2, (int)P.rb[k + 1], ws_, st_);
    const int64_t nphi = r1 * P.rA[k] * P.rb[k];
    const double nb = norm(Phisr_[k], nphi, true);
    normb_[k - 1] = nb;
    scale(Phisr_[k], nphi, 1.0 / nb);
    nrmsc_ = nrmsc_ * normb_[k - 1] / (1.0 * normx_[k - 1]);
    if (!last && P.rz > 0) {
      const int64_t rzk = rz_[k], rzk1 = rz_[k + 1];
      phi_bck_rhs(Phiz_[k + 1], z_[k], x_[k], Phiz_[k], (int)rzk, (int)r1, (int)Mk, (int)rzk1, (int)r2, ws_, st_);
      phi_bck_A_rect(Phizr_[k + 1], z_[k], A_[k], b_[k], Phizr_[k], (int)rzk, (int)P.rA[k], (int)P.rb[k],
                     (int)Mk, (int)P.K[k], (int)P.rA[k + 1], (int)rzk1, (int)P.rb[k + 1], ws_, st_);
      scale(Phizr_[k], rzk * P.rA[k] * P.rb[k], 1.0 / normb_[k - 1]);
    }
  }

  // left-to-right step at core k (_amen.py:265-388)
  void lr_step(int k, bool last, double& max_dx) {
    size_t mk = ws_.mark();
    const int64_t Mk = P.M[k], r1 = rx_[k], r2 = rx_[k + 1];
    const int64_t rows = r1 * Mk, Nl = rows * r2;
    double* sol = ws_.alloc<double>(Nl);
    local_AB(k, Phisr_[k], (int)r1, Phisr_[k + 1], (int)r2, sol);
    scale(sol, Nl, nrmsc_);
    const double nsol = norm(sol, Nl, false);
    diffnorm(sol, x_[k], Nl, nullptr, slot_, rs_, st_);
    const double dx = read1(slot_) / nsol;
    max_dx = (max_dx > dx) ? max_dx : dx;  // the reference's max(dx, max_dx), NaN included

    double* u;
    int64_t ldu, r;
    double* vmat;  // r2 x r, row-major
    if (k < d - 1) {
      const int64_t kc = std::min(rows, r2);
      double* U = ws_.alloc<double>(rows * kc);
      double* sv = ws_.alloc<double>(kc);
      double* V = ws_.alloc<double>(r2 * kc);
      svd(sol, r2, 1, rows, r2, U, sv, V, ws_, st_);
      read_to_host(host(), sv, sizeof(double) * kc, st_);
      const double tol = nsol * O.eps / std::pow((double)d, last ? 0.5 : 1.5);
      r = rank_chop(host(), (int)kc, tol);
      r = std::min<int64_t>(r, std::min<int64_t>(kc, P.caps[k + 1]));
      u = U; ldu = kc;
      vmat = ws_.alloc<double>(r2 * r);
      launch<KScaleCols>(dim3(grid_1d(r2 * r)), dim3(256), 0, st_, {V, kc, sv, vmat, r, r2, (int)r});
      TT_LAUNCH_CHECK();
    } else {
      const int64_t kq = std::min(rows, r2);
      double* Q = ws_.alloc<double>(rows * kq);
      double* R = ws_.alloc<double>(kq * r2);
      qr(sol, r2, 1, rows, r2, Q, R, ws_, st_);
      u = Q; ldu = kq; r = kq;
      vmat = ws_.alloc<double>(r2 * r);
      copy2d(R, 1, r2, vmat, r, r2, r, st_);  // vmat = R^T
    }
    // truncated solution tmp = u vmat^T (rows x r2)
    double* tmp = ws_.alloc<double>(Nl);
    {
      GemmDesc g;
      g.M = (int)rows; g.N = (int)r2; g.K = (int)r;
      g.a = rm(u, ldu); g.b = rmT(vmat, r); g.c = orm(tmp, r2);
      gemm_f64(g, ws_, st_);
    }
    if (!last && P.rz > 0) {
      // z_k from the residual proxy of the truncated solution
      const int64_t rzk = rz_[k], rzk1 = rz_[k + 1];
      const int64_t zrows = rzk * Mk;
      double* czAB = ws_.alloc<double>(zrows * rzk1);
      double* czx = ws_.alloc<double>(zrows * rzk1);
      local_AB(k, Phizr_[k], (int)rzk, Phizr_[k + 1], (int)rzk1, czAB);
      proj(Phiz_[k], (int)rzk, tmp, (int)r1, Mk, (int)r2, Phiz_[k + 1], (int)rzk1, czx);
      axpby(-1.0, czx, nrmsc_, czAB, zrows * rzk1, st_);
      const int64_t kcz = std::min(zrows, rzk1);
      double* Uz = ws_.alloc<double>(zrows * kcz);
      double* svz = ws_.alloc<double>(kcz);
      double* Vz = ws_.alloc<double>(rzk1 * kcz);
      svd(czAB, rzk1, 1, zrows, rzk1, Uz, svz, Vz, ws_, st_);
      const int64_t kk = std::min<int64_t>(O.kickrank, kcz);
      const int64_t ncol = kk + ((k < d - 1) ? O.kick2 : 0);
      if (ncol > 0) {
        double* cz = ws_.alloc<double>(zrows * ncol);
        if (kk > 0) copy2d(Uz, kcz, 1, cz, ncol, zrows, kk, st_);
        if (ncol > kk) add_randn_cols(cz, zrows, ncol, kk, k);
        const int64_t kq = std::min(zrows, ncol);
        double* Qz = ws_.alloc<double>(zrows * kq);
        double* Rz = ws_.alloc<double>(kq * ncol);
        qr(cz, ncol, 1, zrows, ncol, Qz, Rz, ws_, st_);
        copy(Qz, z_[k], zrows * kq, st_);
        rz_[k + 1] = kq;
      }
    }
    if (k < d - 1) {
      if (!last && P.rz > 0) {
        // enrichment with the residual projected on the z interface (:340-354)
        const int64_t radd = rz_[k + 1];
        double* uk = ws_.alloc<double>(rows * radd);
        double* ux = ws_.alloc<double>(rows * radd);
        local_AB(k, Phisr_[k], (int)r1, Phizr_[k + 1], (int)radd, uk);
        proj(Phis_[k], (int)r1, tmp, (int)r1, Mk, (int)r2, Phiz_[k + 1], (int)radd, ux);
        axpby(-1.0, ux, nrmsc_, uk, rows * radd, st_);
        const int64_t cw = r + radd;
        double* cat = ws_.alloc<double>(rows * cw);
        copy2d(u, ldu, 1, cat, cw, rows, r, st_);
        copy2d(uk, radd, 1, cat + r, cw, rows, radd, st_);
        const int64_t kq = std::min(rows, cw);
        double* Qe = ws_.alloc<double>(rows * kq);
        double* Re = ws_.alloc<double>(kq * cw);
        qr(cat, cw, 1, rows, cw, Qe, Re, ws_, st_);
        double* v2 = ws_.alloc<double>(r2 * kq);
        GemmDesc g;  // v2 = [vmat, 0] Re^T = vmat Re[:, :r]^T
        g.M = (int)r2; g.N = (int)kq; g.K = (int)r;
        g.a = rm(vmat, r); g.b = rmT(Re, cw); g.c = orm(v2, kq);
        gemm_f64(g, ws_, st_);
        u = Qe; ldu = kq; vmat = v2; r = kq;
      }
      TT_CHECK(r <= P.Rb[k + 1], ERR_UNSUPPORTED,
               "the product needs rank " + std::to_string(r) + " at interface " + std::to_string(k + 1) +
                   "; ranks above " + std::to_string(P.Rb[k + 1]) +
                   " are not supported at this rank capacity (ttamen_rank_capacity raises it up to " +
                   std::to_string(kMaxRank) + "; or lower rmax)");
      // v = vmat^T x_{k+1}  (r x M' x r3)
      const int64_t n1 = P.M[k + 1], r3 = rx_[k + 2];
      {
        GemmDesc g;
        g.M = (int)r; g.N = (int)(n1 * r3); g.K = (int)r2;
        g.a = rmT(vmat, r); g.b = rm(x_[k + 1], n1 * r3); g.c = orm(spare_, n1 * r3);
        gemm_f64(g, ws_, st_);
      }
      nrmsc_ = nrmsc_ * 1.0 * normx_[k] / normb_[k];
      const double nn = norm(spare_, r * n1 * r3, true);
      scale(spare_, r * n1 * r3, 1.0 / nn);
      normx_[k] *= nn;
      copy2d(u, ldu, 1, x_[k], r, rows, r, st_);
      std::swap(x_[k + 1], spare_);
      rx_[k + 1] = r;
      phi_fwd_rhs(Phis_[k], x_[k], x_[k], Phis_[k + 1], (int)r1, (int)r1, (int)Mk, (int)r, (int)r, ws_, st_);
      phi_fwd_A_rect(Phisr_[k], x_[k], Apm_[k], b_[k], Phisr_[k + 1], (int)r1, (int)P.rA[k], (int)P.rb[k],
                     (int)Mk, (int)P.K[k], (int)P.rA[k + 1], (int)r, (int)P.rb[k + 1], ws_, st_);
      const int64_t nphi = r * P.rA[k + 1] * P.rb[k + 1];
      const double nb = norm(Phisr_[k + 1], nphi, true);
      normb_[k] = nb;
      scale(Phisr_[k + 1], nphi, 1.0 / nb);
      nrmsc_ = nrmsc_ * normb_[k] / (1.0 * normx_[k]);
      if (!last && P.rz > 0) {
        const int64_t rzk = rz_[k], rzk1 = rz_[k + 1];
        phi_fwd_rhs(Phiz_[k], z_[k], x_[k], Phiz_[k + 1], (int)rzk, (int)r1, (int)Mk, (int)rzk1, (int)r, ws_, st_);
        phi_fwd_A_rect(Phizr_[k], z_[k], Apm_[k], b_[k], Phizr_[k + 1], (int)rzk, (int)P.rA[k], (int)P.rb[k],
                       (int)Mk, (int)P.K[k], (int)P.rA[k + 1], (int)rzk1, (int)P.rb[k + 1], ws_, st_);
        scale(Phizr_[k + 1], rzk1 * P.rA[k + 1] * P.rb[k + 1], 1.0 / normb_[k]);
      }
    } else {
      copy(tmp, x_[k], Nl, st_);
    }
    ws_.release(mk);
  }
};

size_t product_workspace_bound(const ProductPlan& P, const ProductOptions& O) {
  Arena meas;
  {
    ProductSweep s(P, O, meas, nullptr);
    std::vector<const double*> dummy(P.d, nullptr);
    s.allocate(dummy.data(), dummy.data(), nullptr, nullptr, nullptr);
  }
  const double R = (double)P.maxR, M = (double)P.maxM, K = (double)P.maxK;
  const double ra = (double)P.maxrA, rb = (double)P.maxrb, rz = (double)std::max<int64_t>(P.rz, 1);
  const double rows = R * M, cw = R + rz;
  double doubles = 0;
  doubles += 2 * R * ra * rb * (K + M);              // local product / interface intermediates
  doubles += 6 * rows * R;                           // sol, tmp, SVD U and Q, products
  doubles += 4 * R * R + 48.0 * (R + 2) * (R / 2 + 2) * 2;  // SVD factors and rotation log
  doubles += 3 * rows * cw + cw * cw + 2 * rows * rz;        // enrichment
  doubles += 8 * rz * M * rz + 4 * rz * rz;                  // residual proxy
  const size_t step = (size_t)(doubles * 8.0 * 1.25);
  return meas.peak() + step + (48ull << 20) + (16ull << 20);
}

size_t product_output_elems(const ProductPlan& P) {
  size_t out = 0;
  for (int k = 0; k < P.d; ++k) out += (size_t)(P.Rb[k] * P.M[k] * P.Rb[k + 1]);
  return out;
}

void product_run(const ProductPlan& P, const ProductOptions& O, const double* const* A,
                 const double* const* b, const double* const* x0, const int64_t* rx0,
                 const double* const* z0, void* workspace, size_t workspace_bytes, double* x_out,
                 size_t x_out_elems, int64_t* rx_out, int64_t* offsets, ttamen_stats* stats,
                 hipStream_t st) {
  auto t0 = std::chrono::high_resolution_clock::now();
  TT_CHECK(A && b && x_out && rx_out && offsets && workspace, ERR_INVALID, "null argument");
  linalg_init();
  Arena ws(workspace, workspace_bytes);
  ProductSweep s(P, O, ws, st);
  s.allocate(A, b, x0, rx0, z0);
  s.run();
  s.write_out(x_out, x_out_elems, rx_out, offsets);
  if (stats) {
    *stats = s.stats_;
    stats->elapsed_s = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
  }
}

}  // namespace
}  // namespace ttamen

using namespace ttamen;

extern "C" {

size_t ttamen_amen_mv_workspace_size(int d, const int64_t* M, const int64_t* K, const int64_t* rA,
                                     const int64_t* rb, const int64_t* rx0, const int64_t* rmax,
                                     int64_t kickrank, int64_t kick2) {
  size_t out = 0;
  int rc = abi_call([&] {
    ProductPlan P = make_product_plan(d, M, K, rA, rb, rx0, rmax, kickrank, kick2);
    ProductOptions O;
    O.kickrank = kickrank; O.kick2 = kick2;
    out = product_workspace_bound(P, O);
  });
  return rc == 0 ? out : 0;
}

size_t ttamen_amen_mv_output_elems(int d, const int64_t* M, const int64_t* K, const int64_t* rx0,
                                   const int64_t* rmax, int64_t kickrank, int64_t kick2) {
  size_t out = 0;
  int rc = abi_call([&] {
    std::vector<int64_t> ones(d + 1, 1);
    ProductPlan P = make_product_plan(d, M, K, ones.data(), ones.data(), rx0, rmax, kickrank, kick2);
    out = product_output_elems(P);
  });
  return rc == 0 ? out : 0;
}

int ttamen_amen_mv(int d, const int64_t* M, const int64_t* K, const int64_t* rA, const int64_t* rb,
                   const double* const* A_cores, const double* const* b_cores,
                   const double* const* x0_cores, const int64_t* rx0, const double* const* z0_cores,
                   int64_t nswp, double eps, const int64_t* rmax, int64_t kickrank, int64_t kick2,
                   uint64_t seed, int verbose, void* workspace, size_t workspace_bytes, double* x_out,
                   size_t x_out_elems, int64_t* rx_out, int64_t* offsets_out, ttamen_stats* stats,
                   void* stream) {
  return abi_call([&] {
    TT_CHECK(nswp >= 0 && eps >= 0, ERR_INVALID, "invalid solver parameters");
    ProductPlan P = make_product_plan(d, M, K, rA, rb, x0_cores ? rx0 : nullptr, rmax, kickrank, kick2);
    ProductOptions O;
    O.nswp = nswp; O.eps = eps; O.kickrank = kickrank; O.kick2 = kick2; O.seed = seed; O.verbose = verbose;
    product_run(P, O, A_cores, b_cores, x0_cores, rx0, z0_cores, workspace, workspace_bytes, x_out,
                x_out_elems, rx_out, offsets_out, stats, (hipStream_t)stream);
  });
}

size_t ttamen_amen_mm_workspace_size(int d, const int64_t* M, const int64_t* K, const int64_t* N,
                                     const int64_t* rA, const int64_t* rB, const int64_t* rx0,
                                     const int64_t* rmax, int64_t kickrank, int64_t kick2) {
  size_t out = 0;
  int rc = abi_call([&] {
    TT_CHECK(d >= 1 && M && K && N && rA, ERR_INVALID, "null argument");
    std::vector<int64_t> MN(d), KN(d);
    size_t ops = 0;
    for (int k = 0; k < d; ++k) {
      MN[k] = M[k] * N[k];
      KN[k] = K[k] * N[k];
      ops += ((size_t)(rA[k] * MN[k] * KN[k] * rA[k + 1]) * sizeof(double) + 255) / 256 * 256;
    }
    ProductPlan P = make_product_plan(d, MN.data(), KN.data(), rA, rB, rx0, rmax, kickrank, kick2);
    ProductOptions O;
    O.kickrank = kickrank; O.kick2 = kick2;
    out = ops + product_workspace_bound(P, O);
  });
  return rc == 0 ? out : 0;
}

int ttamen_amen_mm(int d, const int64_t* M, const int64_t* K, const int64_t* N, const int64_t* rA,
                   const int64_t* rB, const double* const* A_cores, const double* const* B_cores,
                   const double* const* X0_cores, const int64_t* rx0, const double* const* z0_cores,
                   int64_t nswp, double eps, const int64_t* rmax, int64_t kickrank, int64_t kick2,
                   uint64_t seed, int verbose, void* workspace, size_t workspace_bytes, double* x_out,
                   size_t x_out_elems, int64_t* rx_out, int64_t* offsets_out, ttamen_stats* stats,
                   void* stream) {
  return abi_call([&] {
    TT_CHECK(d >= 1 && M && K && N && rA && rB && A_cores && B_cores && workspace, ERR_INVALID, "null argument");
    TT_CHECK(nswp >= 0 && eps >= 0, ERR_INVALID, "invalid solver parameters");
    hipStream_t st = (hipStream_t)stream;
    std::vector<int64_t> MN(d), KN(d);
    for (int k = 0; k < d; ++k) {
      TT_CHECK(M[k] >= 1 && K[k] >= 1 && N[k] >= 1 && A_cores[k], ERR_INVALID, "bad operand cores");
      MN[k] = M[k] * N[k];
      KN[k] = K[k] * N[k];
    }
    ProductPlan P = make_product_plan(d, MN.data(), KN.data(), rA, rB, X0_cores ? rx0 : nullptr, rmax,
                                      kickrank, kick2);
    // the operator A (x) I_N at the front of the workspace
    char* w = (char*)workspace;
    size_t used = 0;
    std::vector<const double*> D(d);
    for (int k = 0; k < d; ++k) {
      const int64_t total = rA[k] * MN[k] * KN[k] * rA[k + 1];
      const size_t bytes = ((size_t)total * sizeof(double) + 255) / 256 * 256;
      TT_CHECK(used + bytes < workspace_bytes, ERR_WORKSPACE, "workspace too small");
      double* Dk = (double*)(w + used);
      launch<KKronEye>(dim3(grid_1d(total)), dim3(256), 0, st,
                       {A_cores[k], Dk, (int)rA[k], (int)M[k], (int)K[k], (int)N[k], (int)rA[k + 1]});
      TT_LAUNCH_CHECK();
      D[k] = Dk;
      used += bytes;
    }
    ProductOptions O;
    O.nswp = nswp; O.eps = eps; O.kickrank = kickrank; O.kick2 = kick2; O.seed = seed; O.verbose = verbose;
    product_run(P, O, D.data(), B_cores, X0_cores, rx0, z0_cores, w + used, workspace_bytes - used, x_out,
                x_out_elems, rx_out, offsets_out, stats, st);
  });
}

}  // extern "C"
