// DMRG approximation of a TT matrix-vector product y ~= A x (reference cpp/dmrg_mv.h:23-230, the
// native path of torchtt/_dmrg.py:22-45; Python twin :47-206).
//
// Each sweep orthogonalises y right to left (QR of every core, R merged into its left
// neighbour, the interface Phi_k = y_k^T A_k x_k Phi_k+1 updated), then visits the supercores
// (k, k+1) left to right:
//   W       = the projection of A x onto the interfaces of y at (k, k+1)
//             (W = y_k y_k+1 itself in the last sweep),
//   delta_k = ||W - y_k y_k+1|| / ||W||, which grows or shrinks the per-core rank enlargement,
//   W       ~= U S V^T truncated by rank_chop at ||W|| eps / d^1.5 (d^0.5 in the last sweep)
//             plus the enlargement (none in the last sweep), capped by rmax,
//   y_k     = the left factor, with kickrank normal columns appended and re-orthogonalised by
//             QR in every sweep but the final one (nswp - 1); y_k+1 = the matching right factor,
//   Phi_k+1 = y_k^T A_k x_k Phi_k.
// The sweep after the one whose max delta fell below eps is the last.
//
// Mapping: W is the product F1 F2 of its two halves, F1[(l,m),(S,R)] = Phi_k A_k x_k (stages
// one and two of the forward interface update) and F2[(S,R),(m',L)] = A_k+1 x_k+1 Phi_k+2 (one
// GEMM and one batched GEMM), so no six-tensor contraction is materialised; Phi_k+1 = y_k^T F1
// reuses F1. When the inner rank c = rA_k+1 rx_k+1 is below both sides of W, the SVD is taken
// through the factors (QR of F1 and of F2^T, SVD of the c x c product of the triangles): exact,
// cheaper, and every QR stays within the LDS-resident Cholesky limit (kMaxSmall columns). The
// factored SVD has c singular values where the reference's dense SVD has min(rows, cols), the
// rest zero; rank_chop sees them padded with zeros, and a rank above c is completed with
// orthonormal directions of zero weight (the dead-column completion of the QR), as the dense
// SVD's trailing singular vectors are.
//
// Random numbers: the kick columns come from the device's SplitMix64 / Box-Muller counter
// stream (stream id (1 << 40) | (sweep << 20) | k, element i * kickrank + j); an initial guess of
// rank 2 when none is passed, from stream (2 << 40) | k.
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

constexpr uint64_t kKickStream = uint64_t(1) << 40;
constexpr uint64_t kGuessStream = uint64_t(2) << 40;

struct DmrgPlan {
  int d = 0;
  std::vector<int64_t> M, N, rA, rx, caps, Rb;
  int64_t kick = 0;
};

// Rank bounds per interface: y's ranks never exceed the left product of mode sizes, the
// right product (or rmax) plus the kick, nor kMaxSmall (every QR of the sweep keeps its
// triangle in LDS); the initial guess's ranks are allowed as they come.
DmrgPlan make_dmrg_plan(int d, const int64_t* M, const int64_t* N, const int64_t* rA, const int64_t* rx,
                        const int64_t* ry0, const int64_t* caps, int64_t kickrank) {
  TT_CHECK(d >= 2, ERR_INVALID, "dmrg_mv needs at least two cores");
  TT_CHECK(M && N && rA && rx && caps, ERR_INVALID, "null argument");
  TT_CHECK(kickrank >= 0, ERR_INVALID, "kickrank must be non-negative");
  DmrgPlan p;
  p.d = d;
  p.M.assign(M, M + d);
  p.N.assign(N, N + d);
  p.rA.assign(rA, rA + d + 1);
  p.rx.assign(rx, rx + d + 1);
  p.caps.assign(caps, caps + d + 1);
  p.kick = kickrank;
  for (int k = 0; k < d; ++k) TT_CHECK(M[k] >= 1 && N[k] >= 1, ERR_INVALID, "mode sizes must be positive");
  for (int k = 0; k <= d; ++k) TT_CHECK(rA[k] >= 1 && rx[k] >= 1, ERR_INVALID, "ranks must be positive");
  TT_CHECK(rA[0] == 1 && rA[d] == 1 && rx[0] == 1 && rx[d] == 1, ERR_INVALID, "boundary ranks must be 1");
  for (int k = 1; k < d; ++k) TT_CHECK(caps[k] >= 1, ERR_INVALID, "rank caps must be positive");
  if (ry0) {
    TT_CHECK(ry0[0] == 1 && ry0[d] == 1, ERR_INVALID, "boundary ranks of the initial guess must be 1");
    for (int k = 1; k < d; ++k)
      TT_CHECK(ry0[k] >= 1 && ry0[k] <= kMaxSmall, ERR_UNSUPPORTED,
               "initial-guess ranks must lie in [1, " + std::to_string(kMaxSmall) + "]");
  }
  const int64_t big = int64_t(1) << 40;
  std::vector<int64_t> right(d + 1, 1);
  for (int k = d - 1; k >= 0; --k) right[k] = std::min<int64_t>(right[k + 1] * M[k], big);
  p.Rb.assign(d + 1, 1);
  for (int k = 1; k < d; ++k) {
    int64_t b = std::min<int64_t>(std::min<int64_t>(p.Rb[k - 1] * M[k - 1], big),
                                  std::min(caps[k], right[k]) + kickrank);
    b = std::min<int64_t>(b, kMaxSmall);
    b = std::max<int64_t>(b, ry0 ? ry0[k] : 2);
    p.Rb[k] = std::max<int64_t>(b, 1);
  }
  return p;
}

struct DmrgOptions {
  int64_t nswp = 20, kickrank = 4;
  double eps = 1e-12;
  uint64_t seed = 0;
  int verbose = 0;
};

class DmrgSweep {
 public:
  DmrgSweep(const DmrgPlan& p, const DmrgOptions& o, Arena& ws, hipStream_t st)
      : P(p), O(o), ws_(ws), st_(st), d(p.d) {}

  void allocate(const double* const* A, const double* const* x, const double* const* y0, const int64_t* ry0) {
    A_.assign(A, A + d);
    x_.assign(x, x + d);
    y0_ = y0;
    ry0_ = ry0;
    Apm_.resize(d);
    y_.resize(d);
    int64_t maxcore = 1;
    for (int k = 0; k < d; ++k) maxcore = std::max(maxcore, P.Rb[k] * P.M[k] * P.Rb[k + 1]);
    // every core slot holds the largest core: the right-to-left step rotates the slots
    // (the product y_k-1 R^T is written to the spare slot, which becomes y_k-1)
    for (int k = 0; k < d; ++k) {
      Apm_[k] = ws_.alloc<double>(P.rA[k] * P.M[k] * P.N[k] * P.rA[k + 1]);
      y_[k] = ws_.alloc<double>(maxcore);
    }
    spare_ = ws_.alloc<double>(maxcore);
    Phis_.resize(d + 1);
    for (int k = 0; k <= d; ++k) Phis_[k] = ws_.alloc<double>(P.Rb[k] * P.rA[k] * P.rx[k]);
    slot_ = ws_.alloc<double>(8);
    rs_.partial = ws_.alloc<double>(kRedBlocks * 64 + 64);
    rs_.counter = ws_.alloc<unsigned>(64);
  }

  void run() {
    if (ws_.measuring()) return;
    zero_words(rs_.counter, 64, st_);
    for (int k = 0; k < d; ++k)
      permute_core_A_rect(A_[k], Apm_[k], (int)P.rA[k], (int)P.M[k], (int)P.N[k], (int)P.rA[k + 1], st_);
    ry_.assign(d + 1, 1);
    if (y0_) {
      for (int k = 0; k <= d; ++k) ry_[k] = ry0_[k];
      for (int k = 0; k < d; ++k) copy(y0_[k], y_[k], ry_[k] * P.M[k] * ry_[k + 1], st_);
    } else {
      for (int k = 1; k < d; ++k) ry_[k] = 2;  // dmrg_mv.h:43-51
      for (int k = 0; k < d; ++k) {
        const int64_t n = ry_[k] * P.M[k] * ry_[k + 1];
        launch<KRandn>(dim3(grid_1d(n)), dim3(256), 0, st_, {y_[k], n, O.seed, kGuessStream | (uint64_t)k});
        TT_LAUNCH_CHECK();
      }
    }
    fill(Phis_[0], 1, 1.0, st_);
    fill(Phis_[d], 1, 1.0, st_);
    delta_.assign(d - 1, 1.0);
    delta_prev_.assign(d - 1, 1.0);
    r_enl_.assign(d, 2);
    bool last = false;
    int swp;
    for (swp = 0; swp < O.nswp; ++swp) {
      if (O.verbose) printf("Sweep %d\n", swp);
      for (int k = d - 1; k > 0; --k) rl_step(k);
      for (int k = 0; k < d - 1; ++k) lr_step(k, swp, last);
      stats_.sweeps = swp + 1;
      stats_.max_res_last = *std::max_element(delta_.begin(), delta_.end());
      if (last) break;
      if (stats_.max_res_last < O.eps) last = true;
      delta_prev_ = delta_;
    }
    if (O.verbose) {
      printf("Finished after %d sweeps: rank [", stats_.sweeps ? (int)stats_.sweeps : 0);
      for (auto r : ry_) printf(" %lld", (long long)r);
      printf(" ]\n");
      fflush(stdout);
    }
    for (auto r : ry_) stats_.max_rank = std::max<int64_t>(stats_.max_rank, r);
  }

  void write_out(double* out, size_t out_elems, int64_t* ry_out, int64_t* offsets) {
    size_t off = 0;
    for (int k = 0; k < d; ++k) {
      const int64_t n = ry_[k] * P.M[k] * ry_[k + 1];
      TT_CHECK(off + (size_t)n <= out_elems, ERR_WORKSPACE, "output buffer too small");
      copy(y_[k], out + off, n, st_);
      offsets[k] = (int64_t)off;
      off += (size_t)n;
    }
    for (int k = 0; k <= d; ++k) ry_out[k] = ry_[k];
    TT_HIP(hipStreamSynchronize(st_));
  }

  ttamen_stats stats_{};

 private:
  const DmrgPlan& P;
  const DmrgOptions& O;
  Arena& ws_;
  hipStream_t st_;
  int d;
  std::vector<const double*> A_, x_;
  const double* const* y0_ = nullptr;
  const int64_t* ry0_ = nullptr;
  std::vector<double*> Apm_, y_, Phis_;
  double* spare_ = nullptr;
  double* slot_ = nullptr;
  RedScratch rs_;
  std::vector<int64_t> ry_, r_enl_;
  std::vector<double> delta_, delta_prev_;

  double* host() {  // pinned staging for device scalars and singular values (kMaxSmall of them)
    static thread_local double* h = nullptr;
    if (!h) TT_HIP(hipHostMalloc((void**)&h, sizeof(double) * 8192, hipHostMallocDefault));
    return h;
  }
  double read1(const double* dev) {
    read_to_host(host(), dev, sizeof(double), st_);
    return host()[0];
  }

  // right-to-left step at core k (dmrg_mv.h:87-108): y_k^T = Q R, y_k = Q^T, y_k-1 = y_k-1 R^T
  void rl_step(int k) {
    size_t mk = ws_.mark();
    const int64_t Mk = P.M[k], r1 = ry_[k], m = Mk * ry_[k + 1];
    const int64_t kq = std::min(m, r1);
    double* Q = ws_.alloc<double>(m * kq);
    double* R = ws_.alloc<double>(kq * r1);
    qr(y_[k], 1, m, m, r1, Q, R, ws_, st_);
    const int64_t r0 = ry_[k - 1], np = P.M[k - 1];
    {
      GemmDesc g;
      g.M = (int)(r0 * np); g.N = (int)kq; g.K = (int)r1;
      g.a = rm(y_[k - 1], r1); g.b = rmT(R, r1); g.c = orm(spare_, kq);
      gemm_f64(g, ws_, st_);
    }
    std::swap(y_[k - 1], spare_);
    copy2d(Q, 1, kq, y_[k], m, kq, m, st_);  // y_k = Q^T (kq, M, ry_k+1)
    ry_[k] = kq;
    phi_bck_A_rect(Phis_[k + 1], y_[k], A_[k], x_[k], Phis_[k], (int)kq, (int)P.rA[k], (int)P.rx[k], (int)Mk,
                   (int)P.N[k], (int)P.rA[k + 1], (int)ry_[k + 1], (int)P.rx[k + 1], ws_, st_);
    ws_.release(mk);
  }

  // F2[(s,r)][(m,L)] = sum A[s,m,n,S] x[r,n,R] Phi[L,S,R] for core j (the right half of W)
  void right_half(int j, double* F2) {
    const int s = (int)P.rA[j], r = (int)P.rx[j], nm = (int)P.M[j], nk = (int)P.N[j];
    const int S = (int)P.rA[j + 1], R = (int)P.rx[j + 1], L = (int)ry_[j + 1];
    size_t mk = ws_.mark();
    double* u1 = ws_.alloc<double>((size_t)r * nk * S * L);
    {
      GemmDesc g;  // u1[(r,n)][S][L] = sum_R x[(r,n)][R] Phi[L][S][R]   (batch S)
      g.M = r * nk; g.N = L; g.K = R; g.B2 = S;
      g.a.p = x_[j]; g.a.r = R; g.a.c = 1;
      g.b.p = Phis_[j + 1]; g.b.r = 1; g.b.c = (int64_t)S * R; g.b.b2 = R;
      g.c.p = u1; g.c.r = (int64_t)S * L; g.c.c = 1; g.c.b2 = L;
      gemm_f64(g, ws_, st_);
    }
    {
      GemmDesc g;  // F2[s][r][m][L] = sum_{(n,S)} A[s][m][(n,S)] u1[r][(n,S)][L]   (batch r, s)
      g.M = nm; g.N = L; g.K = nk * S; g.B1 = r; g.B2 = s;
      g.a.p = A_[j]; g.a.r = (int64_t)nk * S; g.a.c = 1; g.a.b2 = (int64_t)nm * nk * S;
      g.b.p = u1; g.b.r = L; g.b.c = 1; g.b.b1 = (int64_t)nk * S * L;
      g.c.p = F2; g.c.r = L; g.c.c = 1; g.c.b1 = (int64_t)nm * L; g.c.b2 = (int64_t)r * nm * L;
      gemm_f64(g, ws_, st_);
    }
    ws_.release(mk);
  }

  // left-to-right step at supercore (k, k+1) (dmrg_mv.h:111-210)
  void lr_step(int k, int swp, bool last) {
    size_t mk = ws_.mark();
    const int64_t Mk = P.M[k], Mk1 = P.M[k + 1];
    const int64_t m1 = ry_[k] * Mk, n2 = Mk1 * ry_[k + 2], rmid = ry_[k + 1];
    const int64_t kd = std::min(m1, n2);
    double* Wp = ws_.alloc<double>(m1 * n2);
    {
      GemmDesc g;  // W_prev = y_k y_k+1
      g.M = (int)m1; g.N = (int)n2; g.K = (int)rmid;
      g.a = rm(y_[k], rmid); g.b = rm(y_[k + 1], n2); g.c = orm(Wp, n2);
      gemm_f64(g, ws_, st_);
    }
    const double* W = Wp;
    const double *fa = y_[k], *fb = y_[k + 1];
    double* F1 = nullptr;
    int64_t c = rmid;
    if (!last) {
      c = P.rA[k + 1] * P.rx[k + 1];
      F1 = ws_.alloc<double>(m1 * c);
      double* F2 = ws_.alloc<double>(c * n2);
      phi_fwd_half_rect(Phis_[k], Apm_[k], x_[k], F1, (int)ry_[k], (int)P.rA[k], (int)P.rx[k], (int)Mk,
                        (int)P.N[k], (int)P.rA[k + 1], (int)P.rx[k + 1], ws_, st_);
      right_half(k + 1, F2);
      double* Wn = ws_.alloc<double>(m1 * n2);
      GemmDesc g;  // W = F1 F2
      g.M = (int)m1; g.N = (int)n2; g.K = (int)c;
      g.a = rm(F1, c); g.b = rm(F2, n2); g.c = orm(Wn, n2);
      gemm_f64(g, ws_, st_);
      W = Wn;
      fa = F1;
      fb = F2;
    }
    norm2(W, m1 * n2, slot_, rs_, st_, false);
    const double b = read1(slot_);
    if (b > 0) {
      diffnorm(W, Wp, m1 * n2, nullptr, slot_, rs_, st_);
      delta_[k] = read1(slot_) / b;
    } else {
      delta_[k] = 0.0;
    }
    if (delta_[k] / delta_prev_[k] >= 1 && delta_[k] > O.eps) r_enl_[k] += 1;
    if (delta_[k] / delta_prev_[k] < 0.1 && delta_[k] < O.eps) r_enl_[k] = std::max<int64_t>(r_enl_[k] - 1, 1);

    // SVD W = U diag(sv) V^T with ns singular triplets (U: m1 x ns, V: n2 x ns, row-major)
    double *U, *sv, *V;
    int64_t ns;
    if (c < kd && c <= kMaxSmall) {
      ns = c;
      double* Q1 = ws_.alloc<double>(m1 * c);
      double* R1 = ws_.alloc<double>(c * c);
      double* Q2 = ws_.alloc<double>(n2 * c);
      double* R2 = ws_.alloc<double>(c * c);
      qr(fa, c, 1, m1, c, Q1, R1, ws_, st_);
      qr(fb, 1, n2, n2, c, Q2, R2, ws_, st_);
      double* Cm = ws_.alloc<double>(c * c);
      {
        GemmDesc g;  // Cm = R1 R2^T
        g.M = (int)c; g.N = (int)c; g.K = (int)c;
        g.a = rm(R1, c); g.b = rmT(R2, c); g.c = orm(Cm, c);
        gemm_f64(g, ws_, st_);
      }
      double* Uc = ws_.alloc<double>(c * c);
      double* Vc = ws_.alloc<double>(c * c);
      sv = ws_.alloc<double>(c);
      svd(Cm, c, 1, c, c, Uc, sv, Vc, ws_, st_);
      U = ws_.alloc<double>(m1 * c);
      V = ws_.alloc<double>(n2 * c);
      GemmDesc g;
      g.M = (int)m1; g.N = (int)c; g.K = (int)c;
      g.a = rm(Q1, c); g.b = rm(Uc, c); g.c = orm(U, c);
      gemm_f64(g, ws_, st_);
      GemmDesc h;
      h.M = (int)n2; h.N = (int)c; h.K = (int)c;
      h.a = rm(Q2, c); h.b = rm(Vc, c); h.c = orm(V, c);
      gemm_f64(h, ws_, st_);
    } else {
      TT_CHECK(kd <= kMaxSmall, ERR_UNSUPPORTED,
               "dmrg_mv: supercore " + std::to_string(k) + " is " + std::to_string(m1) + " x " + std::to_string(n2) +
                   " with inner rank " + std::to_string(c) + "; SVDs above " + std::to_string(kMaxSmall) +
                   " columns are not supported");
      ns = kd;
      U = ws_.alloc<double>(m1 * kd);
      V = ws_.alloc<double>(n2 * kd);
      sv = ws_.alloc<double>(kd);
      svd(W, n2, 1, m1, n2, U, sv, V, ws_, st_);
    }
    read_to_host(host(), sv, sizeof(double) * ns, st_);
    const double thr = b * O.eps / std::pow((double)d, last ? 0.5 : 1.5);
    // The dense SVD has min(m1, n2) = kd values, zeros past ns. For thr > 0 rank_chop walks down
    // through zero tails to ns - 1 and then sees the same sums, so the ns values decide; for
    // thr <= 0 its answer (kd - 1, or 1 for a zero W) needs the padded length.
    int64_t chop;
    if (thr > 0) {
      chop = rank_chop(host(), (int)ns, thr);
    } else {
      std::vector<double> s(kd, 0.0);
      for (int64_t i = 0; i < ns; ++i) s[i] = host()[i];
      chop = rank_chop(s.data(), (int)kd, thr);
    }
    int64_t r_new = chop + (last ? 0 : r_enl_[k]);
    r_new = std::min(std::min(r_new, kd), P.caps[k + 1]);
    r_new = std::max<int64_t>(r_new, 1);
    const int64_t kick = (swp < O.nswp - 1) ? O.kickrank : 0;
    const bool need_qr = kick > 0 || r_new > ns;
    if (need_qr && m1 >= r_new + kick && r_new + kick > kMaxSmall) {
      const int64_t lim = kMaxSmall - kick;
      TT_CHECK(std::min(chop, P.caps[k + 1]) <= lim, ERR_UNSUPPORTED,
               "dmrg_mv: rank " + std::to_string(std::min(chop, P.caps[k + 1])) + " at interface " +
                   std::to_string(k + 1) + " exceeds " + std::to_string(lim) + " (lower rmax)");
      r_new = lim;  // the enlargement is cut back, the truncation rank is kept
    }
    const int64_t rr = std::min(r_new, ns);
    double* W2s = ws_.alloc<double>(n2 * rr);  // V[:, :rr] diag(s[:rr])
    launch<KScaleCols>(dim3(grid_1d(n2 * rr)), dim3(256), 0, st_, {V, ns, sv, W2s, rr, n2, (int)rr});
    TT_LAUNCH_CHECK();
    if (need_qr) {
      // [U_r, 0, kick] = Q Rc; y_k = Q, y_k+1 = Rc [W2, 0, 0]^T (dmrg_mv.h:177-185)
      const int64_t ncol = r_new + kick;
      double* cat = ws_.alloc<double>(m1 * ncol);
      fill(cat, m1 * ncol, 0.0, st_);
      copy2d(U, ns, 1, cat, ncol, m1, rr, st_);
      if (kick > 0) {
        double* t = ws_.alloc<double>(m1 * kick);
        launch<KRandn>(dim3(grid_1d(m1 * kick)), dim3(256), 0, st_,
                       {t, m1 * kick, O.seed, kKickStream | ((uint64_t)swp << 20) | (uint64_t)k});
        TT_LAUNCH_CHECK();
        copy2d(t, kick, 1, cat + r_new, ncol, m1, kick, st_);
      }
      const int64_t kq = std::min(m1, ncol);
      TT_CHECK(kq <= P.Rb[k + 1], ERR_UNSUPPORTED, "dmrg_mv: rank above the planned bound");
      double* Qc = ws_.alloc<double>(m1 * kq);
      double* Rc = ws_.alloc<double>(kq * ncol);
      qr(cat, ncol, 1, m1, ncol, Qc, Rc, ws_, st_);
      copy(Qc, y_[k], m1 * kq, st_);
      GemmDesc g;  // y_k+1 = Rc[:, :rr] W2s^T
      g.M = (int)kq; g.N = (int)n2; g.K = (int)rr;
      g.a = rm(Rc, ncol); g.b = rmT(W2s, rr); g.c = orm(y_[k + 1], n2);
      gemm_f64(g, ws_, st_);
      r_new = kq;
    } else {
      TT_CHECK(r_new <= P.Rb[k + 1], ERR_UNSUPPORTED, "dmrg_mv: rank above the planned bound");
      copy2d(U, ns, 1, y_[k], r_new, m1, r_new, st_);
      copy2d(W2s, 1, rr, y_[k + 1], n2, rr, n2, st_);  // y_k+1 = W2s^T
    }
    if (O.verbose) {
      printf("\tcore %d: delta %g rank %lld -> %lld\n", k, delta_[k], (long long)ry_[k + 1], (long long)r_new);
      fflush(stdout);
    }
    ry_[k + 1] = r_new;
    // Phi_k+1 = y_k^T Phi_k A_k x_k (dmrg_mv.h:200-209)
    if (F1) {
      GemmDesc g;
      g.M = (int)r_new; g.N = (int)c; g.K = (int)m1;
      g.a = rmT(y_[k], r_new); g.b = rm(F1, c); g.c = orm(Phis_[k + 1], c);
      gemm_f64(g, ws_, st_);
    } else {
      phi_fwd_A_rect(Phis_[k], y_[k], Apm_[k], x_[k], Phis_[k + 1], (int)ry_[k], (int)P.rA[k], (int)P.rx[k],
                     (int)Mk, (int)P.N[k], (int)P.rA[k + 1], (int)r_new, (int)P.rx[k + 1], ws_, st_);
    }
    ws_.release(mk);
  }
};

// Upper bound of the step temporaries at supercore k plus the persistent buffers.
size_t dmrg_workspace_bound(const DmrgPlan& P, const DmrgOptions& O) {
  Arena meas;
  {
    DmrgSweep s(P, O, meas, nullptr);
    std::vector<const double*> dummy(P.d, nullptr);
    s.allocate(dummy.data(), dummy.data(), nullptr, nullptr);
  }
  const double kick = (double)O.kickrank;
  double step = 0;
  for (int k = 0; k + 1 < P.d; ++k) {
    const double m1 = (double)(P.Rb[k] * P.M[k]), n2 = (double)(P.M[k + 1] * P.Rb[k + 2]);
    const double rm_ = (double)P.Rb[k + 1];
    const double c = std::max<double>((double)(P.rA[k + 1] * P.rx[k + 1]), rm_);
    const double kd = std::min(m1, n2), cs = std::min<double>(c, kMaxSmall);
    const double ncol = std::min<double>(kd, kMaxSmall) + kick;
    double t = 0;
    t += 2 * m1 * n2;                                        // W_prev, W
    t += (m1 + n2) * c;                                      // F1, F2
    t += (double)(P.Rb[k] * P.rA[k] * P.N[k] * P.rx[k + 1]);  // forward half stage 1
    t += (double)(P.rx[k + 1] * P.N[k + 1] * P.rA[k + 2] * P.Rb[k + 2]);  // right half stage 1
    t += 2 * (m1 + n2) * cs + 8 * cs * cs;                   // factored SVD
    t += std::max(m1, n2) * std::min<double>(kd, kMaxSmall) * 2 + (m1 + n2) * std::min<double>(kd, kMaxSmall);  // direct SVD
    t += n2 * std::min<double>(kd, kMaxSmall) + 3 * m1 * ncol + ncol * ncol;  // W2s, kick
    t += 4.0 * kMaxSmall * kMaxSmall + 48.0 * (kMaxSmall + 2) * (kMaxSmall / 2 + 2) * 2;  // QR / Jacobi internals
    step = std::max(step, t);
  }
  for (int k = 1; k < P.d; ++k) {  // right-to-left step
    const double m = (double)(P.M[k] * P.Rb[k + 1]), r1 = (double)P.Rb[k];
    double t = 2 * m * r1 + r1 * r1;
    t += (double)(P.rx[k] * P.N[k] * P.rA[k + 1] * P.Rb[k + 1]) + (double)(P.rx[k] * P.rA[k] * P.M[k] * P.Rb[k + 1]);
    t += 4.0 * kMaxSmall * kMaxSmall;
    step = std::max(step, t);
  }
  return meas.peak() + (size_t)(step * 8.0 * 1.25) + (64ull << 20);
}

size_t dmrg_output_elems(const DmrgPlan& P) {
  size_t out = 0;
  for (int k = 0; k < P.d; ++k) out += (size_t)(P.Rb[k] * P.M[k] * P.Rb[k + 1]);
  return out;
}

}  // namespace
}  // namespace ttamen

using namespace ttamen;

extern "C" {

size_t ttamen_dmrg_mv_workspace_size(int d, const int64_t* M, const int64_t* N, const int64_t* rA,
                                     const int64_t* rx, const int64_t* ry0, const int64_t* rmax,
                                     int64_t kickrank) {
  size_t out = 0;
  int rc = abi_call([&] {
    DmrgPlan P = make_dmrg_plan(d, M, N, rA, rx, ry0, rmax, kickrank);
    DmrgOptions O;
    O.kickrank = kickrank;
    out = dmrg_workspace_bound(P, O);
  });
  return rc == 0 ? out : 0;
}

size_t ttamen_dmrg_mv_output_elems(int d, const int64_t* M, const int64_t* ry0, const int64_t* rmax,
                                   int64_t kickrank) {
  size_t out = 0;
  int rc = abi_call([&] {
    TT_CHECK(d >= 2 && M, ERR_INVALID, "null argument");
    std::vector<int64_t> ones(d + 1, 1);
    DmrgPlan P = make_dmrg_plan(d, M, M, ones.data(), ones.data(), ry0, rmax, kickrank);
    out = dmrg_output_elems(P);
  });
  return rc == 0 ? out : 0;
}

int ttamen_dmrg_mv(int d, const int64_t* M, const int64_t* N, const int64_t* rA, const int64_t* rx,
                   const double* const* A_cores, const double* const* x_cores,
                   const double* const* y0_cores, const int64_t* ry0, int64_t nswp, double eps,
                   const int64_t* rmax, int64_t kickrank, uint64_t seed, int verbose, void* workspace,
                   size_t workspace_bytes, double* y_out, size_t y_out_elems, int64_t* ry_out,
                   int64_t* offsets_out, ttamen_stats* stats, void* stream) {
  return abi_call([&] {
    auto t0 = std::chrono::high_resolution_clock::now();
    TT_CHECK(A_cores && x_cores && workspace && y_out && ry_out && offsets_out, ERR_INVALID, "null argument");
    TT_CHECK(nswp >= 0 && eps >= 0, ERR_INVALID, "invalid solver parameters");
    TT_CHECK(!y0_cores || ry0, ERR_INVALID, "initial-guess ranks are required with its cores");
    DmrgPlan P = make_dmrg_plan(d, M, N, rA, rx, y0_cores ? ry0 : nullptr, rmax, kickrank);
    DmrgOptions O;
    O.nswp = nswp; O.eps = eps; O.kickrank = kickrank; O.seed = seed; O.verbose = verbose;
    linalg_init();
    hipStream_t st = (hipStream_t)stream;
    Arena ws(workspace, workspace_bytes);
    DmrgSweep s(P, O, ws, st);
    s.allocate(A_cores, x_cores, y0_cores, ry0);
    s.run();
    s.write_out(y_out, y_out_elems, ry_out, offsets_out);
    if (stats) {
      *stats = s.stats_;
      stats->elapsed_s = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    }
  });
}

}  // extern "C"
