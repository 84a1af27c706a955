// TT rounding on the GPU (the reference's native round_this, cpp/compression.h:3-41, with
// lr_orthogonal of cpp/ortho.h:77-101 and rank_chop of cpp/ortho.h:16-44):
//   1. left-to-right orthogonalisation: core_i = Q (QR of the (r_i n_i) x r_{i+1} unfolding),
//      R merged into core_{i+1};
//   2. right-to-left truncation: SVD of the r_i x (n_i r_{i+1}) unfolding, rank from the
//      singular values (relative tolerance eps / sqrt(d-1), cap rmax), U S merged into core_{i-1}.
// The QR, SVD and GEMMs are this library's kernels; the singular values of each step are read
// by the host to choose the rank (one synchronisation per core).
#include "ttamen_internal.h"
#include "kernels_blas1.h"
#include "../../include/ttamen.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace ttamen {

void linalg_init();

// Same rule as the reference's rank_chop (cpp/ortho.h:16-44), including its results for a
// zero vector (1) and for eps <= 0 (n - 1).
int rank_chop(const double* s, int n, double eps) {
  double nrm = 0.0;
  for (int i = 0; i < n; ++i) nrm += s[i] * s[i];
  int r = n - 1;
  if (std::sqrt(nrm) == 0.0) return 1;
  if (eps <= 0.0) return r;
  while (r > 0) {
    double sum = 0.0;
    for (int k = r; k < n; ++k) sum += s[k] * s[k];
    if (sum >= eps * eps) break;
    --r;
  }
  r += 1;
  return r > 0 ? r : 1;
}

// Rounds in place in slots c[k] (capacity: the input core sizes); R is updated.
static void tt_round_impl(int d, const int64_t* N, std::vector<int64_t>& R, std::vector<double*>& c,
                          double epsilon, int64_t rmax, Arena& ws, double* host, hipStream_t st) {
  if (d <= 1) return;
  const bool live = !ws.measuring();  // a measuring arena only sizes the workspace
  // 1. left-to-right QR
  for (int i = 0; i < d - 1; ++i) {
    size_t mk = ws.mark();
    const int64_t m = R[i] * N[i], n = R[i + 1];
    const int64_t kq = std::min(m, n);
    double* Q = ws.alloc<double>(m * kq);
    double* Rm = ws.alloc<double>(kq * n);
    qr(c[i], n, 1, m, n, Q, Rm, ws, st);
    if (live) copy(Q, c[i], m * kq, st);
    // core_{i+1} <- Rm (kq x n) * core_{i+1} (n x N[i+1] R[i+2])
    const int64_t cols = N[i + 1] * R[i + 2];
    double* t = ws.alloc<double>(kq * cols);
    GemmDesc g;
    g.M = (int)kq; g.N = (int)cols; g.K = (int)n;
    g.a = rm(Rm, n); g.b = rm(c[i + 1], cols); g.c = orm(t, cols);
    gemm_f64(g, ws, st);
    if (live) copy(t, c[i + 1], kq * cols, st);
    R[i + 1] = kq;
    ws.release(mk);
  }
  // 2. right-to-left truncated SVD
  const double eps = epsilon / std::sqrt((double)d - 1.0);
  for (int i = d - 1; i > 0; --i) {
    size_t mk = ws.mark();
    const int64_t m = R[i], n = N[i] * R[i + 1];
    const int64_t k = std::min(m, n);
    double* U = ws.alloc<double>(m * k);
    double* s = ws.alloc<double>(k);
    double* V = ws.alloc<double>(n * k);
    svd(c[i], n, 1, m, n, U, s, V, ws, st);
    int64_t rnew = k;
    if (live) {
      read_to_host(host, s, sizeof(double) * k, st);
      double nrm = 0.0;
      for (int64_t j = 0; j < k; ++j) nrm += host[j] * host[j];
      const int rc = rank_chop(host, (int)k, eps * std::sqrt(nrm));
      rnew = std::min<int64_t>(rmax, rc);
    }
    // core_{i-1} <- core_{i-1} ((R[i-1] N[i-1]) x m) * U[:, :rnew] diag(s[:rnew])
    const int64_t rows = R[i - 1] * N[i - 1];
    double* us = ws.alloc<double>(m * rnew);
    if (live) {
      const int64_t tot = m * rnew;
      const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((tot + 1023) / 1024, 2048));
      launch<KScaleCols>(dim3(grid), dim3(256), 0, st, {U, k, s, us, rnew, m, (int)rnew});
      TT_LAUNCH_CHECK();
    }
    double* t = ws.alloc<double>(rows * rnew);
    GemmDesc h;
    h.M = (int)rows; h.N = (int)rnew; h.K = (int)m;
    h.a = rm(c[i - 1], m); h.b = rm(us, rnew); h.c = orm(t, rnew);
    gemm_f64(h, ws, st);
    if (live) {
      copy(t, c[i - 1], rows * rnew, st);
      // core_i <- V[:, :rnew]^T  (rnew x n)
      copy2d(V, 1, k, c[i], n, rnew, n, st);
    }
    R[i] = rnew;
    ws.release(mk);
  }
}

static size_t round_ws_bound(int d, const int64_t* N, const int64_t* R) {
  Arena meas;
  std::vector<double*> c(d);
  std::vector<int64_t> Rv(R, R + d + 1);
  for (int k = 0; k < d; ++k) c[k] = meas.alloc<double>(R[k] * N[k] * R[k + 1]);
  tt_round_impl(d, N, Rv, c, 1e-12, 1 << 30, meas, nullptr, nullptr);
  return meas.peak() + (16ull << 20);
}

}  // namespace ttamen

using namespace ttamen;

extern "C" {

size_t ttamen_tt_round_workspace_size(int d, const int64_t* N, const int64_t* R) {
  try {
    if (d < 1 || !N || !R) return 0;
    for (int k = 0; k < d; ++k)
      if (N[k] < 1 || R[k] < 1 || R[k + 1] < 1) return 0;
    return round_ws_bound(d, N, R);
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return 0;
  }
}

int ttamen_tt_round(int d, const int64_t* N, const int64_t* R, const double* const* cores,
                    double eps, int64_t rmax, void* workspace, size_t workspace_bytes, double* out,
                    size_t out_elems, int64_t* r_out, int64_t* offsets_out, void* stream) {
  try {
    TT_CHECK(d >= 1 && N && R && cores && workspace && out && r_out && offsets_out, ERR_INVALID,
             "null argument");
    TT_CHECK(R[0] == 1 && R[d] == 1, ERR_INVALID, "boundary ranks must be 1");
    TT_CHECK(rmax >= 1, ERR_INVALID, "rmax must be positive");
    for (int k = 0; k < d; ++k) TT_CHECK(N[k] >= 1 && R[k] >= 1, ERR_INVALID, "sizes must be positive");
    hipStream_t st = (hipStream_t)stream;
    linalg_init();
    Arena ws(workspace, workspace_bytes);
    std::vector<double*> c(d);
    size_t total = 0;
    for (int k = 0; k < d; ++k) {
      const int64_t n = R[k] * N[k] * R[k + 1];
      c[k] = ws.alloc<double>(n);
      copy(cores[k], c[k], n, st);
      total += (size_t)n;
    }
    TT_CHECK(out_elems >= total, ERR_WORKSPACE, "out too small");
    std::vector<int64_t> Rv(R, R + d + 1);
    static thread_local double* host = nullptr;
    if (!host) TT_HIP(hipHostMalloc((void**)&host, sizeof(double) * 4096, hipHostMallocDefault));
    // one k of singular values is read per core: at most min(r_i, n_i r_{i+1}) <= 140 here
    tt_round_impl(d, N, Rv, c, eps, rmax, ws, host, st);
    size_t off = 0;
    for (int k = 0; k < d; ++k) {
      const int64_t n = Rv[k] * N[k] * Rv[k + 1];
      copy(c[k], out + off, n, st);
      offsets_out[k] = (int64_t)off;
      off += (size_t)n;
    }
    for (int k = 0; k <= d; ++k) r_out[k] = Rv[k];
    TT_HIP(hipStreamSynchronize(st));
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ERR_INVALID;
  }
}

// Full tensor of a TT vector (reference cpp/full.h:3-12, exported as torchttcpp.tt_full): the cores
// are multiplied left to right, t_k = t_k-1 (n_0..n_k-1 x r_k) times core_k (r_k x n_k r_k+1), one GEMM
// per core, ping-ponging between two workspace buffers; the last product is written to out.
static int64_t full_stage_elems(int d, const int64_t* N, const int64_t* R) {
  int64_t worst = 1, rows = 1;
  for (int k = 0; k + 1 < d; ++k) {
    rows *= N[k];
    worst = std::max<int64_t>(worst, rows * R[k + 1]);
  }
  return worst;
}

size_t ttamen_tt_full_workspace_size(int d, const int64_t* N, const int64_t* R) {
  int64_t stage = 0;
  int rc = abi_call([&] {
    TT_CHECK(d >= 1 && N && R, ERR_INVALID, "null argument");
    stage = full_stage_elems(d, N, R);
  });
  return rc == 0 ? (size_t)(2 * stage + 64) * sizeof(double) + (64u << 20) : 0;
}

int ttamen_tt_full(int d, const int64_t* N, const int64_t* R, const double* const* cores, void* workspace,
                   size_t workspace_bytes, double* out, size_t out_elems, void* stream) {
  return abi_call([&] {
    TT_CHECK(d >= 1 && N && R && cores && workspace && out, ERR_INVALID, "null argument");
    TT_CHECK(R[0] == 1 && R[d] == 1, ERR_INVALID, "boundary ranks must be 1");
    double total = 1.0;
    for (int k = 0; k < d; ++k) {
      TT_CHECK(N[k] >= 1 && R[k + 1] >= 1 && cores[k], ERR_INVALID, "bad cores");
      total *= (double)N[k];
    }
    TT_CHECK(total <= (double)out_elems, ERR_WORKSPACE, "output buffer too small");
    TT_CHECK(total * (double)std::max<int64_t>(1, *std::max_element(R, R + d + 1)) < 2147483647.0,
             ERR_UNSUPPORTED, "full tensor too large");
    hipStream_t st = (hipStream_t)stream;
    if (d == 1) {
      copy(cores[0], out, N[0], st);
      TT_HIP(hipStreamSynchronize(st));
      return;
    }
    Arena ws(workspace, workspace_bytes);
    const int64_t stage = full_stage_elems(d, N, R);
    double* buf[2] = {ws.alloc<double>(stage), ws.alloc<double>(stage)};
    const double* t = cores[0];  // n_0 x r_1
    int64_t rows = N[0];
    for (int k = 1; k < d; ++k) {
      double* dst = (k == d - 1) ? out : buf[k & 1];
      GemmDesc g;
      g.M = (int)rows; g.N = (int)(N[k] * R[k + 1]); g.K = (int)R[k];
      g.a = rm(t, R[k]); g.b = rm(cores[k], N[k] * R[k + 1]); g.c = orm(dst, N[k] * R[k + 1]);
      gemm_f64(g, ws, st);
      t = dst;
      rows *= N[k];
    }
    TT_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
