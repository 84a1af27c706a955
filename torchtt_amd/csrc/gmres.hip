// Device-resident restarted GMRES(m) for the AMEn local systems.
//
// Reference: cpp/gmres.h:15-161 (gmres_single / gmres / Givens rotation) and its
// caller cpp/amen_solve.h:337-358. The reference keeps the Hessenberg matrix,
// the Givens rotations and every dot product on the host (one device->host sync
// per dot). Here the whole Arnoldi process stays on the device:
//   * orthogonalisation is classical Gram-Schmidt with selective re-orthogonalisation
//     (a second pass only when the first cancels more than 1 - 1/sqrt(2) of ||w||);
//     each pass is one dot kernel and one update kernel streaming the Krylov basis,
//     with deterministic block partials reduced by the last block;
//   * the Givens update of the Hessenberg column, the residual estimate |beta_k+1|
//     and the convergence test run in one thread of the normalisation kernel;
//   * a device flag `done` makes every kernel of later iterations (including the
//     local-operator GEMMs) return immediately, so the host can enqueue iteration
//     k+1 before it has read the flag of iteration k (one sync per iteration,
//     overlapped with the next iteration's work).
// Vectors are in the working precision T (double, or float for the mixed-precision
// inner solve); dot products and the Hessenberg are accumulated in double.
#include "ttamen_internal.h"
#include "gmres.h"
#include <mutex>
#include <vector>
#include "cohort.h"
#include "kernels_gmres.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace ttamen {

// ---------------------------------------------------------------- host driver
GmresDev gmres_alloc(Arena& ws, int m, int64_t N, int esize) {
  GmresDev g;
  g.m = m;
  g.Q = ws.alloc<char>((size_t)(m + 1) * N * esize);
  g.w = ws.alloc<char>((size_t)N * esize);
  g.H = ws.alloc<double>((size_t)(m + 1) * m);
  g.beta = ws.alloc<double>(m + 1);
  g.cs = ws.alloc<double>(m);
  g.sn = ws.alloc<double>(m);
  g.y = ws.alloc<double>(m);
  g.dstate = ws.alloc<double>(16);
  g.istate = ws.alloc<int>(32);  // 16 ints of state, then xtra (4 doubles at byte 64)
  g.xtra = reinterpret_cast<double*>(g.istate + 16);
  g.partial = ws.alloc<double>((size_t)kGmresMaxBlocks * (2 * kGmresMaxCols + 4));
  return g;
}

// Timing of the orthogonalisation (outside cohorts, while enabled): HIP events around the three
// launches of each Arnoldi step (CGS dots, update with the second pass's dots, close), and their
// algorithmic bytes (2 k + 7) N w for step k: the dots read k+1 basis vectors and w, the update
// reads them and w and writes w, the close reads w and writes the new basis vector. Steps enqueued
// after convergence (no-ops) are dropped once the cycle's executed step count is read back.
namespace {
struct OrthoProfile {
  std::mutex mu;
  bool on = false;
  double steps = 0, ms = 0, bytes = 0;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (pool.empty()) {
      hipEvent_t e;
      TT_HIP(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
} g_oprof;
}  // namespace

void ortho_profile_enable(bool on) {
  std::lock_guard<std::mutex> lk(g_oprof.mu);
  g_oprof.on = on;
  g_oprof.steps = g_oprof.ms = g_oprof.bytes = 0;
}

void ortho_profile_read(double* out3) {
  std::lock_guard<std::mutex> lk(g_oprof.mu);
  out3[0] = g_oprof.steps;
  out3[1] = g_oprof.ms;
  out3[2] = g_oprof.bytes;
}

template <typename T>
GmresResult gmres_run(GmresDev g, int64_t N, T* x, const T* rhs, const double* norm_rhs,
                      const double* dnorm, double real_tol, int resets, const GmresOp<T>& op,
                      RedScratch rs, HostSync& hs, hipStream_t st) {
  TT_CHECK(g.m >= 1 && g.m + 1 <= kGmresMaxCols, ERR_UNSUPPORTED,
           "gmres: local_iterations must be in [1, " + std::to_string(kGmresMaxCols - 1) + "]");
  T* Q = (T*)g.Q;
  T* w = (T*)g.w;
  const int nb = gm_blocks(N);
  const int* done = g.istate + GM_DONE;
  double* P1 = g.partial;                                        // (kc + 1) rows
  double* P2 = P1 + (size_t)kGmresMaxBlocks * (kGmresMaxCols + 1);  // kc rows
  double* Pn1 = P2 + (size_t)kGmresMaxBlocks * kGmresMaxCols;
  GmresResult res;
  int state[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool have_state = false;

  launch<KGmresSetup>(dim3(1), dim3(1), 0, st, {g, norm_rhs, dnorm, real_tol});
  TT_LAUNCH_CHECK();
  const bool cohort = current_recorder() != nullptr;
  const int poll = std::max(1, cohort_params().gmres_poll);
  bool prof;
  {
    std::lock_guard<std::mutex> lk(g_oprof.mu);
    prof = g_oprof.on && !cohort;
  }
  struct Tag { hipEvent_t e0, e1; int k; bool fused; };
  std::vector<Tag> tags;
  for (int cyc = 0; cyc < resets; ++cyc) {
    // r = rhs - Op(x); first cycle starts from x = 0 so r = rhs.
    const T* r = rhs;
    if (cyc > 0) {
      launch<KGmresRearm>(dim3(1), dim3(1), 0, st, {g});
      TT_LAUNCH_CHECK();
      op.apply(x, w, nullptr);
      op.axpby(1.0, rhs, -1.0, w, N);  // w = rhs - Op(x)
      op.norm(w, N, g.dstate + GM_RNORM);
      r = w;
    }
    launch<KGmresBegin<T>>(dim3(nb), dim3(kGB), 0, st, {r, Q, N, g});
    TT_LAUNCH_CHECK();
    int k = 0;
    bool stop = false;
    for (k = 0; k < g.m && !stop; ++k) {
      T* qk = Q + (int64_t)k * N;
      T* qn = Q + (int64_t)(k + 1) * N;
      int nb1 = 0;  // > 0: the operator wrote the first pass's dot shares itself
      if (op.apply_step) nb1 = op.apply_step(qk, w, done, cyc, k, Q, k + 1, P1);
      else if (op.apply_dots) nb1 = op.apply_dots(qk, w, done, Q, k + 1, P1);
      else op.apply(qk, w, done);
      double* Hcol = g.H + (int64_t)k * (g.m + 1);
      const double* Pw = P1 + (size_t)(k + 1) * kGmresMaxBlocks;  // row kc of P1: w.w
      // pass 1: project out the basis
      Tag tag{nullptr, nullptr, k, nb1 > 0};
      if (prof) {
        std::lock_guard<std::mutex> lk(g_oprof.mu);
        tag.e0 = g_oprof.get();
        tag.e1 = g_oprof.get();
        TT_HIP(hipEventRecord(tag.e0, st));
      }
      if (nb1 == 0) launch<KCgsDots<T>>(dim3(nb), dim3(kGB), 0, st, {Q, N, k + 1, w, P1, done});
      launch<KCgsUpdate<T>>(dim3(nb), dim3(kGB), 0, st, {Q, N, k + 1, w, P1, Hcol, Pn1, P2, done, nb1});
      // the close also stores the done flag into the host mailbox slot of this step, so the stream carries
      // no copy of it (hs.mark records the event that hs.wait waits for)
      launch<KGmresClose<T>>(dim3(nb), dim3(kGB), 0, st,
                             {Q, N, k + 1, w, qn, Pw, Pn1, P2, Hcol, g, k, done, g.istate + GM_REORTHS,
                              Pn1 + kGmresMaxBlocks, nb1, hs.slot(k)});
      TT_LAUNCH_CHECK();
      if (prof) {
        TT_HIP(hipEventRecord(tag.e1, st));
        tags.push_back(tag);
      }
      hs.mark(k, st);
      if (cohort) {
        // Inside a cohort a host read suspends the member until its queue has run: poll
        // every gmres_poll iterations (iterations enqueued after convergence are no-ops).
        if ((k + 1) % poll == 0 && hs.wait(k, st)) stop = true;
      } else if (k >= 1 && hs.wait(k - 1, st)) {
        // Read the flag of the previous iteration while this one runs.
        stop = true;
      }
    }
    // Flag of the last enqueued iteration.
    if (!cohort) (void)hs.wait(k - 1, st);
    launch<KGmresY>(dim3(1), dim3(1), 0, st, {g});
    launch<KGmresXupd<T>>(dim3(nb), dim3(kGB), 0, st, {Q, N, x, g});
    TT_LAUNCH_CHECK();
    if (op.post) op.post(g.xtra);
    // Did this cycle converge? (the whole state in one read: flag, columns, iterations, ..., and
    // the caller's xtra values)
    read_to_host(hs.host_int(), g.istate, op.post ? 64 + sizeof(double) * 4 : sizeof(int) * 8, st);
    for (int i = 0; i < 8; ++i) state[i] = hs.host_int()[i];
    if (op.post) std::memcpy(res.xtra, hs.host_int() + 16, sizeof(double) * 4);
    have_state = true;
    res.cycle_steps.push_back(state[GM_KCOLS]);
    if (prof) {  // the stream has reached the state read: every event of this cycle is complete
      std::lock_guard<std::mutex> lk(g_oprof.mu);
      for (const Tag& t : tags) {
        if (t.k < state[GM_KCOLS]) {
          float ms = 0;
          TT_HIP(hipEventElapsedTime(&ms, t.e0, t.e1));
          g_oprof.ms += ms;
          g_oprof.steps += 1;
          // the dots (k + 2 vectors) run inside the operator's kernel when fused
          g_oprof.bytes += (t.fused ? (double)t.k + 5.0 : 2.0 * t.k + 7.0) * (double)N * sizeof(T);
        }
        g_oprof.pool.push_back(t.e0);
        g_oprof.pool.push_back(t.e1);
      }
      tags.clear();
    }
    res.cycles = cyc + 1;
    if (state[GM_FLAG]) break;
  }
  if (!have_state) {
    read_to_host(hs.host_int(), g.istate, sizeof(int) * 8, st);
    for (int i = 0; i < 8; ++i) state[i] = hs.host_int()[i];
  }
  res.flag = state[GM_FLAG];
  res.iters = state[GM_ITERS];
  res.reorths = state[GM_REORTHS];
  return res;
}

template GmresResult gmres_run<double>(GmresDev, int64_t, double*, const double*, const double*,
                                       const double*, double, int, const GmresOp<double>&,
                                       RedScratch, HostSync&, hipStream_t);
template GmresResult gmres_run<float>(GmresDev, int64_t, float*, const float*, const double*,
                                      const double*, double, int, const GmresOp<float>&,
                                      RedScratch, HostSync&, hipStream_t);

}  // namespace ttamen
