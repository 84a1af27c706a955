// Tensor-train contractions of the AMEn sweep, each lowered to one to three
// strided GEMMs on the MFMA kernel (no transposes in memory).
//
// Index conventions follow cpp/amen_solve.h of the reference:
//   vector core        x[r, n, R]           (row-major)
//   operator core      A[s, m, n, S]        (row m, column n)
//   A-interface        Phi[l, s, r]         (l: test/output side, r: trial/x side)
//   rhs interface      Phib[b, r]
#include "ttamen_internal.h"
#include "kernels_matvec.h"
#include "kernels_mvtail.h"

#include <atomic>
#include <mutex>

namespace ttamen {

// ---------------------------------------------------------------- fused local operator
// Launch plan of KLocalMv (kernels_matvec.h) for one shape, or ok = false when the shape is
// outside what the kernel takes (then the three-GEMM form below runs):
//   BL rows l per workgroup: 8 / s (at least 1, at most l1), so that stage 1 reuses each loaded x
//       element for about 8 products; NLS = the PhiL slice width (4, 8 or 16) >= BL s
//   BR columns R per workgroup: 16 / BL (the 16 stage-2 columns), at most 32 / S (stage 3 keeps
//       S BR / 4 <= 8 fragments in registers) and at most r2; P = ceil(r2 / BR) <= 18 partial sums
//       (ranks up to 144 with 8 columns R per workgroup)
//   LDS: PhiL slice + T1 (s n rows) + T2 (n S rows), within the 160 KB of a CU
struct MvPlan {
  bool ok = false;
  int BL = 1, BR = 1, nlb = 1, P = 1, NLS = 4, KS = 0, MS = 0, KST = 0, RB = 0;
  size_t lds = 0;
};
constexpr size_t kMvMaxLds = 160 * 1024 - 1024;

static MvPlan mv_plan(int l1, int s, int r1, int n, int S, int l2, int r2, size_t elem) {
  MvPlan p;
  if (s < 1 || s > 16 || S < 1 || S > 32 || n < 1 || l1 < 1 || l2 < 1 || r1 < 1 || r2 < 1) return p;
  p.BL = std::max(1, std::min(s <= 8 ? 8 / s : 1, l1));
  p.NLS = p.BL * s <= 4 ? 4 : (p.BL * s <= 8 ? 8 : 16);
  p.BR = std::max(1, std::min({kMvCols / p.BL, 32 / S, r2}));
  p.nlb = (l1 + p.BL - 1) / p.BL;
  p.P = (r2 + p.BR - 1) / p.BR;
  if (p.P > 18) return p;
  p.KS = s * n;
  p.MS = n * S;
  p.KST = (p.KS + 3) / 4;
  p.RB = (p.MS + 15) / 16;
  p.lds = elem * ((size_t)r1 * p.NLS + (size_t)(4 * p.KST + 16 * p.RB) * kMvPitch);
  if (p.lds > kMvMaxLds) return p;
  p.ok = true;
  return p;
}

size_t frag_a_elems(int MS, int KS) { return (size_t)((MS + 15) / 16) * ((KS + 3) / 4) * 64; }

template <typename T>
static void pack_frag_a_t(const T* Apm, T* Apk, int MS, int KS, hipStream_t st) {
  typename KPackFragA<T>::Args a{Apm, Apk, MS, KS, (KS + 3) / 4, (MS + 15) / 16};
  launch<KPackFragA<T>>(dim3(grid_1d((int64_t)frag_a_elems(MS, KS))), dim3(256), 0, st, a);
  TT_LAUNCH_CHECK();
}
void pack_frag_a(const double* Apm, double* Apk, int MS, int KS, hipStream_t st) {
  pack_frag_a_t<double>(Apm, Apk, MS, KS, st);
}
void pack_frag_a_f32(const float* Apm, float* Apk, int MS, int KS, hipStream_t st) {
  pack_frag_a_t<float>(Apm, Apk, MS, KS, st);
}

template <typename T, int NLS>
static void launch_mv(const MvArgs<T>& a, const MvPlan& pl, hipStream_t st) {
  static std::once_flag once;
  std::call_once(once, [] {
    if (allow_lds<KLocalMv<T, NLS>>((int)kMvMaxLds) != hipSuccess) (void)hipGetLastError();
  });
  launch<KLocalMv<T, NLS>>(dim3(pl.nlb * pl.P, a.nvec), dim3(KLocalMv<T, NLS>::kThreads), pl.lds, st, a);
  TT_LAUNCH_CHECK();
}

// Local-operator form (ttamen_matvec_mode): 0 (default) stage 1 as a GEMM and stages 2-3 as one kernel
// (KMvTail, kernels_mvtail.h; FP64 and FP32, l2 <= 144, else the GEMMs); 1 the fused one-launch kernel
// where its plan allows it; 2 three GEMMs.
// Measured on MI355X (largest batch system, FP32, ranks ~30-60, n = 60): the one-launch kernel takes
// 38 us per application against 24 us for the three GEMMs, and the 256-system batch runs at 14.5
// against 37.7 solves/s. In-kernel timestamps show why: a workgroup's four stages each wait on one
// round trip to memory (3.6-11 us each at these shapes, 17-28 us per workgroup end to end), and the
// stages of one workgroup cannot overlap, while the GEMMs spread the same flops over many short tiles
// on every CU.
static std::atomic<int> g_matvec_mode{0};
int matvec_mode(int mode) {
  const int prev = g_matvec_mode.load();
  if (mode >= 0) g_matvec_mode.store(mode);
  return prev;
}

// Runs the fused kernel when it is selected, the plan allows it and the packed core is there
// (returns true), else returns false. With a measuring arena it allocates what the fused form needs and returns false,
// so the caller also measures the three-GEMM form (the workspace covers either).
template <typename T>
static bool fused_matvec(const T* PhiL, const T* Apk, const T* PhiR, int l1, int s, int r1, int n, int S,
                         int l2, int r2, const T* x, T* y, int nvec, unsigned* cnt, int cnt_cap, Arena& ws,
                         hipStream_t st, const int* skip) {
  if (!ws.measuring() && (!Apk || g_matvec_mode.load() != 1)) return false;
  const MvPlan pl = mv_plan(l1, s, r1, n, S, l2, r2, sizeof(T));
  if (!pl.ok) return false;
  const size_t mk = ws.mark();
  MvArgs<T> a;
  a.PhiL = PhiL; a.Apk = Apk; a.PhiR = PhiR; a.x = x; a.y = y; a.part = nullptr; a.cnt = nullptr; a.skip = skip;
  a.l1 = l1; a.s = s; a.r1 = r1; a.n = n; a.S = S; a.l2 = l2; a.r2 = r2; a.nvec = nvec;
  a.BL = pl.BL; a.BR = pl.BR; a.nlb = pl.nlb; a.P = pl.P;
  a.KS = pl.KS; a.MS = pl.MS; a.KST = pl.KST; a.RB = pl.RB; a.pad = 0;
  const int64_t need_cnt = (int64_t)nvec * pl.nlb;
  bool zero_cnt = false;
  if (pl.P > 1) {
    a.part = ws.alloc<T>((size_t)pl.P * nvec * l1 * n * l2);
    if (cnt && need_cnt <= cnt_cap) {
      a.cnt = cnt;
    } else {
      a.cnt = ws.alloc<unsigned>((size_t)need_cnt);
      zero_cnt = true;
    }
  }
  if (ws.measuring()) {
    ws.release(mk);
    return false;
  }
  if (zero_cnt) zero_words(a.cnt, need_cnt, st);
  switch (pl.NLS) {
    case 4: launch_mv<T, 4>(a, pl, st); break;
    case 8: launch_mv<T, 8>(a, pl, st); break;
    default: launch_mv<T, 16>(a, pl, st); break;
  }
  ws.release(mk);
  return true;
}

// Local operator y[v] = PhiL * A * PhiR * x[v]   (reference: matvecs.h:196-212,
// amen_solve.h:17-27, contraction 'lsr,smnS,LSR,rnR->lmL').
//   stage 1 (batch v,s):  t1[v][l][s][n][R] = sum_r PhiL[l,s,r] x[v][r][n][R]
//   stage 2 (batch v,l):  t2[v][l][m][S][R] = sum_{s,n} Apm[(m,S),(s,n)] t1[v][l][(s,n)][R]
//   stage 3 (one GEMM):   y[(v,l,m)][L]     = sum_{S,R} t2[(v,l,m)][(S,R)] PhiR[L,(S,R)]
// Stages 2 and 3 of the local operator as one KMvTail launch (t1[v][l][(s,n)][R] -> y); false when the
// shape is outside the kernel (l2 above 16 kTailMaxLB).
bool matvec_skips_zero_tiles(int l2, int K, bool bf16) {
  return !bf16 && g_matvec_mode.load() == 0 && (l2 + 15) / 16 <= kTailMaxLB && K <= kTileMaskMaxK;
}

template <typename T>
static void op_tile_masks_t(const T* Apm, int nm, int S, int K, uint64_t* out, hipStream_t st) {
  TT_CHECK(K >= 1 && K <= kTileMaskMaxK && nm >= 1 && S >= 1, ERR_INVALID, "op_tile_masks: shape");
  launch<KTileMask<T>>(dim3((nm + 15) / 16, S), dim3(64), 0, st, {Apm, out, nm, S, K});
  TT_LAUNCH_CHECK();
}
void op_tile_masks(const double* Apm, int nm, int S, int K, uint64_t* out, hipStream_t st) {
  op_tile_masks_t<double>(Apm, nm, S, K, out, st);
}
void op_tile_masks(const float* Apm, int nm, int S, int K, uint64_t* out, hipStream_t st) {
  op_tile_masks_t<float>(Apm, nm, S, K, out, st);
}

template <typename T>
static bool mv_tail(const T* t1, const T* Apm, const T* PhiR, T* y, int64_t nvl, int K, int r2, int nm, int S,
                    int l2, Arena& ws, hipStream_t st, const int* skip, MvDots* dots, const uint64_t* tmask) {
  const int LB = (l2 + 15) / 16;
  // a measuring arena returns false, so that the caller also sizes the GEMM form (either may run)
  if (LB > kTailMaxLB || nvl > INT32_MAX || (nm + 15) / 16 > 65535 || ws.measuring()) return false;
  typename KMvTail<T>::Args a;
  a.t1 = t1; a.Apm = Apm; a.PhiR = PhiR; a.y = y; a.skip = skip;
  a.tmask = K <= kTileMaskMaxK ? tmask : nullptr;
  a.K = K; a.r2 = r2; a.nm = nm; a.S = S; a.l2 = l2; a.LB = LB;
  a.Qd = nullptr; a.P = nullptr; a.N = 0; a.kc = 0;
  const int64_t nblocks = nvl * ((nm + 15) / 16);
  if (dots && dots->kc > 0 && dots->kc + 1 <= kGmresMaxCols && nblocks <= kGmresMaxBlocks &&
      dots->N == nvl * nm * l2) {
    a.Qd = (const T*)dots->Q; a.P = dots->P; a.N = dots->N; a.kc = dots->kc;
    dots->nb = (int)nblocks;
  }
  const int vw = (int)(16 / sizeof(T));
  a.apv = (K % vw == 0 && reinterpret_cast<uintptr_t>(Apm) % 16 == 0) ? 1 : 0;
  launch<KMvTail<T>>(dim3((unsigned)nvl, (nm + 15) / 16), dim3(KMvTail<T>::kThreads), 0, st, a);
  TT_LAUNCH_CHECK();
  return true;
}

template <typename T>
static void matvec_impl(const T* PhiL, const T* Apm, const T* PhiR, int l1, int s, int r1, int n,
                        int nm, int S, int l2, int r2, const T* x, T* y, int nvec, Arena& ws,
                        hipStream_t st, const int* skip,
                        void (*gemm)(const GemmDesc&, Arena&, hipStream_t), bool tail, MvDots* dots,
                        const uint64_t* tmask) {
  // n: mode of x (columns of the operator core), nm: mode of y (rows; nm == n when square)
  const int64_t nx = (int64_t)r1 * n * r2;
  const int64_t t1v = (int64_t)l1 * s * n * r2;  // per vector
  const int64_t t2v = (int64_t)l1 * nm * S * r2;
  size_t mk = ws.mark();
  T* t1 = ws.alloc<T>(t1v * nvec);
  T* t2 = ws.alloc<T>(t2v * nvec);
  {
    GemmDesc g;
    g.M = l1; g.N = n * r2; g.K = r1; g.B1 = nvec; g.B2 = s;
    g.a.p = PhiL; g.a.r = (int64_t)s * r1; g.a.c = 1; g.a.b1 = 0; g.a.b2 = r1;
    g.b.p = x; g.b.r = (int64_t)n * r2; g.b.c = 1; g.b.b1 = nx; g.b.b2 = 0;
    g.c.p = t1; g.c.r = (int64_t)s * n * r2; g.c.c = 1; g.c.b1 = t1v; g.c.b2 = (int64_t)n * r2;
    g.skip = skip;
    gemm(g, ws, st);
  }
  if (tail && g_matvec_mode.load() == 0 &&
      mv_tail<T>(t1, Apm, PhiR, y, (int64_t)nvec * l1, s * n, r2, nm, S, l2, ws, st, skip,
                 nvec == 1 ? dots : nullptr, tmask)) {
    ws.release(mk);
    return;
  }
  {
    GemmDesc g;
    g.M = nm * S; g.N = r2; g.K = s * n; g.B1 = nvec * l1; g.B2 = 1;
    g.a.p = Apm; g.a.r = (int64_t)s * n; g.a.c = 1; g.a.b1 = 0;
    g.b.p = t1; g.b.r = r2; g.b.c = 1; g.b.b1 = (int64_t)s * n * r2;
    g.c.p = t2; g.c.r = r2; g.c.c = 1; g.c.b1 = (int64_t)nm * S * r2;
    g.skip = skip;
    gemm(g, ws, st);
  }
  {
    GemmDesc g;
    g.M = nvec * l1 * nm; g.N = l2; g.K = S * r2;
    g.a.p = t2; g.a.r = (int64_t)S * r2; g.a.c = 1;
    g.b.p = PhiR; g.b.r = 1; g.b.c = (int64_t)S * r2;
    g.c.p = y; g.c.r = l2; g.c.c = 1;
    g.skip = skip;
    gemm(g, ws, st);
  }
  ws.release(mk);
}

void local_matvec(const LocalOp& op, const double* x, double* y, int nvec, Arena& ws,
                  hipStream_t st, const int* skip, MvDots* dots) {
  if (dots) dots->nb = 0;
  const bool square = op.n_out == 0 || op.n_out == op.n;
  if (square && fused_matvec<double>(op.PhiL, op.Apk, op.PhiR, op.l1, op.s, op.r1, op.n, op.S, op.l2, op.r2,
                                     x, y, nvec, op.cnt, op.cnt_cap, ws, st, skip))
    return;
  matvec_impl<double>(op.PhiL, op.Apm, op.PhiR, op.l1, op.s, op.r1, op.n, op.n_out ? op.n_out : op.n,
                      op.S, op.l2, op.r2, x, y, nvec, ws, st, skip, gemm_f64, true, dots, op.tmask);
}

void local_matvec_f32(const LocalOpF32& op, const float* x, float* y, int nvec, Arena& ws,
                      hipStream_t st, const int* skip, MvDots* dots) {
  if (dots) dots->nb = 0;
  if (!op.bf16 && fused_matvec<float>(op.PhiL, op.Apk, op.PhiR, op.l1, op.s, op.r1, op.n, op.S, op.l2, op.r2,
                                      x, y, nvec, op.cnt, op.cnt_cap, ws, st, skip))
    return;
  matvec_impl<float>(op.PhiL, op.Apm, op.PhiR, op.l1, op.s, op.r1, op.n, op.n, op.S, op.l2, op.r2,
                     x, y, nvec, ws, st, skip, op.bf16 ? gemm_bf16 : gemm_f32, !op.bf16, dots, op.tmask);
}

size_t local_matvec_ws(const LocalOp& op, int nvec) {
  Arena a;
  a.alloc<double>((size_t)op.l1 * op.s * op.n * op.r2 * nvec);
  a.alloc<double>((size_t)op.l1 * (op.n_out ? op.n_out : op.n) * op.S * op.r2 * nvec);
  size_t peak = a.peak();
  // the fused form: partial sums and counters (on top of the packed core, which callers add)
  const MvPlan pl = mv_plan(op.l1, op.s, op.r1, op.n, op.S, op.l2, op.r2, sizeof(double));
  if (pl.ok && pl.P > 1)
    peak = std::max(peak, (size_t)pl.P * nvec * op.l1 * op.n * op.l2 * sizeof(double) +
                              (size_t)nvec * pl.nlb * sizeof(unsigned) + 512);
  return peak + 64 * 1024 * 1024;  // generous split-K headroom
}

size_t local_matvec_fused_bytes(int l1, int s, int r1, int n, int S, int l2, int r2, int nvec, size_t elem) {
  const MvPlan pl = mv_plan(l1, s, r1, n, S, l2, r2, elem);
  if (!pl.ok || pl.P <= 1) return 0;
  return (size_t)pl.P * nvec * l1 * n * l2 * elem + (size_t)nvec * pl.nlb * sizeof(unsigned) + 512;
}

typedef void (*GemmFn)(const GemmDesc&, Arena&, hipStream_t);

// Phi forward (reference amen_solve.h:58-68, 'lsr,lML,sMNS,rNR->LSR'):
//   t2[l][(M,S)][R] = stages 1-2 of the local operator with PhiL = Phi, x = xR
//   Pn[L][(S,R)]    = sum_{(l,M)} xL[(l,M)][L] t2[(l,M)][(S,R)]
template <typename T>
static void phi_fwd_half_t(const T* Phi, const T* Apm, const T* xR, T* t2, int l, int s, int r, int nm, int nk,
                           int S, int R, Arena& ws, hipStream_t st, GemmFn gemm) {
  size_t mk = ws.mark();
  T* t1 = ws.alloc<T>((size_t)l * s * nk * R);
  {
    GemmDesc g;
    g.M = l; g.N = nk * R; g.K = r; g.B1 = 1; g.B2 = s;
    g.a.p = Phi; g.a.r = (int64_t)s * r; g.a.c = 1; g.a.b2 = r;
    g.b.p = xR; g.b.r = (int64_t)nk * R; g.b.c = 1;
    g.c.p = t1; g.c.r = (int64_t)s * nk * R; g.c.c = 1; g.c.b2 = (int64_t)nk * R;
    gemm(g, ws, st);
  }
  {
    GemmDesc g;
    g.M = nm * S; g.N = R; g.K = s * nk; g.B1 = l;
    g.a.p = Apm; g.a.r = (int64_t)s * nk; g.a.c = 1;
    g.b.p = t1; g.b.r = R; g.b.c = 1; g.b.b1 = (int64_t)s * nk * R;
    g.c.p = t2; g.c.r = R; g.c.c = 1; g.c.b1 = (int64_t)nm * S * R;
    gemm(g, ws, st);
  }
  ws.release(mk);
}

template <typename T>
static void phi_fwd_A_t(const T* Phi, const T* xL, const T* Apm, const T* xR, T* Pn, int l, int s, int r, int nm,
                        int nk, int S, int L, int R, Arena& ws, hipStream_t st, GemmFn gemm,
                        const double* inv_scale = nullptr) {
  size_t mk = ws.mark();
  T* t2 = ws.alloc<T>((size_t)l * nm * S * R);
  phi_fwd_half_t<T>(Phi, Apm, xR, t2, l, s, r, nm, nk, S, R, ws, st, gemm);
  {
    GemmDesc g;
    g.M = L; g.N = S * R; g.K = l * nm;
    g.a.p = xL; g.a.r = 1; g.a.c = L;
    g.b.p = t2; g.b.r = (int64_t)S * R; g.b.c = 1;
    g.c.p = Pn; g.c.r = (int64_t)S * R; g.c.c = 1;
    if (inv_scale) { g.scale_ptr = inv_scale; g.scale_mode = 2; }
    gemm(g, ws, st);
  }
  ws.release(mk);
}

// Phi backward (reference amen_solve.h:38-47, 'LSR,lML,sMNS,rNR->lsr'):
//   u1[(r,N)][S][L] = sum_R xR[(r,N)][R] Phi[L][S][R]            (batch S)
//   u2[r][(s,M)][L] = sum_{(N,S)} A[(s,M)][(N,S)] u1[r][(N,S)][L] (batch r)
//   Pp[l][s][r]     = sum_{(M,L)} xL[l][(M,L)] u2[r][s][(M,L)]    (batch s)
template <typename T>
static void phi_bck_A_t(const T* Phi, const T* xL, const T* A, const T* xR, T* Pp, int l, int s, int r, int nm,
                        int nk, int S, int L, int R, Arena& ws, hipStream_t st, GemmFn gemm,
                        const double* inv_scale = nullptr) {
  size_t mk = ws.mark();
  T* u1 = ws.alloc<T>((size_t)r * nk * S * L);
  T* u2 = ws.alloc<T>((size_t)r * s * nm * L);
  {
    GemmDesc g;
    g.M = r * nk; g.N = L; g.K = R; g.B2 = S;
    g.a.p = xR; g.a.r = R; g.a.c = 1;
    g.b.p = Phi; g.b.r = 1; g.b.c = (int64_t)S * R; g.b.b2 = R;
    g.c.p = u1; g.c.r = (int64_t)S * L; g.c.c = 1; g.c.b2 = L;
    gemm(g, ws, st);
  }
  {
    GemmDesc g;
    g.M = s * nm; g.N = L; g.K = nk * S; g.B1 = r;
    g.a.p = A; g.a.r = (int64_t)nk * S; g.a.c = 1;
    g.b.p = u1; g.b.r = L; g.b.c = 1; g.b.b1 = (int64_t)nk * S * L;
    g.c.p = u2; g.c.r = L; g.c.c = 1; g.c.b1 = (int64_t)s * nm * L;
    gemm(g, ws, st);
  }
  {
    GemmDesc g;
    g.M = l; g.N = r; g.K = nm * L; g.B2 = s;
    g.a.p = xL; g.a.r = (int64_t)nm * L; g.a.c = 1;
    g.b.p = u2; g.b.r = 1; g.b.c = (int64_t)s * nm * L; g.b.b2 = (int64_t)nm * L;
    g.c.p = Pp; g.c.r = (int64_t)s * r; g.c.c = 1; g.c.b2 = r;
    if (inv_scale) { g.scale_ptr = inv_scale; g.scale_mode = 2; }
    gemm(g, ws, st);
  }
  ws.release(mk);
}

void phi_fwd_half_rect(const double* Phi, const double* Apm, const double* xR, double* t2, int l, int s,
                       int r, int nm, int nk, int S, int R, Arena& ws, hipStream_t st) {
  phi_fwd_half_t<double>(Phi, Apm, xR, t2, l, s, r, nm, nk, S, R, ws, st, gemm_f64);
}
void phi_fwd_A_rect(const double* Phi, const double* xL, const double* Apm, const double* xR,
                    double* Pn, int l, int s, int r, int nm, int nk, int S, int L, int R,
                    Arena& ws, hipStream_t st, const double* inv_scale) {
  phi_fwd_A_t<double>(Phi, xL, Apm, xR, Pn, l, s, r, nm, nk, S, L, R, ws, st, gemm_f64, inv_scale);
}
void phi_bck_A_rect(const double* Phi, const double* xL, const double* A, const double* xR,
                    double* Pp, int l, int s, int r, int nm, int nk, int S, int L, int R,
                    Arena& ws, hipStream_t st, const double* inv_scale) {
  phi_bck_A_t<double>(Phi, xL, A, xR, Pp, l, s, r, nm, nk, S, L, R, ws, st, gemm_f64, inv_scale);
}
void phi_fwd_A(const double* Phi, const double* xL, const double* Apm, const double* xR,
               double* Pn, int l, int s, int r, int n, int S, int L, int R, Arena& ws,
               hipStream_t st, const double* inv_scale) {
  phi_fwd_A_rect(Phi, xL, Apm, xR, Pn, l, s, r, n, n, S, L, R, ws, st, inv_scale);
}
void phi_bck_A(const double* Phi, const double* xL, const double* A, const double* xR,
               double* Pp, int l, int s, int r, int n, int S, int L, int R, Arena& ws,
               hipStream_t st, const double* inv_scale) {
  phi_bck_A_rect(Phi, xL, A, xR, Pp, l, s, r, n, n, S, L, R, ws, st, inv_scale);
}
// FP32 operands and result, FP32 accumulation on the matrix cores.
void phi_fwd_A_f32(const float* Phi, const float* xL, const float* Apm, const float* xR, float* Pn, int l, int s,
                   int r, int n, int S, int L, int R, Arena& ws, hipStream_t st) {
  phi_fwd_A_t<float>(Phi, xL, Apm, xR, Pn, l, s, r, n, n, S, L, R, ws, st, gemm_f32);
}
void phi_bck_A_f32(const float* Phi, const float* xL, const float* A, const float* xR, float* Pp, int l, int s,
                   int r, int n, int S, int L, int R, Arena& ws, hipStream_t st) {
  phi_bck_A_t<float>(Phi, xL, A, xR, Pp, l, s, r, n, n, S, L, R, ws, st, gemm_f32);
}

// Phib forward (reference amen_solve.h:93-98, 'br,bnB,rnR->BR').
void phi_fwd_rhs(const double* Pb, const double* bc, const double* x, double* Pn, int b, int r,
                 int n, int B, int R, Arena& ws, hipStream_t st, const double* inv_scale) {
  size_t mk = ws.mark();
  double* u = ws.alloc<double>((size_t)r * n * B);
  {
    GemmDesc g;  // u[r][(n,B)] = Pb^T[r][b] bc[b][(n,B)]
    g.M = r; g.N = n * B; g.K = b;
    g.a.p = Pb; g.a.r = 1; g.a.c = r;
    g.b.p = bc; g.b.r = (int64_t)n * B; g.b.c = 1;
    g.c.p = u; g.c.r = (int64_t)n * B; g.c.c = 1;
    gemm_f64(g, ws, st);
  }
  {
    GemmDesc g;  // Pn[B][R] = sum_{(r,n)} u[(r,n)][B] x[(r,n)][R]
    g.M = B; g.N = R; g.K = r * n;
    g.a.p = u; g.a.r = 1; g.a.c = B;
    g.b.p = x; g.b.r = R; g.b.c = 1;
    g.c.p = Pn; g.c.r = R; g.c.c = 1;
    if (inv_scale) { g.scale_ptr = inv_scale; g.scale_mode = 2; }
    gemm_f64(g, ws, st);
  }
  ws.release(mk);
}

// Phib backward (reference amen_solve.h:78-83, 'BR,bnB,rnR->br').
void phi_bck_rhs(const double* Pb, const double* bc, const double* x, double* Pp, int b, int r,
                 int n, int B, int R, Arena& ws, hipStream_t st, const double* inv_scale) {
  size_t mk = ws.mark();
  double* u = ws.alloc<double>((size_t)b * n * R);
  {
    GemmDesc g;  // u[(b,n)][R] = bc[(b,n)][B] Pb[B][R]
    g.M = b * n; g.N = R; g.K = B;
    g.a.p = bc; g.a.r = B; g.a.c = 1;
    g.b.p = Pb; g.b.r = R; g.b.c = 1;
    g.c.p = u; g.c.r = R; g.c.c = 1;
    gemm_f64(g, ws, st);
  }
  {
    GemmDesc g;  // Pp[b][r] = sum_{(n,R)} u[b][(n,R)] x[r][(n,R)]
    g.M = b; g.N = r; g.K = n * R;
    g.a.p = u; g.a.r = (int64_t)n * R; g.a.c = 1;
    g.b.p = x; g.b.r = 1; g.b.c = (int64_t)n * R;
    g.c.p = Pp; g.c.r = r; g.c.c = 1;
    if (inv_scale) { g.scale_ptr = inv_scale; g.scale_mode = 2; }
    gemm_f64(g, ws, st);
  }
  ws.release(mk);
}

// Local right-hand side (reference amen_solve.h:303-304, 'br,bnB,BR->rnR'), scaled
// by a device scalar (the sweep's norm correction nrmsc).
void assemble_rhs(const double* PL, const double* bc, const double* PR, double* out, int b, int r,
                  int n, int B, int R, const double* scale_ptr, Arena& ws, hipStream_t st) {
  size_t mk = ws.mark();
  double* u = ws.alloc<double>((size_t)r * n * B);
  {
    GemmDesc g;  // u[r][(n,B)] = PL^T[r][b] bc[b][(n,B)]
    g.M = r; g.N = n * B; g.K = b;
    g.a.p = PL; g.a.r = 1; g.a.c = r;
    g.b.p = bc; g.b.r = (int64_t)n * B; g.b.c = 1;
    g.c.p = u; g.c.r = (int64_t)n * B; g.c.c = 1;
    gemm_f64(g, ws, st);
  }
  {
    GemmDesc g;  // out[(r,n)][R] = u[(r,n)][B] PR[B][R]
    g.M = r * n; g.N = R; g.K = B;
    g.a.p = u; g.a.r = B; g.a.c = 1;
    g.b.p = PR; g.b.r = R; g.b.c = 1;
    g.c.p = out; g.c.r = R; g.c.c = 1;
    if (scale_ptr) { g.scale_ptr = scale_ptr; g.scale_mode = 1; }
    gemm_f64(g, ws, st);
  }
  ws.release(mk);
}

}  // namespace ttamen
