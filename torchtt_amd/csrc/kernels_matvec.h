// Fused local operator of the AMEn sweep (launch.h form): y = Phi_L * A * Phi_R * x in one launch,
// with the two intermediates kept in LDS instead of round-tripping through HBM.
//
//   y[v][l, m, L] = sum_{s,r,n,S,R} PhiL[l,s,r] A[s,m,n,S] PhiR[L,S,R] x[v][r,n,R]
//   (reference: cpp/matvecs.h:196-212, cpp/amen_solve.h:17-27, 'lsr,smnS,LSR,rnR->lmL')
//
// Work split: one workgroup (16 waves) per (vector v, block of BL rows l, block of BR columns R); its 16
// (l, R) pairs (BL * BR <= 16) are the columns of its share of the middle product.
//   stage 1 (VALU)  T1[(s,n)][(l,R)] = sum_r PhiL[l,s,r] x[v][r,n,R]              -> LDS
//   stage 2 (MFMA)  T2[(m,S)][(l,R)] = sum_(s,n) Apm[(m,S),(s,n)] T1[(s,n)][(l,R)] -> LDS
//   stage 3 (MFMA)  Y[(l,m)][L]      = sum_(S,R) T2[(m,S)][(l,R)] PhiR[L,(S,R)]
// Stage 3 sums over this block's R only: with P = ceil(r2 / BR) > 1 blocks of R, each workgroup
// writes its Y to a partial buffer, and the last workgroup of its (v, l block) to finish (counter)
// adds the P partials in block order into y -- a fixed order, so results are deterministic.
//
// Stage 2's A operand is the operator core packed in MFMA fragment order (pack_frag_a): fragment
// (row block rb, k step ks) is 64 consecutive elements, element `lane` = Apm[16 rb + lane % 16]
// [4 ks + lane / 16], zero outside the matrix. Each wave's fragment loads are thus contiguous.
// FP64 uses v_mfma_f64_16x16x4_f64, FP32 v_mfma_f32_16x16x4_f32 (same fragment shapes).
#pragma once
#include "ttamen_internal.h"
#include "kernels_gemm.h"

namespace ttamen {

constexpr int kMvCols = 16;   // (l, R) columns of one workgroup
constexpr int kMvPitch = 17;  // LDS row pitch of T1 / T2 (odd: rows fall on different banks)

template <typename T>
struct MvArgs {
  const T* PhiL;  // l1 x s x r1
  const T* Apk;   // packed operator core, RB x KST fragments of 64
  const T* PhiR;  // l2 x S x r2
  const T* x;     // nvec x r1 x n x r2
  T* y;           // nvec x l1 x n x l2
  T* part;        // P x nvec x l1 x n x l2 partial sums (P > 1), else null
  unsigned* cnt;  // nvec x nlb counters, zero on entry, left at zero
  const int* skip;
  int l1, s, r1, n, S, l2, r2, nvec;
  int BL, BR, nlb, P;
  int KS, MS, KST, RB;  // KS = s n, MS = n S, KST = ceil(KS / 4), RB = ceil(MS / 16)
  int pad;
};

// Dynamic LDS of one workgroup: PhiL slice (r1 x NLS), T1 (4 KST rows), T2 (16 RB rows).
template <typename T, int NLS>
inline size_t mv_lds_bytes(int r1, int KST, int RB) {
  return sizeof(T) * ((size_t)r1 * NLS + (size_t)(4 * KST + 16 * RB) * kMvPitch);
}

// Packs Apm (MS x KS row-major) into fragment order (see above).
template <typename T>
struct KPackFragA {
  struct Args { const T* Apm; T* Apk; int MS, KS, KST, RB; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& a, const Blk& bk) {
    const int64_t total = (int64_t)a.RB * a.KST * 64;
    for (int64_t i = bk.x * (int64_t)256 + threadIdx.x; i < total; i += (int64_t)bk.gx * 256) {
      const int lane = (int)(i & 63);
      const int64_t f = i >> 6;
      const int ks = (int)(f % a.KST), rb = (int)(f / a.KST);
      const int row = rb * 16 + (lane & 15), k = ks * 4 + (lane >> 4);
      a.Apk[i] = (row < a.MS && k < a.KS) ? a.Apm[(int64_t)row * a.KS + k] : T(0);
    }
  }
};

template <typename T, int NLS>
struct KLocalMv {
  typedef MvArgs<T> Args;
  // 16 waves: the grid has about one workgroup per CU at the solver's shapes, so the waves that
  // hide each other's load latency are those of one workgroup (4 per SIMD).
  static constexpr int kThreads = 1024;
  static __device__ void run(const Args& a, const Blk& bk);
};

template <typename T, int NLS>
__device__ void KLocalMv<T, NLS>::run(const Args& a, const Blk& bk) {
  if (a.skip && *a.skip) return;
  typedef typename AccT<T>::type acc_t;
  constexpr int NT = 1024, NW = NT / 64;
  extern __shared__ __attribute__((aligned(16))) char mv_smem[];
  T* PL = reinterpret_cast<T*>(mv_smem);           // [r1][NLS]: PhiL[l0 + j / s][j % s][r]
  T* T1 = PL + (size_t)a.r1 * NLS;                  // [4 KST][kMvPitch]
  T* T2 = T1 + (size_t)4 * a.KST * kMvPitch;        // [16 RB][kMvPitch]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int lb = (int)(bk.x % a.nlb), p = (int)(bk.x / a.nlb), v = (int)bk.y;
  const int l0 = lb * a.BL, R0 = p * a.BR;
  const int BLv = min(a.BL, a.l1 - l0), BRv = min(a.BR, a.r2 - R0);
  const int n = a.n, s = a.s, S = a.S, r1 = a.r1, r2 = a.r2;
  const int nls = BLv * s;  // valid (l, s) pairs of the PhiL slice
  // Stage-2 k steps and stage-1 x values per load round (FP64: half as many, same registers).
  constexpr int CH = sizeof(T) == 8 ? 8 : 16;
  constexpr int XC = sizeof(T) == 8 ? 8 : 16;

  // Every stage needs data from memory; to pay few round trips, the first loads of stage 2 (this
  // wave's first row block's first CH fragments) and of stage 3 (the PhiR fragments of this wave's
  // first output column block) are issued now, and overlap the PhiL slice and stage 1.
  auto load_frag = [&](T (&f)[CH], int rb, int ks0) {
#pragma unroll
    for (int u = 0; u < CH; ++u)
      f[u] = (rb < a.RB && ks0 + u < a.KST) ? a.Apk[(((int64_t)rb * a.KST + ks0 + u) << 6) + lane] : T(0);
  };
  T fa[CH];
  load_frag(fa, wave, 0);
  const int M3 = BLv * n, K3 = S * a.BR, K3S = (K3 + 3) / 4;
  const int RB3 = (M3 + 15) / 16, CB = (a.l2 + 15) / 16;
  T bfr[8];  // K3S <= 8 (S BR <= 32)
  auto load_phir = [&](int cb) {
    const int L = cb * 16 + lr;
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int k = ks * 4 + lk, Si = k / a.BR, R = k - Si * a.BR;
      bfr[ks] = (ks < K3S && k < K3 && L < a.l2 && R < BRv) ? a.PhiR[((int64_t)L * S + Si) * r2 + R0 + R] : T(0);
    }
  };
  int cur_cb = wave < RB3 * CB ? wave / RB3 : -1;
  if (cur_cb >= 0) load_phir(cur_cb);

  // stage-1 work item = (unit (n_i, R), slice of 4 of the NLS (l, s) pairs): its first XC values of x
  // are loaded before the PhiL slice is written
  constexpr int JS = NLS / 4;
  const T* xv = a.x + (int64_t)v * r1 * n * r2;
  const int64_t xr = (int64_t)n * r2;  // stride of r in x
  const int items = n * a.BR * JS;
  T xs[XC];
  auto load_x = [&](int w, int r0) {
    const int u = w / JS;
    const int ni = u / a.BR, R = u - ni * a.BR;
    const bool in = R < BRv;
    const T* xp = xv + (int64_t)ni * r2 + R0 + R;
#pragma unroll
    for (int q = 0; q < XC; ++q) xs[q] = (in && r0 + q < r1) ? xp[(int64_t)(r0 + q) * xr] : T(0);
  };
  if (tid < items) load_x(tid, 0);

  // ---- PhiL slice, and zero T1's padding rows and its columns past BL * BR (never written below)
  for (int i = tid; i < r1 * NLS; i += NT) {
    const int r = i / NLS, j = i - r * NLS;
    PL[i] = (j < nls) ? a.PhiL[((int64_t)(l0 + j / s) * s + (j % s)) * r1 + r] : T(0);
  }
  for (int i = a.KS * kMvPitch + tid; i < 4 * a.KST * kMvPitch; i += NT) T1[i] = T(0);
  {
    const int w = kMvCols - a.BL * a.BR;
    for (int i = tid; i < a.KS * w; i += NT) T1[(i / w) * kMvPitch + a.BL * a.BR + (i % w)] = T(0);
  }
  __syncthreads();

  // ---- stage 1
  for (int w = tid; w < items; w += NT) {
    const int u = w / JS, js = w - u * JS;
    const int ni = u / a.BR, R = u - ni * a.BR;
    const T* plj = PL + 4 * js;
    T acc[4] = {T(0), T(0), T(0), T(0)};
    for (int r0 = 0; r0 < r1; r0 += XC) {
      if (w != tid || r0 != 0) load_x(w, r0);
#pragma unroll
      for (int q = 0; q < XC; ++q) {
        if (r0 + q < r1) {
          const T* pl = plj + (r0 + q) * NLS;
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[j] += pl[j] * xs[q];
        }
      }
    }
    // T1 row (si, ni) = si * n + ni, column (l, R) = l * BR + R
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int j = 4 * js + jj, l = j / s, si = j - l * s;
      if (l < a.BL) T1[(si * n + ni) * kMvPitch + l * a.BR + R] = acc[jj];
    }
  }
  __syncthreads();

  // ---- stage 2: wave w takes row blocks w, w + 16, ...; K streamed CH k steps per load round,
  // the next round's fragments in flight while the current round feeds the matrix cores
  for (int rb = wave; rb < a.RB; rb += NW) {
    acc_t acc = acc_t{0, 0, 0, 0};
    if (rb != wave) load_frag(fa, rb, 0);
    for (int ks0 = 0; ks0 < a.KST; ks0 += 2 * CH) {
      T fb[CH];
      if (ks0 + CH < a.KST) load_frag(fb, rb, ks0 + CH);
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (ks0 + u < a.KST) acc = mfma16(fa[u], T1[((ks0 + u) * 4 + lk) * kMvPitch + lr], acc);
      if (ks0 + 2 * CH < a.KST) load_frag(fa, rb, ks0 + 2 * CH);
      if (ks0 + CH < a.KST) {
#pragma unroll
        for (int u = 0; u < CH; ++u)
          if (ks0 + CH + u < a.KST) acc = mfma16(fb[u], T1[((ks0 + CH + u) * 4 + lk) * kMvPitch + lr], acc);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) T2[(rb * 16 + acc_row<T>(lk, q)) * kMvPitch + lr] = acc[q];
  }
  __syncthreads();

  // ---- stage 3: output blocks (row block of (l, m), column block of L) round-robin over waves,
  // column-block major so that a wave keeps its PhiR fragments while its row blocks change
  T* dst = (a.P > 1) ? a.part + ((int64_t)p * a.nvec + v) * a.l1 * n * a.l2 : a.y + (int64_t)v * a.l1 * n * a.l2;
  dst += (int64_t)l0 * n * a.l2;
  for (int q = wave; q < RB3 * CB; q += NW) {
    const int cb = q / RB3, rb = q - cb * RB3;
    if (cb != cur_cb) {
      cur_cb = cb;
      load_phir(cb);
    }
    const int i = rb * 16 + lr;  // A fragment row of this lane
    const int li = i / n, mi = i - li * n;
    acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      if (ks < K3S) {
        const int k = ks * 4 + lk, Si = k / a.BR, R = k - Si * a.BR;
        const T af = (i < M3 && k < K3) ? T2[(mi * S + Si) * kMvPitch + li * a.BR + R] : T(0);
        acc = mfma16(af, bfr[ks], acc);
      }
    }
    const int L = cb * 16 + lr;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int row = rb * 16 + acc_row<T>(lk, qq);
      if (row < M3 && L < a.l2) {
        const int lo = row / n, mo = row - lo * n;
        dst[((int64_t)lo * n + mo) * a.l2 + L] = acc[qq];
      }
    }
  }
  if (a.P <= 1) return;

  // ---- the last workgroup of this (v, l block) adds the P partials in block order. Release: each
  // wave waits until its partial stores have reached the L2 (s_waitcnt), the barrier orders that
  // before thread 0's device-scope fence (which writes the L2's dirty lines back) and the counter;
  // acquire: the last workgroup fences before it reads the other workgroups' partials. One fence
  // per workgroup: a fence in every wave (16 L2 write-backs per workgroup) cost 20-90 us.
  __shared__ bool last;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(a.cnt + (int64_t)v * a.nlb + lb, 1u);
    last = (prev == (unsigned)a.P - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  const int64_t ysz = (int64_t)a.l1 * n * a.l2, off = (int64_t)l0 * n * a.l2;
  const int64_t cnt = (int64_t)BLv * n * a.l2;
  const T* pp = a.part + (int64_t)v * ysz + off;
  T* yo = a.y + (int64_t)v * ysz + off;
  const int64_t pstride = (int64_t)a.nvec * ysz;
  for (int64_t e = tid; e < cnt; e += NT) {
    // the partials of element e in groups of six loaded together, added in block order
    T sum = pp[e];
    for (int q0 = 1; q0 < a.P; q0 += 6) {
      T part[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) part[q] = (q0 + q < a.P) ? pp[(int64_t)(q0 + q) * pstride + e] : T(0);
#pragma unroll
      for (int q = 0; q < 6; ++q)
        if (q0 + q < a.P) sum += part[q];
    }
    yo[e] = sum;
  }
  if (tid == 0) a.cnt[(int64_t)v * a.nlb + lb] = 0u;
}

}  // namespace ttamen
