// Kernels of the device-resident GMRES (bodies in the launch.h form), shared by the
// GMRES driver and the cohort's merged launch kind.
#pragma once
#include "ttamen_internal.h"
#include <type_traits>
#include "gmres.h"
#include <algorithm>

namespace ttamen {

constexpr int kGB = 256;  // threads per block of the streaming kernels

inline int gm_blocks(int64_t N) {  // about two elements per thread, at most kGmresMaxBlocks blocks
  int64_t b = (N + kGB * 2 - 1) / (kGB * 2);
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kGmresMaxBlocks));
}

__device__ __forceinline__ double gm_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ bool gm_skip(const int* a) { return a && *a; }

// out[i] = sum_b P[i * kGmresMaxBlocks + b] over b < nb, for i < rows; every thread of
// the block gets the results in LDS. Wave wv reduces rows wv, wv + 4, ...; lane l adds the
// partials l, l + 64, l + 128, l + 192 (four independent loads), then the wave's xor tree.
// Fixed order: deterministic, and the same in every block.
__device__ __forceinline__ void gm_reduce_rows(const double* P, int rows, int nb, double* out) {
  static_assert(kGmresMaxBlocks <= 4 * 64, "four partials per lane");
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  constexpr int W = kGB / 64;
  for (int i = wv; i < rows; i += 2 * W) {  // rows i and i + W: eight loads in flight
    const bool two = i + W < rows;
    const double* r0 = P + (int64_t)i * kGmresMaxBlocks;
    const double* r1 = P + (int64_t)(two ? i + W : i) * kGmresMaxBlocks;
    double p[4], q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = lane + 64 * u < nb;
      p[u] = in ? r0[lane + 64 * u] : 0.0;
      q[u] = in ? r1[lane + 64 * u] : 0.0;
    }
    const double v0 = gm_wave_sum((p[0] + p[1]) + (p[2] + p[3]));
    const double v1 = gm_wave_sum((q[0] + q[1]) + (q[2] + q[3]));
    if (lane == 0) {
      out[i] = v0;
      if (two) out[i + W] = v1;
    }
  }
  __syncthreads();
}

// out[0] = sum of the nba partials of row Pa, out[1] = of the nbb partials of row Pb (each as
// gm_reduce_rows does it), in one round: wave 0 reduces Pa, wave 1 reduces Pb.
__device__ __forceinline__ void gm_reduce_pair(const double* Pa, int nba, const double* Pb, int nbb, double* out) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (wv < 2) {
    const double* row = wv == 0 ? Pa : Pb;
    const int nb = wv == 0 ? nba : nbb;
    double p[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) p[u] = (lane + 64 * u < nb) ? row[lane + 64 * u] : 0.0;
    const double v = gm_wave_sum((p[0] + p[1]) + (p[2] + p[3]));
    if (lane == 0) out[wv] = v;
  }
  __syncthreads();
}

// Per-block partial dot products of the rows Q_0..Q_{rows-1} (and optionally of w itself)
// with the block's chunk [lo, hi) of w: P[i * kGmresMaxBlocks + blk]. The block walks the chunk
// once per group of kDotRows rows, each thread holding one accumulator per row of the group, so
// that a thread has the group's loads in flight together (the chunk of w is re-read from cache).
constexpr int kDotRows = 8;
template <typename T>
__device__ __forceinline__ void gm_chunk_dots(const T* Q, int64_t N, int rows, const T* w,
                                              bool with_ww, int64_t lo, int64_t hi, double* P,
                                              unsigned blk) {
  __shared__ double red[kGB / 64][kDotRows];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nr = rows + (with_ww ? 1 : 0);
  for (int i0 = 0; i0 < nr; i0 += kDotRows) {
    const T* src[kDotRows];
#pragma unroll
    for (int q = 0; q < kDotRows; ++q) {
      const int i = min(i0 + q, nr - 1);  // rows past the last repeat it; not stored
      src[q] = (i < rows) ? Q + (int64_t)i * N : w;
    }
    double acc[kDotRows];
#pragma unroll
    for (int q = 0; q < kDotRows; ++q) acc[q] = 0.0;
    for (int64_t j = lo + t; j < hi; j += kGB) {
      const double wj = (double)w[j];
      double v[kDotRows];
#pragma unroll
      for (int q = 0; q < kDotRows; ++q) v[q] = (double)src[q][j];
#pragma unroll
      for (int q = 0; q < kDotRows; ++q) acc[q] += v[q] * wj;
    }
#pragma unroll
    for (int q = 0; q < kDotRows; ++q) {
      const double s = gm_wave_sum(acc[q]);
      if (lane == 0) red[wv][q] = s;
    }
    __syncthreads();
    if (t < kDotRows && i0 + t < nr) {
      double s = 0.0;
      for (int w4 = 0; w4 < kGB / 64; ++w4) s += red[w4][t];
      P[(int64_t)(i0 + t) * kGmresMaxBlocks + blk] = s;
    }
    __syncthreads();
  }
}

// w[lo:hi] -= sum_i h_i Q_i[lo:hi]; returns this thread's share of ||w[lo:hi]||^2.
template <typename T>
__device__ __forceinline__ double gm_chunk_update(const T* Q, int64_t N, int kc, T* w,
                                                  const double* h, int64_t lo, int64_t hi) {
  double acc = 0.0;
  for (int64_t j = lo + threadIdx.x; j < hi; j += kGB) {
    double v = (double)w[j];
    int i = 0;
    for (; i + kDotRows <= kc; i += kDotRows) {  // a group's loads in flight together, same order
      double q[kDotRows];
#pragma unroll
      for (int u = 0; u < kDotRows; ++u) q[u] = (double)Q[(int64_t)(i + u) * N + j];
#pragma unroll
      for (int u = 0; u < kDotRows; ++u) v -= h[i + u] * q[u];
    }
    for (; i < kc; ++i) v -= h[i] * (double)Q[(int64_t)i * N + j];
    w[j] = (T)v;
    acc += v * v;
  }
  return acc;
}

// Arnoldi step k, orthogonalisation in three kernels (classical Gram-Schmidt, selective second
// pass with the DGKS criterion 1/sqrt(2); Daniel, Gragg, Kaufman, Stewart, Math. Comp. 30
// (1976)), all deterministic (fixed-order partials reduced by every consumer block):
//   KCgsDots     P1 = [Q^T w ; w.w] partials (or the local operator's fused kernel writes them: nb1 partials
//                per row, one per workgroup of that kernel, instead of one per block of these kernels)
//   KCgsUpdate   h = sum P1, w' = w - Q h, Pn1 = ||w'||^2 partials and P2 = Q^T w' partials
//                (the second pass's dots, from the chunk the block just updated)
//   KGmresClose  if ||w'|| < ||w|| / sqrt(2): h2 = sum P2, w'' = w' - Q h2, H(:,k) += h2 and
//                ||w''||^2 = ||w'||^2 - ||h2||^2; q_{k+1} = w'' / ||w''||; Givens on H(:,k).
//                When ||h2||^2 > ||w'||^2 / 2 that difference cancels (w' nearly in span(Q), near
//                Krylov breakdown): ||w''||^2 is then reduced from w'' itself -- per-block partials,
//                summed in block order by the last block to finish, which normalises q_{k+1} and
//                applies the Givens rotations (the reference computes this norm explicitly,
//                cpp/gmres.h:106).
template <typename T>
struct KCgsDots {
  struct Args { const T* Q; int64_t N; int kc; const T* w; double* P; const int* skip; };
  static constexpr int kThreads = kGB;
  static constexpr int kMega = std::is_same<T, double>::value ? 4 : 5;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk) {
    if (gm_skip(a.skip)) return;
    const int64_t chunk = (a.N + bk.gx - 1) / bk.gx;
    const int64_t lo = bk.x * chunk, hi = min(a.N, lo + chunk);
    gm_chunk_dots(a.Q, a.N, a.kc, a.w, true, lo, hi, a.P, bk.x);
  }
};

template <typename T>
struct KCgsUpdate {
  struct Args {
    const T* Q; int64_t N; int kc; T* w; const double* P1; double* Hcol; double* Pn1; double* P2;
    const int* skip; int nb1;  // partials per row of P1 (0: one per block of this launch)
  };
  static constexpr int kThreads = kGB;
  static constexpr int kMega = std::is_same<T, double>::value ? 6 : 7;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk) {
    if (gm_skip(a.skip)) return;
    __shared__ double hs[kGmresMaxCols + 1];
    __shared__ double sh[kGB / 64];
    const int kc = a.kc;
    gm_reduce_rows(a.P1, kc, a.nb1 > 0 ? a.nb1 : (int)bk.gx, hs);
    if (bk.x == 0)
      for (int i = threadIdx.x; i < kc; i += kGB) a.Hcol[i] = hs[i];
    const int64_t chunk = (a.N + bk.gx - 1) / bk.gx;
    const int64_t lo = bk.x * chunk, hi = min(a.N, lo + chunk);
    double acc = gm_chunk_update(a.Q, a.N, kc, a.w, hs, lo, hi);
    acc = gm_wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();  // also orders this block's writes of w before the dots below
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < kGB / 64; ++i) t += sh[i];
      a.Pn1[bk.x] = t;
    }
    gm_chunk_dots(a.Q, a.N, kc, (const T*)a.w, false, lo, hi, a.P2, bk.x);
  }
};

template <typename T>
struct KGmresClose {
  struct Args {
    const T* Q; int64_t N; int kc; T* w; T* qn; const double* Pw; const double* Pn1;
    const double* P2; double* Hcol; GmresDev g; int k; const int* done_in; int* reorth_count;
    double* Pex;  // per-block partials of ||w''||^2 (exact-norm path)
    int nb1;      // partials in Pw (0: one per block of this launch)
    int* host_done;  // pinned host word that receives the done flag once the step is closed, or null
  };
  static constexpr int kThreads = kGB;
  static constexpr int kMega = std::is_same<T, double>::value ? 8 : 9;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk) {
    if (*a.done_in) {
      if (bk.x == 0 && threadIdx.x == 0 && a.host_done) *(volatile int*)a.host_done = 1;
      return;
    }
    __shared__ double h2[kGmresMaxCols + 1];
    __shared__ double nn[2];
    const int kc = a.kc;
    // Whether the second (re-orthogonalisation) pass runs: the first pass cancelled more than
    // 1 - 1/sqrt(2) of ||w|| (Daniel, Gragg, Kaufman and Stewart, Math. Comp. 30 (1976)).
    gm_reduce_pair(a.Pw, a.nb1 > 0 ? a.nb1 : (int)bk.gx, a.Pn1, (int)bk.gx, nn);  // ||w||^2, ||w'||^2
    const bool reorth = !(sqrt(nn[1]) > 0.7071067811865476 * sqrt(nn[0]));
    double hn2 = nn[1];
    const int64_t chunk = (a.N + bk.gx - 1) / bk.gx;
    const int64_t lo = bk.x * chunk, hi = min(a.N, lo + chunk);
    bool exact = false;  // the same decision in every block (same partials, same order)
    if (reorth) {
      gm_reduce_rows(a.P2, kc, bk.gx, h2);
      double s2 = 0.0;
      for (int i = 0; i < kc; ++i) s2 += h2[i] * h2[i];
      exact = s2 > 0.5 * hn2;
      hn2 -= s2;
      const double part = gm_chunk_update(a.Q, a.N, kc, a.w, h2, lo, hi);
      if (bk.x == 0) {
        for (int i = threadIdx.x; i < kc; i += kGB) a.Hcol[i] += h2[i];
        if (threadIdx.x == 0 && a.reorth_count) *a.reorth_count += 1;
      }
      if (exact) {
        // ||w''||^2 of this block's chunk; the last block to finish adds the partials
        __shared__ double sw[kGB / 64];
        __shared__ bool last;
        const double v = gm_wave_sum(part);
        if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
        __syncthreads();
        // release: each wave's stores (w'', H, the partial) have reached memory before the counter
        if (threadIdx.x == 0) a.Pex[bk.x] = (sw[0] + sw[1]) + (sw[2] + sw[3]);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0) {
          __threadfence();
          const int prev = atomicAdd(a.g.istate + GM_CNT, 1);
          last = (prev == (int)bk.gx - 1);
        }
        __syncthreads();
        if (!last) return;
        __threadfence();
        if (threadIdx.x == 0) {
          double t = 0.0;
          for (unsigned b = 0; b < bk.gx; ++b) t += ((volatile double*)a.Pex)[b];
          sw[0] = t;
          a.g.istate[GM_CNT] = 0;
          a.g.istate[GM_EXACT] += 1;
        }
        __syncthreads();
        hn2 = sw[0];
      }
      __syncthreads();
    }
    const double hn = sqrt(hn2 > 0.0 ? hn2 : 0.0);
    const double inv = (hn > 0.0) ? 1.0 / hn : 0.0;
    if (exact) {  // the last block normalises the whole vector and closes the step
      for (int64_t j = threadIdx.x; j < a.N; j += kGB) a.qn[j] = (T)((double)((volatile T*)a.w)[j] * inv);
    } else {
      for (int64_t j = lo + threadIdx.x; j < hi; j += kGB) a.qn[j] = (T)((double)a.w[j] * inv);
      if (bk.x != 0) return;
    }
    // Givens rotations of column k (block 0): its entries, the earlier rotations and the
    // scalars are staged in LDS by all threads at once, then one thread applies them in order.
    __shared__ double hc[kGmresMaxCols + 1], cs[kGmresMaxCols], sn[kGmresMaxCols], sc[3];
    const GmresDev g = a.g;
    const int k = a.k;
    double* h = g.H + (int64_t)k * (g.m + 1);
    for (int i = threadIdx.x; i < k; i += kGB) {
      hc[i] = h[i];
      cs[i] = g.cs[i];
      sn[i] = g.sn[i];
    }
    if (threadIdx.x == kGB - 1) hc[k] = h[k];
    if (threadIdx.x == kGB - 2) sc[0] = g.beta[k];
    if (threadIdx.x == kGB - 3) sc[1] = g.dstate[GM_BNORM];
    if (threadIdx.x == kGB - 4) sc[2] = g.dstate[GM_THR];
    __syncthreads();
    if (threadIdx.x == 0) {
      hc[k + 1] = hn;
      for (int i = 0; i < k; ++i) {
        double t = cs[i] * hc[i] + sn[i] * hc[i + 1];
        hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
        hc[i] = t;
      }
      double den = sqrt(hc[k] * hc[k] + hc[k + 1] * hc[k + 1]);
      double c = 1.0, s = 0.0;
      if (den > 0.0) { c = hc[k] / den; s = hc[k + 1] / den; }
      hc[k] = c * hc[k] + s * hc[k + 1];
      hc[k + 1] = 0.0;
      g.cs[k] = c;
      g.sn[k] = s;
      const double bk0 = sc[0];
      const double bk1 = -s * bk0;
      g.beta[k + 1] = bk1;
      g.beta[k] = c * bk0;
      const double err = fabs(bk1) / sc[1];
      g.istate[GM_KCOLS] = k + 1;
      g.istate[GM_ITERS] += 1;
      g.dstate[GM_ERR] = err;
      if (err <= sc[2]) {
        g.istate[GM_FLAG] = 1;
        g.istate[GM_DONE] = 1;
      } else if (k + 1 >= g.m) {
        g.istate[GM_DONE] = 1;  // cycle exhausted (flag stays 0)
      }
      if (a.host_done) *(volatile int*)a.host_done = g.istate[GM_DONE];
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= k + 1; i += kGB) h[i] = hc[i];
  }
};

// Start of a cycle: r already holds rhs - Op(x0) and dstate[GM_RNORM] = ||r||.
// q_0 = r / ||r||, beta = (||r||, 0, ...), done = (||r|| <= 0).
template <typename T>
struct KGmresBegin {
  struct Args { const T* r; T* q0; int64_t N; GmresDev g; };
  static constexpr int kThreads = kGB;
  static constexpr int kMega = std::is_same<T, double>::value ? 10 : 11;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};
template <typename T>
__device__ __forceinline__ void KGmresBegin<T>::run(const Args& a, const Blk& bk) {
  const T* r = a.r; T* q0 = a.q0; const int64_t N = a.N; const GmresDev g = a.g;
  const double rn = g.dstate[GM_RNORM];
  const double inv = (rn > 0.0) ? 1.0 / rn : 0.0;
  for (int64_t j = bk.x * (int64_t)kGB + threadIdx.x; j < N; j += (int64_t)bk.gx * kGB)
    q0[j] = (T)((double)r[j] * inv);
  if (bk.x == 0) {
    for (int i = threadIdx.x; i <= g.m; i += kGB) g.beta[i] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
      g.beta[0] = rn;
      g.istate[GM_KCOLS] = 0;
      const int zero = !(rn > 0.0);
      g.istate[GM_DONE] = zero;
      if (zero) g.istate[GM_FLAG] = 1;
    }
  }
}

// thr = real_tol * norm_rhs / ||drhs||,  bnorm = ||drhs||  (reference amen_solve.h:340-343)
struct KGmresSetup {
  struct Args { GmresDev g; const double* norm_rhs; const double* dnorm; double real_tol; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 14;
  static __device__ __forceinline__ void run(const Args& a, const Blk&);
};
__device__ __forceinline__ void KGmresSetup::run(const Args& a, const Blk&) {
  if (threadIdx.x != 0) return;
  const GmresDev g = a.g; const double* norm_rhs = a.norm_rhs; const double* dnorm = a.dnorm;
  const double real_tol = a.real_tol;
  g.dstate[GM_BNORM] = *dnorm;
  g.dstate[GM_THR] = real_tol * (*norm_rhs) / (*dnorm);
  g.dstate[GM_RNORM] = *dnorm;
  g.istate[GM_FLAG] = 0;
  g.istate[GM_ITERS] = 0;
  g.istate[GM_DONE] = 0;
  g.istate[GM_KCOLS] = 0;
  g.istate[GM_REORTHS] = 0;
  g.istate[GM_EXACT] = 0;
  g.istate[GM_CNT] = 0;
}

// y = H(0:kc,0:kc) \ beta(0:kc) (upper triangular), one thread.
struct KGmresY {
  struct Args { GmresDev g; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 15;
  static __device__ __forceinline__ void run(const Args& a, const Blk&);
};
// y is kept in LDS during the back substitution and written out at the end: with no global store
// inside the loop, H's loads do not wait behind the previous row's store of y.
__device__ __forceinline__ void KGmresY::run(const Args& a, const Blk&) {
  __shared__ double ys[kGmresMaxCols];
  const GmresDev g = a.g;
  const int kc = g.istate[GM_KCOLS];
  const int ld = g.m + 1;
  if (threadIdx.x == 0) {
    for (int i = kc - 1; i >= 0; --i) {
      double s = g.beta[i];
      for (int j = i + 1; j < kc; ++j) s -= g.H[(int64_t)j * ld + i] * ys[j];
      const double d = g.H[(int64_t)i * ld + i];
      ys[i] = (d != 0.0) ? s / d : 0.0;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kc; i += blockDim.x) g.y[i] = ys[i];
}

// x += sum_{i<kc} y_i Q_i
template <typename T>
struct KGmresXupd {
  struct Args { const T* Q; int64_t N; T* x; GmresDev g; };
  static constexpr int kThreads = kGB;
  static constexpr int kMega = std::is_same<T, double>::value ? 12 : 13;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};
template <typename T>
__device__ __forceinline__ void KGmresXupd<T>::run(const Args& a, const Blk& bk) {
  const T* Q = a.Q; const int64_t N = a.N; T* x = a.x; const GmresDev g = a.g;
  const int kc = g.istate[GM_KCOLS];
  __shared__ double y[kGmresMaxCols];
  for (int i = threadIdx.x; i < kc; i += kGB) y[i] = g.y[i];
  __syncthreads();
  for (int64_t j = bk.x * (int64_t)kGB + threadIdx.x; j < N; j += (int64_t)bk.gx * kGB) {
    double v = (double)x[j];
    for (int i = 0; i < kc; ++i) v += y[i] * (double)Q[(int64_t)i * N + j];
    x[j] = (T)v;
  }
}

struct KGmresRearm {
  struct Args { GmresDev g; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 16;
  static __device__ __forceinline__ void run(const Args& a, const Blk&) {
    if (threadIdx.x == 0) a.g.istate[GM_DONE] = 0;
  }
};

}  // namespace ttamen
