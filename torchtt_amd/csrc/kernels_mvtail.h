// Stages 2 and 3 of the local operator in one launch (launch.h form):
//
//   t2^T[R][m] = sum_k t1[v,l][k][R] Apm[(m,S)][k]              k = (s,n), one 16x16 block per (R block, S)
//   y[v,l,m,L] = sum_{S,R} PhiR[L][S][R] t2^T[R][m]               (reference: cpp/matvecs.h:196-212,
//                                                                 cpp/amen_solve.h:17-27)
//
// The stage-2 product is computed transposed so that its MFMA result registers are, as they stand,
// the B operand of the stage-3 MFMAs: in a 16x16 result block, lane l holds rows R = dR(l/16, v) of
// column m = l%16 in register v, and the B operand of one 16x16x4 step is row k = l/16 of column
// l%16, so register v feeds step v with k <-> R = dR(l/16, v). No intermediate leaves the registers.
//
// Workgroup = (vector v and row l of the output, block of 16 rows m); 4 waves. An item is one R block
// with a group of SM values of S: its K loop loads the t1 fragment once and uses it against the SM
// operator rows (SM independent accumulator chains); its K steps are software-pipelined DEPTH deep
// (every step's loads are issued DEPTH steps ahead, the step index clamped at the last full step so
// that the loads are unconditional). The waves take items round-robin and keep partial sums of
// y^T[L][m] for all L blocks; the four partial sums are added in LDS in a fixed order,
// (p0 + p2) + (p1 + p3), so the result is deterministic.
//
// Operand layouts: t1[v][l][k][R] (k = s n + n', R fastest), Apm[(m, S)][k] row-major (k fastest),
// PhiR[L][S][R], y[v][l][m][L]. Within a 16-deep K step lane group g's element for MFMA jj is
// k = kofs(g, jj) (a permutation of the 16: the sum over k is the same, and contiguous per lane so
// that operator rows load as vectors when K allows).
//
// Operator tiles. The operator cores of finite-difference and finite-element operators in TT form are
// block-sparse: their n x n blocks A[s, :, :, S] are identities, stencils, diagonals or zero. With the
// optional tile masks (KTileMask), an item runs only the K steps where its 16 operator rows hold a
// nonzero for one of its S values, and a chain with no nonzero tile skips stage 3. A skipped step would
// have added exact zero products to the accumulators, so the results are the same as the dense loop's
// (up to the sign of a zero).
#pragma once
#include "ttamen_internal.h"
#include "kernels_gemm.h"
#include "gmres.h"

namespace ttamen {

constexpr int kTailMaxLB = 9;  // l2 <= 144

template <typename T> __device__ __forceinline__ int tail_kofs(int g, int jj);
template <> __device__ __forceinline__ int tail_kofs<float>(int g, int jj) { return 4 * g + jj; }
template <> __device__ __forceinline__ int tail_kofs<double>(int g, int jj) { return 2 * g + (jj & 1) + 8 * (jj >> 1); }

// Row R (within a 16x16 result block) held by lane group g in result register v.
template <typename T> __device__ __forceinline__ int tail_rrow(int g, int v);
template <> __device__ __forceinline__ int tail_rrow<double>(int g, int v) { return g + 4 * v; }
template <> __device__ __forceinline__ int tail_rrow<float>(int g, int v) { return 4 * g + v; }

typedef double tail_d2 __attribute__((ext_vector_type(2)));

// The four k values kofs(g, 0..3) of a lane from a k-contiguous row (16-byte aligned vector loads).
template <typename T>
__device__ __forceinline__ void tail_vec4(const T* p, T* out) {
  if constexpr (std::is_same<T, float>::value) {
    const f4 v = *reinterpret_cast<const f4*>(p);
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; out[3] = v[3];
  } else {
    const tail_d2 v0 = *reinterpret_cast<const tail_d2*>(p);
    const tail_d2 v1 = *reinterpret_cast<const tail_d2*>(p + 8);
    out[0] = v0[0]; out[1] = v0[1]; out[2] = v1[0]; out[3] = v1[1];
  }
}

// LDS of the partial-sum tree: 2 x kTailMaxLB x 256 values, one buffer for every instantiation.
__device__ __forceinline__ char* tail_lds() {
  __shared__ __attribute__((aligned(16))) char smem[2 * kTailMaxLB * 256 * sizeof(double)];
  return smem;
}

template <typename T>
struct KMvTail {
  struct Args {
    const T* t1;
    const T* Apm;
    const T* PhiR;
    T* y;
    const int* skip;
    const uint64_t* tmask;  // tile masks (KTileMask) of Apm, or null: every tile; only for K <= kTileMaskMaxK
    int K, r2, nm, S, l2;
    int LB;   // ceil(l2 / 16) <= kTailMaxLB
    int apv;  // operator rows load as vectors (K a multiple of the vector width, aligned base)
    // Optional (kc > 0; one vector, grid x grid.y <= kGmresMaxBlocks workgroups): the first pass of the
    // GMRES orthogonalisation on the output, P[i * kGmresMaxBlocks + blk] = this workgroup's share of Q_i . y for the
    // basis rows i < kc and of y . y in row kc (blk = bk.x * grid.y + bk.y; Q_i = Qd + i * N).
    const T* Qd;
    double* P;
    int64_t N;
    int kc;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};

// P[i * kGmresMaxBlocks + blk] = sum_{e < E} Q_i[off + e] y[e] for i < kc, and y . y in row kc, with y (E values) in
// LDS. Wave w takes the rows w, w + 4, ..., four at a time (their loads in flight together); lane l adds
// the elements e = l, l + 64, ... of a row in order, then the wave's xor tree gives the row's share: no
// barrier, a fixed order. Accumulated in FP64, as the separate dot kernel does.
template <typename T>
__device__ __forceinline__ void mvtail_dots(const typename KMvTail<T>::Args& a, const T* y, int E, int64_t off,
                                            int blk) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nr = a.kc + 1;
  for (int i0 = w; i0 < nr; i0 += 16) {  // rows i0, i0 + 4, i0 + 8, i0 + 12: w, w + 4, ... over the rounds
    const T* src[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 4 * q < nr ? i0 + 4 * q : nr - 1;  // w's rows: i0, i0 + 4, i0 + 8, i0 + 12
      src[q] = (i < a.kc) ? a.Qd + (int64_t)i * a.N + off : nullptr;
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = lane; e < E; e += 64) {
      const double ye = (double)y[e];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] += (src[q] ? (double)src[q][e] : ye) * ye;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double s = acc[q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0 && i0 + 4 * q < nr) a.P[(int64_t)(i0 + 4 * q) * kGmresMaxBlocks + blk] = s;
    }
  }
}

template <typename T, int LB, int SM, int DEPTH, bool APV>
__device__ __forceinline__ void mvtail_body(const typename KMvTail<T>::Args& a, const Blk& bk) {
  typedef typename AccT<T>::type acc_t;
  T* red = reinterpret_cast<T*>(tail_lds());
  const int64_t vl = bk.x;
  const int m0 = bk.y * 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  const T* t1 = a.t1 + vl * a.K * a.r2;
  const int RB = (a.r2 + 15) / 16;
  const int SG = (a.S + SM - 1) / SM;
  const int items = RB * SG;
  const int mrow = min(m0 + r, a.nm - 1);
  const int K = a.K;
  const int64_t r2 = a.r2;
  const bool masked = a.tmask != nullptr;
  const uint64_t* tmask = a.tmask + (masked ? (int64_t)bk.y * a.S : 0);

  acc_t yacc[LB];
#pragma unroll
  for (int lb = 0; lb < LB; ++lb) yacc[lb] = acc_t{0, 0, 0, 0};

  for (int it = wave; it < items; it += 4) {
    const int rb = it / SG, s0 = (it - rb * SG) * SM;
    const T* arow = t1 + min(rb * 16 + r, a.r2 - 1);
    const T* brow[SM];
#pragma unroll
    for (int q = 0; q < SM; ++q) brow[q] = a.Apm + ((int64_t)mrow * a.S + min(s0 + q, a.S - 1)) * K;
    acc_t d[SM];
#pragma unroll
    for (int q = 0; q < SM; ++q) d[q] = acc_t{0, 0, 0, 0};
    // Steps of the K loop: bit j of mq[q] marks step j (k in [16 j, 16 j + 16)) as one where chain q's
    // operator rows hold a nonzero (chains past S have none); the loop runs the steps of any chain.
    // Without masks, every step.
    const int nsteps = K / 16;
    uint64_t mq[SM];
    uint64_t any = 0;
#pragma unroll
    for (int q = 0; q < SM; ++q) {
      mq[q] = (s0 + q < a.S) ? (masked ? tmask[s0 + q] : ~0ull) : 0ull;
      any |= mq[q];
    }
    uint64_t rem = masked ? (any & (nsteps >= 64 ? ~0ull : ((1ull << nsteps) - 1))) : 0ull;
    const int cnt = masked ? __popcll(rem) : nsteps;
    int pos = 0, lastj = 0;
    auto next = [&]() -> int {  // the next step to load; past the last one, the last one again
      if (masked) {
        if (rem) {
          lastj = __builtin_ctzll(rem);
          rem &= rem - 1;
        }
      } else {
        lastj = min(pos, nsteps - 1);
        ++pos;
      }
      return lastj;
    };
    struct Frag { T a[4]; T b[SM][4]; };
    auto load = [&](Frag& f, int j) {
      const int kt = 16 * j;
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) f.a[jj] = arow[(int64_t)(kt + tail_kofs<T>(g, jj)) * r2];
#pragma unroll
      for (int q = 0; q < SM; ++q) {
        if constexpr (APV) {
          tail_vec4<T>(brow[q] + kt + tail_kofs<T>(g, 0), f.b[q]);
        } else {
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) f.b[q][jj] = brow[q][kt + tail_kofs<T>(g, jj)];
        }
      }
    };
    auto mma = [&](const Frag& f) {
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int q = 0; q < SM; ++q) d[q] = mfma16(f.a[jj], f.b[q][jj], d[q]);
    };
    if (cnt > 0) {
      Frag f[DEPTH];
#pragma unroll
      for (int dd = 0; dd < DEPTH; ++dd) load(f[dd], next());
      const int full = cnt / DEPTH * DEPTH;
      for (int s = 0; s < full; s += DEPTH) {
#pragma unroll
        for (int dd = 0; dd < DEPTH; ++dd) {
          mma(f[dd]);
          load(f[dd], next());
        }
      }
#pragma unroll
      for (int dd = 0; dd < DEPTH - 1; ++dd)
        if (full + dd < cnt) mma(f[dd]);
    }
    // the last, partial K step (elements past K are zero)
    if (16 * nsteps < K && (!masked || ((any >> nsteps) & 1ull))) {
      Frag f;
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int k = 16 * nsteps + tail_kofs<T>(g, jj);
        const bool in = k < K;
        const int kc = in ? k : 0;
        const T va = arow[(int64_t)kc * r2];
        f.a[jj] = in ? va : T(0);
#pragma unroll
        for (int q = 0; q < SM; ++q) {
          const T vb = brow[q][kc];
          f.b[q][jj] = in ? vb : T(0);
        }
      }
      mma(f);
    }
    // stage 3: y^T[L][m] += sum_{R in block rb} PhiR[L][S][R] t2^T[R][m]
#pragma unroll
    for (int q = 0; q < SM; ++q) {
      if (mq[q] != 0) {
#pragma unroll
        for (int lb = 0; lb < LB; ++lb) {
          const T* prow = a.PhiR + ((int64_t)min(lb * 16 + r, a.l2 - 1) * a.S + s0 + q) * r2;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int R = rb * 16 + tail_rrow<T>(g, v);
            const T pv = prow[min(R, a.r2 - 1)];
            yacc[lb] = mfma16(R < a.r2 ? pv : T(0), d[q][v], yacc[lb]);
          }
        }
      }
    }
  }

  // (p0 + p2) + (p1 + p3) through LDS, then wave 0 writes its 16 rows m of y.
  auto slot = [&](int s, int lb, int v) -> T& { return red[((s * LB + lb) * 4 + v) * 64 + lane]; };
  if (wave >= 2) {
#pragma unroll
    for (int lb = 0; lb < LB; ++lb)
#pragma unroll
      for (int v = 0; v < 4; ++v) slot(wave - 2, lb, v) = yacc[lb][v];
  }
  __syncthreads();
  if (wave < 2) {
#pragma unroll
    for (int lb = 0; lb < LB; ++lb)
#pragma unroll
      for (int v = 0; v < 4; ++v) yacc[lb][v] = yacc[lb][v] + slot(wave, lb, v);
  }
  __syncthreads();
  if (wave == 1) {
#pragma unroll
    for (int lb = 0; lb < LB; ++lb)
#pragma unroll
      for (int v = 0; v < 4; ++v) slot(0, lb, v) = yacc[lb][v];
  }
  __syncthreads();
  // With the fused dots, the output block also goes to the second slot's LDS (free by now), row-major
  // as in y, for all four waves to read.
  T* ych = red + LB * 256;
  if (wave == 0 && m0 + r < a.nm) {
    T* y = a.y + (vl * a.nm + m0 + r) * a.l2;
#pragma unroll
    for (int lb = 0; lb < LB; ++lb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int L = lb * 16 + acc_row<T>(g, v);
        if (L < a.l2) {
          const T val = yacc[lb][v] + slot(0, lb, v);
          y[L] = val;
          if (a.kc > 0) ych[r * a.l2 + L] = val;
        }
      }
  }
  if (a.kc > 0) {
    __syncthreads();
    const int rows = min(16, a.nm - m0);
    mvtail_dots<T>(a, ych, rows * a.l2, (vl * a.nm + m0) * a.l2, (int)(bk.x * bk.gy + bk.y));
  }
}

template <typename T, int SM, int DEPTH, bool APV>
__device__ __forceinline__ void mvtail_lb(const typename KMvTail<T>::Args& a, const Blk& bk) {
  switch (a.LB) {
    case 1: mvtail_body<T, 1, SM, DEPTH, APV>(a, bk); break;
    case 2: mvtail_body<T, 2, SM, DEPTH, APV>(a, bk); break;
    case 3: mvtail_body<T, 3, SM, DEPTH, APV>(a, bk); break;
    case 4: mvtail_body<T, 4, SM, DEPTH, APV>(a, bk); break;
    case 5: mvtail_body<T, 5, SM, DEPTH, APV>(a, bk); break;
    case 6: mvtail_body<T, 6, SM, DEPTH, APV>(a, bk); break;
    case 7: mvtail_body<T, 7, SM, DEPTH, APV>(a, bk); break;
    case 8: mvtail_body<T, 8, SM, DEPTH, APV>(a, bk); break;
    default: mvtail_body<T, 9, SM, DEPTH, APV>(a, bk); break;
  }
}

// S groups of 2 in both precisions (the item order, hence the summation order, is the same wherever the
// kernel runs); K steps prefetched 2 deep in FP64 (244 VGPRs: 2 waves per SIMD) and 3 deep in FP32
// (164 VGPRs: 3 waves per SIMD).
// Tile masks of a permuted operator core Apm[(m, S)][k] (nm S rows of K values, K <= kTileMaskMaxK):
// out[mb S + S'] has bit j set when some row m of block mb (rows 16 mb .. 16 mb + 15 below nm) of S'
// holds a nonzero (or a NaN) at some k in [16 j, 16 j + 16). One wave per (mb, S'), lane j takes step j.
template <typename T>
struct KTileMask {
  struct Args {
    const T* Apm;
    uint64_t* out;
    int nm, S, K;
  };
  static constexpr int kThreads = 64;
  static __device__ void run(const Args& a, const Blk& bk) {
    const int j = threadIdx.x, mb = bk.x, sp = bk.y;
    bool nz = false;
    if (16 * j < a.K) {
      const int k1 = min(16 * j + 16, a.K);
      for (int m = 16 * mb; m < min(16 * mb + 16, a.nm); ++m) {
        const T* row = a.Apm + ((int64_t)m * a.S + sp) * a.K;
        for (int k = 16 * j; k < k1; ++k) nz = nz || !(row[k] == T(0));
      }
    }
    const uint64_t bits = __ballot(nz);
    if (j == 0) a.out[(int64_t)mb * a.S + sp] = bits;
  }
};

template <typename T>
__device__ __forceinline__ void KMvTail<T>::run(const Args& a, const Blk& bk) {
  if (a.skip && *a.skip) return;
  constexpr int D = std::is_same<T, double>::value ? 2 : 3;
  if (a.apv) mvtail_lb<T, 2, D, true>(a, bk);
  else mvtail_lb<T, 2, D, false>(a, bk);
}

}  // namespace ttamen
