// Streaming / reduction / scalar kernel bodies (launch.h form) of the AMEn sweep; the
// host wrappers are in blas1.hip.
#pragma once
#include "ttamen_internal.h"
#include <type_traits>

namespace ttamen {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-level sum of 256 threads; result valid in thread 0.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

// Grid-stride loop over [0, n) with the virtual block coordinates.
#define TT_GRID_LOOP(i, n)                                                                  \
  for (int64_t i = bk.x * (int64_t)blockDim.x + threadIdx.x; i < (n);                     \
       i += (int64_t)bk.gx * blockDim.x)

// Last-block-finalises reduction of sum(x^2) (or sum((x-y)^2)).
template <typename T>
struct KSumsq {
  struct Args {
    const T* x; const T* y; int64_t n; double* partial; unsigned* counter; double* out;
    const double* den; int guard; double* mul; const int* skip;
    double* scale_x;  // if set: the last block multiplies scale_x[0 .. n) by 1 / result (a small vector)
  };
  static constexpr int kThreads = 256;
  static constexpr int kMega = std::is_same<T, double>::value ? 17 : 18;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    if (p.skip && *p.skip) return;
    __shared__ double sh[4];
    __shared__ bool last;
    __shared__ double parts[256];
    double acc = 0.0;
    // Grid-stride order, four strides' loads in flight at once.
    const int64_t stride = (int64_t)bk.gx * 256;
    for (int64_t i0 = bk.x * 256 + threadIdx.x; i0 < p.n; i0 += 4 * stride) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t i = i0 + u * stride;
        v[u] = 0.0;
        if (i < p.n) {
          v[u] = (double)p.x[i];
          if (p.y) v[u] -= (double)p.y[i];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i0 + u * stride < p.n) acc += v[u] * v[u];
    }
    double bs = block_sum256(acc, sh);
    if (threadIdx.x == 0) {
      p.partial[bk.x] = bs;
      __threadfence();
      unsigned prev = atomicAdd(p.counter, 1u);
      last = (prev == bk.gx - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // The last block: the threads fetch 256 partials at a time, thread 0 adds them in block order.
    volatile double* vp = p.partial;
    double s = 0.0;
    for (unsigned b0 = 0; b0 < bk.gx; b0 += 256) {
      if (b0 + threadIdx.x < bk.gx) parts[threadIdx.x] = vp[b0 + threadIdx.x];
      __syncthreads();
      if (threadIdx.x == 0) {
        const unsigned nb = min(256u, bk.gx - b0);
        for (unsigned b = 0; b < nb; ++b) s += parts[b];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double r = sqrt(s);
      if (p.den) r = r / (*p.den);
      if (p.guard && !(r > 0.0)) r = 1.0;
      *p.out = r;
      if (p.mul) *p.mul = (*p.mul) * r;
      *p.counter = 0u;
      sh[0] = r;
    }
    if (p.scale_x) {
      __syncthreads();
      const double inv = 1.0 / sh[0];
      for (int64_t i = threadIdx.x; i < p.n; i += 256) p.scale_x[i] *= inv;
    }
  }
};

struct KScaleSlot {
  struct Args { double* x; int64_t n; const double* slot; int inv; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 19;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    double s = *p.slot;
    if (p.inv) s = 1.0 / s;
    TT_GRID_LOOP(i, p.n) p.x[i] *= s;
  }
};

struct KAxpby {
  struct Args { double a; const double* x; double b; double* y; int64_t n; const int* skip; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 20;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    if (p.skip && *p.skip) return;
    TT_GRID_LOOP(i, p.n) p.y[i] = p.a * p.x[i] + p.b * p.y[i];
  }
};

struct KAxpbyF32 {
  struct Args { double a; const float* x; double b; float* y; int64_t n; const int* skip; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 21;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    if (p.skip && *p.skip) return;
    TT_GRID_LOOP(i, p.n) p.y[i] = (float)(p.a * (double)p.x[i] + p.b * (double)p.y[i]);
  }
};

// y (FP64) += x (FP32)
// y = (base ? base : y) + x, x in FP32
struct KAddF32F64 {
  struct Args { const float* x; const double* base; double* y; int64_t n; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 22;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    if (p.base) {
      TT_GRID_LOOP(i, p.n) p.y[i] = p.base[i] + (double)p.x[i];
    } else {
      TT_GRID_LOOP(i, p.n) p.y[i] += (double)p.x[i];
    }
  }
};

struct KSub {
  struct Args { const double* x; const double* y; double* z; int64_t n; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 23;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.z[i] = p.x[i] - p.y[i];
  }
};

struct KCopy {
  struct Args { const double* x; double* y; int64_t n; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 24;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.y[i] = p.x[i];
  }
};

struct KFill {
  struct Args { double* x; int64_t n; double v; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 25;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.x[i] = p.v;
  }
};

struct KZero32 {
  struct Args { uint32_t* x; int64_t n; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 26;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.x[i] = 0u;
  }
};

struct KEye {
  struct Args { double* x; int64_t m, n; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 27;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(t, p.m * p.n) {
      int64_t i = t / p.n, j = t - i * p.n;
      p.x[t] = (i == j) ? 1.0 : 0.0;
    }
  }
};

struct KCopy2d {
  struct Args { const double* src; int64_t sr, sc; double* dst; int64_t ldd, m, n; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 28;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(t, p.m * p.n) {
      int64_t i = t / p.n, j = t - i * p.n;
      p.dst[i * p.ldd + j] = p.src[i * p.sr + j * p.sc];
    }
  }
};

template <typename A, typename B>
struct KCast {
  struct Args { const A* x; B* y; int64_t n; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = std::is_same<A, double>::value ? 29 : 30;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.y[i] = (B)p.x[i];
  }
};

template <typename T>
struct KPermA {
  struct Args { const double* A; T* P; int s, nm, nk, S; };  // A[s][nm][nk][S] -> P[nm][S][s][nk]
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    const int s = p.s, nm = p.nm, nk = p.nk, S = p.S;
    const int64_t total = (int64_t)s * nm * nk * S;
    TT_GRID_LOOP(t, total) {
      // t indexes P: ((m*S + Si)*s + si)*nk + ki
      int64_t q = t;
      int ki = (int)(q % nk); q /= nk;
      int si = (int)(q % s); q /= s;
      int Si = (int)(q % S); q /= S;
      int mi = (int)q;
      p.P[t] = (T)p.A[(((int64_t)si * nm + mi) * nk + ki) * S + Si];
    }
  }
};

// B[(l,m,L),(r,n,R)] = sum_{s,S} PhiL[l,s,r] A[s,m,n,S] PhiR[L,S,R]
struct KAssemble {
  struct Args {
    const double* PhiL; const double* A; const double* PhiR; double* B; int64_t ld;
    int r1, s, n, S, r2;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    const int r1 = p.r1, s = p.s, n = p.n, S = p.S, r2 = p.r2;
    const int64_t N = (int64_t)r1 * n * r2;
    TT_GRID_LOOP(t, N * N) {
      int64_t row = t / N, col = t - row * N;
      int L = (int)(row % r2); int64_t q = row / r2; int m = (int)(q % n); int l = (int)(q / n);
      int R = (int)(col % r2); q = col / r2; int nn = (int)(q % n); int r = (int)(q / n);
      double acc = 0.0;
      for (int si = 0; si < s; ++si) {
        double pl = p.PhiL[((int64_t)l * s + si) * r1 + r];
        if (pl == 0.0) continue;
        for (int Si = 0; Si < S; ++Si)
          acc += pl * p.A[(((int64_t)si * n + m) * n + nn) * S + Si] *
                 p.PhiR[((int64_t)L * S + Si) * r2 + R];
      }
      p.B[row * p.ld + col] = acc;
    }
  }
};

// ---------------------------------------------------------------- scalar program
struct KNrmsc {
  struct Args { double* S; int in, ia, ix, ib, mode; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 31;
  static __device__ __forceinline__ void run(const Args& p, const Blk&) {
    if (threadIdx.x != 0) return;
    double* S = p.S;
    if (p.mode == 0) S[p.in] = S[p.in] * S[p.ia] * S[p.ix] / S[p.ib];
    else S[p.in] = S[p.in] * S[p.ib] / (S[p.ia] * S[p.ix]);
  }
};

struct KMaxdiv {
  struct Args { double* S; int dst, src, den; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 32;
  static __device__ __forceinline__ void run(const Args& p, const Blk&) {
    if (threadIdx.x != 0) return;
    double v = p.S[p.src] / p.S[p.den];
    if (!(p.S[p.dst] >= v)) p.S[p.dst] = v;
  }
};

struct KSset {
  struct Args { double* S; int dst; double v; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 33;
  static __device__ __forceinline__ void run(const Args& p, const Blk&) {
    if (threadIdx.x == 0) p.S[p.dst] = p.v;
  }
};

// S[dst] = S[a] / S[b]
struct KSdiv {
  struct Args { double* S; int dst, a, b; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 34;
  static __device__ __forceinline__ void run(const Args& p, const Blk&) {
    if (threadIdx.x == 0) p.S[p.dst] = p.S[p.a] / p.S[p.b];
  }
};

// Inner tolerance of a mixed-precision refinement step, times ||r32||:
//   S[dst] = min(0.5, max(floor, real_tol * S[nrhs] / S[rn])) * S[r32]
// (the step's GMRES aims at the target itself; the FP64 residual decides whether another step runs)
struct KTauScale {
  struct Args { double* S; int dst, nrhs, rn, r32; double real_tol, floor; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 38;
  static __device__ __forceinline__ void run(const Args& p, const Blk&) {
    if (threadIdx.x != 0) return;
    const double t = fmin(0.5, fmax(p.floor, p.real_tol * p.S[p.nrhs] / p.S[p.rn]));
    p.S[p.dst] = t * p.S[p.r32];
  }
};

// ---------------------------------------------------------------- host transfers
// Device -> pinned host copy done by a kernel (used inside a cohort, where the
// copies of many systems become one launch).
struct KStoreHost {
  struct Args { const uint32_t* src; uint32_t* dst; int64_t nwords; };
  static constexpr int kThreads = 64;
  static constexpr int kMega = 35;
  static __device__ __forceinline__ void run(const Args& p, const Blk&) {
    for (int64_t i = threadIdx.x; i < p.nwords; i += blockDim.x) p.dst[i] = p.src[i];
  }
};

// out[i][c] = in[i][c] * s[c]  for c < ncols (in row-major with ld ldi, out ld ldo)
struct KScaleCols {
  struct Args { const double* in; int64_t ldi; const double* s; double* out; int64_t ldo, rows; int ncols; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 36;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk) {
    for (int64_t t = bk.x * (int64_t)blockDim.x + threadIdx.x; t < a.rows * a.ncols;
         t += (int64_t)bk.gx * blockDim.x) {
      int64_t i = t / a.ncols;
      int c = (int)(t - i * a.ncols);
      a.out[i * a.ldo + c] = a.in[i * a.ldi + c] * a.s[c];
    }
  }
};

// Normal random numbers: SplitMix64 counter stream + Box-Muller.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
struct KRandn {
  struct Args { double* x; int64_t n; uint64_t seed, stream_id; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = 37;
  static __device__ __forceinline__ void run(const Args& p, const Blk& bk) {
    for (int64_t i = bk.x * (int64_t)blockDim.x + threadIdx.x; i < p.n;
         i += (int64_t)bk.gx * blockDim.x) {
      uint64_t a = splitmix64(p.seed ^ splitmix64(p.stream_id * 0x100000001B3ull + (uint64_t)i * 2));
      uint64_t b = splitmix64(a + 0x632BE59BD9B4E019ull);
      double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);
      double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
      p.x[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
  }
};

}  // namespace ttamen
