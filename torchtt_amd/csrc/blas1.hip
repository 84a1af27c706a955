// Vector kernels of the AMEn sweep: deterministic norms, scaling by device
// scalars, copies/casts, the operator-core permutation and the dense local
// operator assembly used by the direct local solver.
//
// All of these are HBM-bound streaming kernels; they use grid-stride loops with
// a fixed grid so that reductions combine partials in a fixed order.
#include "ttamen_internal.h"
#include <algorithm>

namespace ttamen {

static inline int grid_for(int64_t n, int per = 256 * 4, int cap = 2048) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

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

// Last-block-finalises reduction of sum(x^2) (or sum((x-y)^2)).
template <typename T>
struct KSumsq {
  struct Args {
    const T* x; const T* y; int64_t n; double* partial; unsigned* counter; double* out;
    const double* den; int guard; double* mul; const int* skip;
  };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    if (p.skip && *p.skip) return;
    __shared__ double sh[4];
    __shared__ bool last;
    double acc = 0.0;
    for (int64_t i = bk.x * 256 + threadIdx.x; i < p.n; i += (int64_t)bk.gx * 256) {
      double v = (double)p.x[i];
      if (p.y) v -= (double)p.y[i];
      acc += v * v;
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
    if (threadIdx.x == 0) {
      double s = 0.0;
      volatile double* vp = p.partial;
      for (unsigned b = 0; b < bk.gx; ++b) s += vp[b];
      double r = sqrt(s);
      if (p.den) r = r / (*p.den);
      if (p.guard && !(r > 0.0)) r = 1.0;
      *p.out = r;
      if (p.mul) *p.mul = (*p.mul) * r;
      *p.counter = 0u;
    }
  }
};

void norm2(const double* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st, bool guard,
           double* mul_slot, const int* skip) {
  launch<KSumsq<double>>(dim3(grid_for(n)), dim3(256), 0, st,
                         {x, nullptr, n, rs.partial, rs.counter, out_slot, nullptr, guard ? 1 : 0,
                          mul_slot, skip});
  TT_LAUNCH_CHECK();
}

void norm2_f32(const float* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st,
               const int* skip) {
  launch<KSumsq<float>>(dim3(grid_for(n)), dim3(256), 0, st,
                        {x, nullptr, n, rs.partial, rs.counter, out_slot, nullptr, 0, nullptr, skip});
  TT_LAUNCH_CHECK();
}

void diffnorm(const double* x, const double* y, int64_t n, const double* den_slot, double* out_slot,
              RedScratch rs, hipStream_t st) {
  launch<KSumsq<double>>(dim3(grid_for(n)), dim3(256), 0, st,
                         {x, y, n, rs.partial, rs.counter, out_slot, den_slot, 0, nullptr, nullptr});
  TT_LAUNCH_CHECK();
}

// Grid-stride loop over [0, n) with the virtual block coordinates.
#define TT_GRID_LOOP(i, n)                                                                  \
  for (int64_t i = bk.x * (int64_t)blockDim.x + threadIdx.x; i < (n);                     \
       i += (int64_t)bk.gx * blockDim.x)

struct KScaleSlot {
  struct Args { double* x; int64_t n; const double* slot; int inv; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    double s = *p.slot;
    if (p.inv) s = 1.0 / s;
    TT_GRID_LOOP(i, p.n) p.x[i] *= s;
  }
};

void scale_inv(double* x, int64_t n, const double* slot, hipStream_t st) {
  if (n <= 0) return;
  launch<KScaleSlot>(dim3(grid_for(n)), dim3(256), 0, st, {x, n, slot, 1});
  TT_LAUNCH_CHECK();
}
void scale_by(double* x, int64_t n, const double* slot, hipStream_t st) {
  if (n <= 0) return;
  launch<KScaleSlot>(dim3(grid_for(n)), dim3(256), 0, st, {x, n, slot, 0});
  TT_LAUNCH_CHECK();
}

struct KAxpby {
  struct Args { double a; const double* x; double b; double* y; int64_t n; const int* skip; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    if (p.skip && *p.skip) return;
    TT_GRID_LOOP(i, p.n) p.y[i] = p.a * p.x[i] + p.b * p.y[i];
  }
};
void axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t st,
           const int* skip) {
  if (n <= 0) return;
  launch<KAxpby>(dim3(grid_for(n)), dim3(256), 0, st, {a, x, b, y, n, skip});
  TT_LAUNCH_CHECK();
}

struct KAxpbyF32 {
  struct Args { double a; const float* x; double b; float* y; int64_t n; const int* skip; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    if (p.skip && *p.skip) return;
    TT_GRID_LOOP(i, p.n) p.y[i] = (float)(p.a * (double)p.x[i] + p.b * (double)p.y[i]);
  }
};
void axpby_f32(double a, const float* x, double b, float* y, int64_t n, hipStream_t st,
               const int* skip) {
  if (n <= 0) return;
  launch<KAxpbyF32>(dim3(grid_for(n)), dim3(256), 0, st, {a, x, b, y, n, skip});
  TT_LAUNCH_CHECK();
}

// y (FP64) += x (FP32)
struct KAddF32F64 {
  struct Args { const float* x; double* y; int64_t n; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.y[i] += (double)p.x[i];
  }
};
void add_f32_to_f64(const float* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  launch<KAddF32F64>(dim3(grid_for(n)), dim3(256), 0, st, {x, y, n});
  TT_LAUNCH_CHECK();
}

struct KSub {
  struct Args { const double* x; const double* y; double* z; int64_t n; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.z[i] = p.x[i] - p.y[i];
  }
};
void sub(const double* x, const double* y, double* z, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  launch<KSub>(dim3(grid_for(n)), dim3(256), 0, st, {x, y, z, n});
  TT_LAUNCH_CHECK();
}

struct KCopy {
  struct Args { const double* x; double* y; int64_t n; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.y[i] = p.x[i];
  }
};
void copy(const double* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0 || x == y) return;
  launch<KCopy>(dim3(grid_for(n)), dim3(256), 0, st, {x, y, n});
  TT_LAUNCH_CHECK();
}

struct KFill {
  struct Args { double* x; int64_t n; double v; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.x[i] = p.v;
  }
};
void fill(double* x, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return;
  launch<KFill>(dim3(grid_for(n)), dim3(256), 0, st, {x, n, v});
  TT_LAUNCH_CHECK();
}

struct KZero32 {
  struct Args { uint32_t* x; int64_t n; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.x[i] = 0u;
  }
};
void zero_words(void* x, int64_t nwords, hipStream_t st) {
  if (nwords <= 0) return;
  launch<KZero32>(dim3(grid_for(nwords)), dim3(256), 0, st, {(uint32_t*)x, nwords});
  TT_LAUNCH_CHECK();
}

struct KEye {
  struct Args { double* x; int64_t m, n; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(t, p.m * p.n) {
      int64_t i = t / p.n, j = t - i * p.n;
      p.x[t] = (i == j) ? 1.0 : 0.0;
    }
  }
};
void eye(double* x, int64_t m, int64_t n, hipStream_t st) {
  if (m * n <= 0) return;
  launch<KEye>(dim3(grid_for(m * n)), dim3(256), 0, st, {x, m, n});
  TT_LAUNCH_CHECK();
}

struct KCopy2d {
  struct Args { const double* src; int64_t sr, sc; double* dst; int64_t ldd, m, n; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(t, p.m * p.n) {
      int64_t i = t / p.n, j = t - i * p.n;
      p.dst[i * p.ldd + j] = p.src[i * p.sr + j * p.sc];
    }
  }
};
void copy2d(const double* src, int64_t sr, int64_t sc, double* dst, int64_t ldd, int64_t m,
            int64_t n, hipStream_t st) {
  if (m * n <= 0) return;
  launch<KCopy2d>(dim3(grid_for(m * n)), dim3(256), 0, st, {src, sr, sc, dst, ldd, m, n});
  TT_LAUNCH_CHECK();
}

template <typename A, typename B>
struct KCast {
  struct Args { const A* x; B* y; int64_t n; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    TT_GRID_LOOP(i, p.n) p.y[i] = (B)p.x[i];
  }
};
void cast_f64_f32(const double* x, float* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  launch<KCast<double, float>>(dim3(grid_for(n)), dim3(256), 0, st, {x, y, n});
  TT_LAUNCH_CHECK();
}
void cast_f32_f64(const float* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  launch<KCast<float, double>>(dim3(grid_for(n)), dim3(256), 0, st, {x, y, n});
  TT_LAUNCH_CHECK();
}

// A[s][m][n][S] -> P[m][S][s][n]
template <typename T>
struct KPermA {
  struct Args { const double* A; T* P; int s, n, S; };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
    const int s = p.s, n = p.n, S = p.S;
    const int64_t total = (int64_t)s * n * n * S;
    TT_GRID_LOOP(t, total) {
      // t indexes P: ((m*S + Si)*s + si)*n + ni
      int64_t q = t;
      int ni = (int)(q % n); q /= n;
      int si = (int)(q % s); q /= s;
      int Si = (int)(q % S); q /= S;
      int mi = (int)q;
      p.P[t] = (T)p.A[(((int64_t)si * n + mi) * n + ni) * S + Si];
    }
  }
};
void permute_core_A(const double* A, double* P, int s, int n, int S, hipStream_t st) {
  int64_t total = (int64_t)s * n * n * S;
  launch<KPermA<double>>(dim3(grid_for(total)), dim3(256), 0, st, {A, P, s, n, S});
  TT_LAUNCH_CHECK();
}
void permute_core_A_f32(const double* A, float* P, int s, int n, int S, hipStream_t st) {
  int64_t total = (int64_t)s * n * n * S;
  launch<KPermA<float>>(dim3(grid_for(total)), dim3(256), 0, st, {A, P, s, n, S});
  TT_LAUNCH_CHECK();
}

// B[(l,m,L),(r,n,R)] = sum_{s,S} PhiL[l,s,r] A[s,m,n,S] PhiR[L,S,R]
struct KAssemble {
  struct Args {
    const double* PhiL; const double* A; const double* PhiR; double* B; int64_t ld;
    int r1, s, n, S, r2;
  };
  static constexpr int kThreads = 256;
  static __device__ void run(const Args& p, const Blk& bk) {
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
void assemble_local(const double* PhiL, const double* A, const double* PhiR, double* B, int64_t ld,
                    int r1, int s, int n, int S, int r2, hipStream_t st) {
  int64_t N = (int64_t)r1 * n * r2;
  launch<KAssemble>(dim3(grid_for(N * N, 256, 4096)), dim3(256), 0, st,
                    {PhiL, A, PhiR, B, ld, r1, s, n, S, r2});
  TT_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- scalar program
struct KNrmsc {
  struct Args { double* S; int in, ia, ix, ib, mode; };
  static constexpr int kThreads = 64;
  static __device__ void run(const Args& p, const Blk&) {
    if (threadIdx.x != 0) return;
    double* S = p.S;
    if (p.mode == 0) S[p.in] = S[p.in] * S[p.ia] * S[p.ix] / S[p.ib];
    else S[p.in] = S[p.in] * S[p.ib] / (S[p.ia] * S[p.ix]);
  }
};
void nrmsc_update(double* S, int i_nrmsc, int i_a, int i_x, int i_b, int mode, hipStream_t st) {
  launch<KNrmsc>(dim3(1), dim3(1), 0, st, {S, i_nrmsc, i_a, i_x, i_b, mode});
  TT_LAUNCH_CHECK();
}
struct KMaxdiv {
  struct Args { double* S; int dst, src, den; };
  static constexpr int kThreads = 64;
  static __device__ void run(const Args& p, const Blk&) {
    if (threadIdx.x != 0) return;
    double v = p.S[p.src] / p.S[p.den];
    if (!(p.S[p.dst] >= v)) p.S[p.dst] = v;
  }
};
void scalar_maxdiv(double* S, int dst, int src, int den, hipStream_t st) {
  launch<KMaxdiv>(dim3(1), dim3(1), 0, st, {S, dst, src, den});
  TT_LAUNCH_CHECK();
}
struct KSset {
  struct Args { double* S; int dst; double v; };
  static constexpr int kThreads = 64;
  static __device__ void run(const Args& p, const Blk&) {
    if (threadIdx.x == 0) p.S[p.dst] = p.v;
  }
};
void scalar_set(double* S, int dst, double v, hipStream_t st) {
  launch<KSset>(dim3(1), dim3(1), 0, st, {S, dst, v});
  TT_LAUNCH_CHECK();
}
// S[dst] = S[a] / S[b]
struct KSdiv {
  struct Args { double* S; int dst, a, b; };
  static constexpr int kThreads = 64;
  static __device__ void run(const Args& p, const Blk&) {
    if (threadIdx.x == 0) p.S[p.dst] = p.S[p.a] / p.S[p.b];
  }
};
void scalar_div(double* S, int dst, int a, int b, hipStream_t st) {
  launch<KSdiv>(dim3(1), dim3(1), 0, st, {S, dst, a, b});
  TT_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- host transfers
// Device -> pinned host copy done by a kernel (used inside a cohort, where the
// copies of many systems become one launch).
struct KStoreHost {
  struct Args { const uint32_t* src; uint32_t* dst; int64_t nwords; };
  static constexpr int kThreads = 64;
  static __device__ void run(const Args& p, const Blk&) {
    for (int64_t i = threadIdx.x; i < p.nwords; i += blockDim.x) p.dst[i] = p.src[i];
  }
};

void post_to_host(void* host, const void* dev, size_t bytes, hipStream_t st) {
  if (bytes == 0) return;
  if (current_recorder()) {
    TT_CHECK(bytes % 4 == 0, ERR_INVALID, "post_to_host: size must be a multiple of 4");
    launch<KStoreHost>(dim3(1), dim3(64), 0, st,
                       {(const uint32_t*)dev, (uint32_t*)host, (int64_t)(bytes / 4)});
    return;
  }
  TT_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
}

void sync_point(hipStream_t st) {
  if (Recorder* r = current_recorder()) {
    r->sync();
    return;
  }
  TT_HIP(hipStreamSynchronize(st));
}

void read_to_host(void* host, const void* dev, size_t bytes, hipStream_t st) {
  post_to_host(host, dev, bytes, st);
  sync_point(st);
}

}  // namespace ttamen
