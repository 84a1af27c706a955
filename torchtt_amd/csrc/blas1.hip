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
__global__ __launch_bounds__(256) void k_sumsq(const T* x, const T* y, int64_t n, double* partial,
                                               unsigned* counter, double* out, const double* den,
                                               int guard, double* mul, const int* skip) {
  if (skip && *skip) return;
  __shared__ double sh[4];
  __shared__ bool last;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double v = (double)x[i];
    if (y) v -= (double)y[i];
    acc += v * v;
  }
  double bs = block_sum256(acc, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = bs;
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (threadIdx.x == 0) {
    double s = 0.0;
    volatile double* vp = partial;
    for (unsigned b = 0; b < gridDim.x; ++b) s += vp[b];
    double r = sqrt(s);
    if (den) r = r / (*den);
    if (guard && !(r > 0.0)) r = 1.0;
    *out = r;
    if (mul) *mul = (*mul) * r;
    *counter = 0u;
  }
}

void norm2(const double* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st, bool guard,
           double* mul_slot, const int* skip) {
  int g = grid_for(n);
  hipLaunchKernelGGL((k_sumsq<double>), dim3(g), dim3(256), 0, st, x, (const double*)nullptr, n,
                     rs.partial, rs.counter, out_slot, (const double*)nullptr, guard ? 1 : 0,
                     mul_slot, skip);
  TT_LAUNCH_CHECK();
}

void norm2_f32(const float* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st,
               const int* skip) {
  int g = grid_for(n);
  hipLaunchKernelGGL((k_sumsq<float>), dim3(g), dim3(256), 0, st, x, (const float*)nullptr, n,
                     rs.partial, rs.counter, out_slot, (const double*)nullptr, 0,
                     (double*)nullptr, skip);
  TT_LAUNCH_CHECK();
}

void diffnorm(const double* x, const double* y, int64_t n, const double* den_slot, double* out_slot,
              RedScratch rs, hipStream_t st) {
  int g = grid_for(n);
  hipLaunchKernelGGL((k_sumsq<double>), dim3(g), dim3(256), 0, st, x, y, n, rs.partial,
                     rs.counter, out_slot, den_slot, 0, (double*)nullptr, (const int*)nullptr);
  TT_LAUNCH_CHECK();
}

__global__ void k_scale_slot(double* x, int64_t n, const double* slot, int inv) {
  double s = *slot;
  if (inv) s = 1.0 / s;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= s;
}

void scale_inv(double* x, int64_t n, const double* slot, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_slot, dim3(grid_for(n)), dim3(256), 0, st, x, n, slot, 1);
  TT_LAUNCH_CHECK();
}
void scale_by(double* x, int64_t n, const double* slot, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_slot, dim3(grid_for(n)), dim3(256), 0, st, x, n, slot, 0);
  TT_LAUNCH_CHECK();
}

__global__ void k_axpby(double a, const double* x, double b, double* y, int64_t n,
                        const int* skip) {
  if (skip && *skip) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
void axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t st,
           const int* skip) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, st, a, x, b, y, n, skip);
  TT_LAUNCH_CHECK();
}

__global__ void k_axpby_f32(double a, const float* x, double b, float* y, int64_t n,
                            const int* skip) {
  if (skip && *skip) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)(a * (double)x[i] + b * (double)y[i]);
}
void axpby_f32(double a, const float* x, double b, float* y, int64_t n, hipStream_t st,
               const int* skip) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby_f32, dim3(grid_for(n)), dim3(256), 0, st, a, x, b, y, n, skip);
  TT_LAUNCH_CHECK();
}

// y (FP64) += x (FP32)
__global__ void k_add_f32_to_f64(const float* x, double* y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += (double)x[i];
}
void add_f32_to_f64(const float* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add_f32_to_f64, dim3(grid_for(n)), dim3(256), 0, st, x, y, n);
  TT_LAUNCH_CHECK();
}

__global__ void k_sub(const double* x, const double* y, double* z, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    z[i] = x[i] - y[i];
}
void sub(const double* x, const double* y, double* z, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, st, x, y, z, n);
  TT_LAUNCH_CHECK();
}

__global__ void k_copy(const double* x, double* y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
void copy(const double* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0 || x == y) return;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, st, x, y, n);
  TT_LAUNCH_CHECK();
}

__global__ void k_fill(double* x, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
void fill(double* x, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, st, x, n, v);
  TT_LAUNCH_CHECK();
}

__global__ void k_eye(double* x, int64_t m, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m * n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = t / n, j = t - i * n;
    x[t] = (i == j) ? 1.0 : 0.0;
  }
}
void eye(double* x, int64_t m, int64_t n, hipStream_t st) {
  if (m * n <= 0) return;
  hipLaunchKernelGGL(k_eye, dim3(grid_for(m * n)), dim3(256), 0, st, x, m, n);
  TT_LAUNCH_CHECK();
}

__global__ void k_copy2d(const double* src, int64_t sr, int64_t sc, double* dst, int64_t ldd,
                         int64_t m, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m * n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = t / n, j = t - i * n;
    dst[i * ldd + j] = src[i * sr + j * sc];
  }
}
void copy2d(const double* src, int64_t sr, int64_t sc, double* dst, int64_t ldd, int64_t m,
            int64_t n, hipStream_t st) {
  if (m * n <= 0) return;
  hipLaunchKernelGGL(k_copy2d, dim3(grid_for(m * n)), dim3(256), 0, st, src, sr, sc, dst, ldd, m,
                     n);
  TT_LAUNCH_CHECK();
}

template <typename A, typename B>
__global__ void k_cast(const A* x, B* y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (B)x[i];
}
void cast_f64_f32(const double* x, float* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(n)), dim3(256), 0, st, x, y, n);
  TT_LAUNCH_CHECK();
}
void cast_f32_f64(const float* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(n)), dim3(256), 0, st, x, y, n);
  TT_LAUNCH_CHECK();
}

// A[s][m][n][S] -> P[m][S][s][n]
template <typename T>
__global__ void k_permA(const double* A, T* P, int s, int n, int S) {
  int64_t total = (int64_t)s * n * n * S;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    // t indexes P: ((m*S + Si)*s + si)*n + ni
    int64_t q = t;
    int ni = (int)(q % n); q /= n;
    int si = (int)(q % s); q /= s;
    int Si = (int)(q % S); q /= S;
    int mi = (int)q;
    P[t] = (T)A[(((int64_t)si * n + mi) * n + ni) * S + Si];
  }
}
void permute_core_A(const double* A, double* P, int s, int n, int S, hipStream_t st) {
  int64_t total = (int64_t)s * n * n * S;
  hipLaunchKernelGGL((k_permA<double>), dim3(grid_for(total)), dim3(256), 0, st, A, P, s, n, S);
  TT_LAUNCH_CHECK();
}
void permute_core_A_f32(const double* A, float* P, int s, int n, int S, hipStream_t st) {
  int64_t total = (int64_t)s * n * n * S;
  hipLaunchKernelGGL((k_permA<float>), dim3(grid_for(total)), dim3(256), 0, st, A, P, s, n, S);
  TT_LAUNCH_CHECK();
}

// B[(l,m,L),(r,n,R)] = sum_{s,S} PhiL[l,s,r] A[s,m,n,S] PhiR[L,S,R]
__global__ void k_assemble(const double* PhiL, const double* A, const double* PhiR, double* B,
                           int64_t ld, int r1, int s, int n, int S, int r2) {
  const int64_t N = (int64_t)r1 * n * r2;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N * N;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t row = t / N, col = t - row * N;
    int L = (int)(row % r2); int64_t q = row / r2; int m = (int)(q % n); int l = (int)(q / n);
    int R = (int)(col % r2); q = col / r2; int nn = (int)(q % n); int r = (int)(q / n);
    double acc = 0.0;
    for (int si = 0; si < s; ++si) {
      double pl = PhiL[((int64_t)l * s + si) * r1 + r];
      if (pl == 0.0) continue;
      for (int Si = 0; Si < S; ++Si)
        acc += pl * A[(((int64_t)si * n + m) * n + nn) * S + Si] * PhiR[((int64_t)L * S + Si) * r2 + R];
    }
    B[row * ld + col] = acc;
  }
}
void assemble_local(const double* PhiL, const double* A, const double* PhiR, double* B, int64_t ld,
                    int r1, int s, int n, int S, int r2, hipStream_t st) {
  int64_t N = (int64_t)r1 * n * r2;
  hipLaunchKernelGGL(k_assemble, dim3(grid_for(N * N, 256, 4096)), dim3(256), 0, st, PhiL, A,
                     PhiR, B, ld, r1, s, n, S, r2);
  TT_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- scalar program
__global__ void k_nrmsc(double* S, int in, int ia, int ix, int ib, int mode) {
  if (mode == 0) S[in] = S[in] * S[ia] * S[ix] / S[ib];
  else S[in] = S[in] * S[ib] / (S[ia] * S[ix]);
}
void nrmsc_update(double* S, int i_nrmsc, int i_a, int i_x, int i_b, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_nrmsc, dim3(1), dim3(1), 0, st, S, i_nrmsc, i_a, i_x, i_b, mode);
  TT_LAUNCH_CHECK();
}
__global__ void k_maxdiv(double* S, int dst, int src, int den) {
  double v = S[src] / S[den];
  if (!(S[dst] >= v)) S[dst] = v;
}
void scalar_maxdiv(double* S, int dst, int src, int den, hipStream_t st) {
  hipLaunchKernelGGL(k_maxdiv, dim3(1), dim3(1), 0, st, S, dst, src, den);
  TT_LAUNCH_CHECK();
}
__global__ void k_sset(double* S, int dst, double v) { S[dst] = v; }
void scalar_set(double* S, int dst, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_sset, dim3(1), dim3(1), 0, st, S, dst, v);
  TT_LAUNCH_CHECK();
}

}  // namespace ttamen
