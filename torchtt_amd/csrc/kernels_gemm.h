// GEMM kernel bodies (launch.h form) on CDNA4 matrix cores; see gemm.hip for the design.
#pragma once
#include "ttamen_internal.h"
#include <type_traits>

namespace ttamen {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct AccT;
template <> struct AccT<double> { typedef d4 type; };
template <> struct AccT<float> { typedef f4 type; };

// Row (within a 16x16 block) of result register v in a lane with lane/16 == lk.
template <typename T> __device__ __forceinline__ int acc_row(int lk, int v);
template <> __device__ __forceinline__ int acc_row<double>(int lk, int v) { return lk + 4 * v; }
template <> __device__ __forceinline__ int acc_row<float>(int lk, int v) { return 4 * lk + v; }

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T>
struct GemmParams {
  int M, N, K, B2, B;
  int64_t kchunk;
  const T* a;
  int64_t ar, ac, ab1, ab2;
  const T* b;
  int64_t br, bc, bb1, bb2;
  T* c;
  int64_t cr, cc, cb1, cb2;
  double alpha, beta;
  const double* scale_ptr;
  int scale_mode;
  const int* skip;
  T* partial;  // split-K scratch: [split][batch][M][N]
  int tile;    // 0: 32x32, 1: 32x64, 2: 64x32, 3: 64x64 (BM x BN); +4: BF16 MFMA (FP32 operands)
};

template <typename T>
__device__ __forceinline__ double eff_alpha(const GemmParams<T>& p) {
  double s = p.alpha;
  if (p.scale_mode == 1) s *= *p.scale_ptr;
  else if (p.scale_mode == 2) s /= *p.scale_ptr;
  return s;
}

// One tile shape; the LDS for the double-buffered A and B tiles is passed in.
template <typename T, int BM, int BN, int BK>
struct GemmTile {
  static constexpr int PA = 1, PB = 1;
  static constexpr size_t kLds = sizeof(T) * 2 * BK * ((BM + PA) + (BN + PB));
  static __device__ __forceinline__ void body(const GemmParams<T>& p, const Blk& bk, char* smem);
};

// All tile shapes are one launch kind (the tile is chosen per launch by p.tile), so that
// a cohort merges the GEMMs of systems whose shapes pick different tiles.
// LDS of the GEMM tiles: one buffer for both precisions (sized for FP64 64x64), so a
// kernel that can run either GEMM holds it once.
__device__ __forceinline__ char* gemm_lds() {
  __shared__ __attribute__((aligned(16))) char smem[2 * 16 * (65 + 65) * sizeof(double)];
  return smem;
}

// FP32 operands, BF16 arithmetic: each element is rounded to BF16 (round to nearest even)
// when it is staged into LDS, the products run on gfx950's v_mfma_f32_16x16x32_bf16 and accumulate
// in FP32. Fragment layout (16x16xK32 block): lane l holds A[i = l%16][k = 8*(l/16) + j] and
// B[k = 8*(l/16) + j][n = l%16], j = 0..7; result register v holds D[4*(l/16) + v][l%16].
// The LDS tiles are stored k-contiguous (A as [m][k], B as [n][k]) so that a lane's eight
// k values are one 16-byte read.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template <int BM, int BN>
struct GemmTileBf16 {
  static constexpr int BK = 32, PK = 8;  // row of BK + PK bf16 = 80 bytes (16-byte aligned reads)
  static constexpr size_t kLds = sizeof(unsigned short) * 2 * (BM + BN) * (BK + PK);
  static __device__ __forceinline__ void body(const GemmParams<float>& p, const Blk& bk, char* smem);
};

template <typename T>
struct KGemm {
  typedef GemmParams<T> Args;
  static constexpr int kThreads = 256;
  static constexpr int kMega = std::is_same<T, double>::value ? 0 : 1;
  static constexpr int kMegaGroup = 1;
  static __device__ __forceinline__ void run(const Args& pp, const Blk& bk) {
    static_assert(GemmTile<T, 64, 64, 16>::kLds <= 2 * 16 * (65 + 65) * sizeof(double), "gemm LDS");
    static_assert(GemmTileBf16<64, 64>::kLds <= 2 * 16 * (65 + 65) * sizeof(double), "gemm LDS");
    char* smem = gemm_lds();
    const GemmParams<T> p = pp;
    if (p.skip && *p.skip) return;
    if constexpr (std::is_same<T, float>::value) {
      if (p.tile >= 4) {
        switch (p.tile) {
          case 4: GemmTileBf16<32, 32>::body(p, bk, smem); break;
          case 5: GemmTileBf16<32, 64>::body(p, bk, smem); break;
          case 6: GemmTileBf16<64, 32>::body(p, bk, smem); break;
          default: GemmTileBf16<64, 64>::body(p, bk, smem); break;
        }
        return;
      }
    }
    switch (p.tile) {
      case 0: GemmTile<T, 32, 32, 16>::body(p, bk, smem); break;
      case 1: GemmTile<T, 32, 64, 16>::body(p, bk, smem); break;
      case 2: GemmTile<T, 64, 32, 16>::body(p, bk, smem); break;
      default: GemmTile<T, 64, 64, 16>::body(p, bk, smem); break;
    }
  }
};

template <typename T, int BM, int BN, int BK>
__device__ __forceinline__ void GemmTile<T, BM, BN, BK>::body(const GemmParams<T>& p, const Blk& bk, char* smem) {
  constexpr int MI = BM / 32, NI = BN / 32;  // 16x16 blocks per wave
  constexpr int AE = BM * BK / 256, BE = BK * BN / 256;
  T (*As)[BK][BM + PA] = reinterpret_cast<T (*)[BK][BM + PA]>(smem);
  T (*Bs)[BK][BN + PB] = reinterpret_cast<T (*)[BK][BN + PB]>(smem + sizeof(T) * 2 * BK * (BM + PA));
  typedef typename AccT<T>::type acc_t;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tiles_n = (p.N + BN - 1) / BN;
  const int tm = bk.x / tiles_n, tn = bk.x - tm * tiles_n;
  const int bidx = bk.y;
  const int b1 = bidx / p.B2, b2 = bidx - b1 * p.B2;
  const int64_t k0 = (int64_t)bk.z * p.kchunk;
  const int64_t k1 = min((int64_t)p.K, k0 + p.kchunk);
  const T* A = p.a + b1 * p.ab1 + b2 * p.ab2;
  const T* B = p.b + b1 * p.bb1 + b2 * p.bb2;
  const int m0 = tm * BM, n0 = tn * BN;
  const bool a_kfast = (p.ac == 1);
  const bool b_nfast = (p.bc == 1);

  T ra[AE], rb[BE];
  // The loop-invariant part of every element this thread stages: its address at k = 0, its
  // k offset within the tile, whether its row (column) exists, and its LDS slot. A K step
  // then adds one uniform offset (kt times the K stride) to each address.
  const T* a_at[AE];
  const T* b_at[BE];
  int a_k[AE], b_k[BE], a_slot[AE], b_slot[BE];
  bool a_in[AE], b_in[BE];
#pragma unroll
  for (int e = 0; e < AE; ++e) {
    int idx = tid + 256 * e, i, kk;
    if (a_kfast) { i = idx / BK; kk = idx - i * BK; }
    else { kk = idx / BM; i = idx - kk * BM; }
    a_in[e] = m0 + i < p.M;
    a_k[e] = kk;
    a_slot[e] = kk * (BM + PA) + i;
    a_at[e] = A + (a_in[e] ? (int64_t)(m0 + i) * p.ar : 0) + (int64_t)kk * p.ac;
  }
#pragma unroll
  for (int e = 0; e < BE; ++e) {
    int idx = tid + 256 * e, j, kk;
    if (b_nfast) { kk = idx / BN; j = idx - kk * BN; }
    else { j = idx / BK; kk = idx - j * BK; }
    b_in[e] = n0 + j < p.N;
    b_k[e] = kk;
    b_slot[e] = kk * (BN + PB) + j;
    b_at[e] = B + (b_in[e] ? (int64_t)(n0 + j) * p.bc : 0) + (int64_t)kk * p.br;
  }
  auto loadA = [&](int64_t kt) {
    const int64_t off = kt * p.ac;
#pragma unroll
    for (int e = 0; e < AE; ++e) ra[e] = (a_in[e] && kt + a_k[e] < k1) ? a_at[e][off] : T(0);
  };
  auto storeA = [&](int buf) {
#pragma unroll
    for (int e = 0; e < AE; ++e) (&As[buf][0][0])[a_slot[e]] = ra[e];
  };
  auto loadB = [&](int64_t kt) {
    const int64_t off = kt * p.br;
#pragma unroll
    for (int e = 0; e < BE; ++e) rb[e] = (b_in[e] && kt + b_k[e] < k1) ? b_at[e][off] : T(0);
  };
  auto storeB = [&](int buf) {
#pragma unroll
    for (int e = 0; e < BE; ++e) (&Bs[buf][0][0])[b_slot[e]] = rb[e];
  };

  acc_t acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = acc_t{0, 0, 0, 0};

  if (k0 < k1) {
    loadA(k0);
    loadB(k0);
    storeA(0);
    storeB(0);
  }
  __syncthreads();
  int cur = 0;
  const int lr = lane & 15, lk = lane >> 4;
  for (int64_t kt = k0; kt < k1; kt += BK) {
    const bool more = kt + BK < k1;
    if (more) {
      loadA(kt + BK);
      loadB(kt + BK);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[MI], bf[NI];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) af[mi] = As[cur][kk + lk][wm * (BM / 2) + mi * 16 + lr];
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) bf[ni] = Bs[cur][kk + lk][wn * (BN / 2) + ni * 16 + lr];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = mfma16(af[mi], bf[ni], acc[mi][ni]);
    }
    if (more) {
      storeA(cur ^ 1);
      storeB(cur ^ 1);
    }
    __syncthreads();
    cur ^= 1;
  }

  if (p.partial) {
    T* P = p.partial + ((int64_t)bk.z * p.B + bidx) * (int64_t)p.M * p.N;
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < NI; ++ni)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          int i = m0 + wm * (BM / 2) + mi * 16 + acc_row<T>(lk, v);
          int j = n0 + wn * (BN / 2) + ni * 16 + lr;
          if (i < p.M && j < p.N) P[(int64_t)i * p.N + j] = acc[mi][ni][v];
        }
    return;
  }
  const double al = eff_alpha(p);
  T* C = p.c + b1 * p.cb1 + b2 * p.cb2;
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int i = m0 + wm * (BM / 2) + mi * 16 + acc_row<T>(lk, v);
        int j = n0 + wn * (BN / 2) + ni * 16 + lr;
        if (i < p.M && j < p.N) {
          T* cp = C + (int64_t)i * p.cr + (int64_t)j * p.cc;
          double val = al * (double)acc[mi][ni][v];
          if (p.beta != 0.0) val += p.beta * (double)(*cp);
          *cp = (T)val;
        }
      }
}

template <int BM, int BN>
__device__ __forceinline__ void GemmTileBf16<BM, BN>::body(const GemmParams<float>& p, const Blk& bk,
                                                          char* smem) {
  constexpr int MI = BM / 32, NI = BN / 32;
  constexpr int AE = BM * BK / 256, BE = BK * BN / 256;
  typedef unsigned short bf;
  bf (*As)[BM][BK + PK] = reinterpret_cast<bf (*)[BM][BK + PK]>(smem);
  bf (*Bs)[BN][BK + PK] = reinterpret_cast<bf (*)[BN][BK + PK]>(smem + sizeof(bf) * 2 * BM * (BK + PK));

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tiles_n = (p.N + BN - 1) / BN;
  const int tm = bk.x / tiles_n, tn = bk.x - tm * tiles_n;
  const int bidx = bk.y;
  const int b1 = bidx / p.B2, b2 = bidx - b1 * p.B2;
  const int64_t k0 = (int64_t)bk.z * p.kchunk;
  const int64_t k1 = min((int64_t)p.K, k0 + p.kchunk);
  const float* A = p.a + b1 * p.ab1 + b2 * p.ab2;
  const float* B = p.b + b1 * p.bb1 + b2 * p.bb2;
  const int m0 = tm * BM, n0 = tn * BN;
  const bool a_kfast = (p.ac == 1);
  const bool b_nfast = (p.bc == 1);

  float ra[AE], rb[BE];
  auto loadA = [&](int64_t kt) {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      int idx = tid + 256 * e, i, kk;
      if (a_kfast) { i = idx / BK; kk = idx - i * BK; }
      else { kk = idx / BM; i = idx - kk * BM; }
      int64_t gi = m0 + i, gk = kt + kk;
      ra[e] = (gi < p.M && gk < k1) ? A[gi * p.ar + gk * p.ac] : 0.0f;
    }
  };
  auto storeA = [&](int buf) {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      int idx = tid + 256 * e, i, kk;
      if (a_kfast) { i = idx / BK; kk = idx - i * BK; }
      else { kk = idx / BM; i = idx - kk * BM; }
      As[buf][i][kk] = bf16_rne(ra[e]);
    }
  };
  auto loadB = [&](int64_t kt) {
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      int idx = tid + 256 * e, j, kk;
      if (b_nfast) { kk = idx / BN; j = idx - kk * BN; }
      else { j = idx / BK; kk = idx - j * BK; }
      int64_t gj = n0 + j, gk = kt + kk;
      rb[e] = (gj < p.N && gk < k1) ? B[gk * p.br + gj * p.bc] : 0.0f;
    }
  };
  auto storeB = [&](int buf) {
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      int idx = tid + 256 * e, j, kk;
      if (b_nfast) { kk = idx / BN; j = idx - kk * BN; }
      else { j = idx / BK; kk = idx - j * BK; }
      Bs[buf][j][kk] = bf16_rne(rb[e]);
    }
  };

  f4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = f4{0, 0, 0, 0};

  if (k0 < k1) {
    loadA(k0);
    loadB(k0);
    storeA(0);
    storeB(0);
  }
  __syncthreads();
  int cur = 0;
  const int lr = lane & 15, lk = lane >> 4;
  for (int64_t kt = k0; kt < k1; kt += BK) {
    const bool more = kt + BK < k1;
    if (more) {
      loadA(kt + BK);
      loadB(kt + BK);
    }
    bf16x8 af[MI], bfr[NI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
      af[mi] = *reinterpret_cast<const bf16x8*>(&As[cur][wm * (BM / 2) + mi * 16 + lr][8 * lk]);
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
      bfr[ni] = *reinterpret_cast<const bf16x8*>(&Bs[cur][wn * (BN / 2) + ni * 16 + lr][8 * lk]);
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < NI; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mi], bfr[ni], acc[mi][ni], 0, 0, 0);
    if (more) {
      storeA(cur ^ 1);
      storeB(cur ^ 1);
    }
    __syncthreads();
    cur ^= 1;
  }

  if (p.partial) {
    float* P = p.partial + ((int64_t)bk.z * p.B + bidx) * (int64_t)p.M * p.N;
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < NI; ++ni)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          int i = m0 + wm * (BM / 2) + mi * 16 + 4 * lk + v;
          int j = n0 + wn * (BN / 2) + ni * 16 + lr;
          if (i < p.M && j < p.N) P[(int64_t)i * p.N + j] = acc[mi][ni][v];
        }
    return;
  }
  const double al = eff_alpha(p);
  float* C = p.c + b1 * p.cb1 + b2 * p.cb2;
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int i = m0 + wm * (BM / 2) + mi * 16 + 4 * lk + v;
        int j = n0 + wn * (BN / 2) + ni * 16 + lr;
        if (i < p.M && j < p.N) {
          float* cp = C + (int64_t)i * p.cr + (int64_t)j * p.cc;
          double val = al * (double)acc[mi][ni][v];
          if (p.beta != 0.0) val += p.beta * (double)(*cp);
          *cp = (float)val;
        }
      }
}

// Sum split-K partials in split order and apply alpha / beta.
template <typename T>
struct KGemmReduce {
  struct Args { GemmParams<T> p; int splits; };
  static constexpr int kThreads = 256;
  static constexpr int kMega = std::is_same<T, double>::value ? 2 : 3;
  static constexpr int kMegaGroup = 1;
  static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};

// sum_{z < splits} P[z * stride] in slice order from 0.0 (the split-K reduction's order), with up to
// 8 slices' loads in flight: a loop that waits for each slice's load before issuing the next pays a
// memory round trip per slice (the partials were written by workgroups of every XCD).
template <typename T>
__device__ __forceinline__ double sum_slices(const T* P, int64_t stride, int splits) {
  double s = 0.0;
  for (int z = 0; z < splits; z += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = (z + u < splits) ? (double)P[(int64_t)(z + u) * stride] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (z + u < splits) s += v[u];
  }
  return s;
}

template <typename T>
__device__ __forceinline__ void KGemmReduce<T>::run(const Args& a, const Blk& bk) {
  const GemmParams<T> p = a.p;
  const int splits = a.splits;
  if (p.skip && *p.skip) return;
  const int64_t MN = (int64_t)p.M * p.N;
  const int64_t total = MN * p.B;
  const double al = eff_alpha(p);
  for (int64_t t = bk.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)bk.gx * blockDim.x) {
    int64_t bidx = t / MN, r = t - bidx * MN;
    int i = (int)(r / p.N), j = (int)(r - (int64_t)i * p.N);
    const double s = sum_slices(p.partial + bidx * MN + r, (int64_t)p.B * MN, splits);
    int b1 = (int)(bidx / p.B2), b2 = (int)(bidx - (int64_t)b1 * p.B2);
    T* cp = p.c + b1 * p.cb1 + b2 * p.cb2 + (int64_t)i * p.cr + (int64_t)j * p.cc;
    double val = al * s;
    if (p.beta != 0.0) val += p.beta * (double)(*cp);
    *cp = (T)val;
  }
}

}  // namespace ttamen
