// Strided, batched GEMM on CDNA4 matrix cores (gfx950).
//
// Every tensor-train contraction of the AMEn sweep (local operator, interface
// updates, right-hand-side assembly, Gram matrices of the QR) is lowered to this
// one kernel by choosing strides, so the operands are never transposed in memory.
//
// Tile: BM x BN x BK per workgroup of 256 threads = 4 wave64s in a 2x2 grid.
// Each wave owns a (BM/2) x (BN/2) sub-tile built from 16x16 MFMA blocks:
//   FP64: v_mfma_f64_16x16x4_f64   (A: 1 double/lane, B: 1 double/lane, D: 4 doubles/lane)
//   FP32: v_mfma_f32_16x16x4_f32   (same fragment shapes in float)
// Fragment layout (16x16xK4 block): lane l holds A[i = l%16][k = l/16] and
// B[k = l/16][j = l%16]; result register v of lane l holds
//   FP32: D[i = 4*(l/16) + v][j = l%16]      FP64: D[i = l/16 + 4*v][j = l%16]
// (the FP64 layout was established on the device with known operands).
// Global -> register -> LDS double buffering: the next K tile is fetched while
// the current one feeds the MFMAs. The global fetch walks whichever operand
// axis has unit stride, so loads are coalesced for either orientation.
// Split-K writes per-split partials that a second kernel sums in a fixed order
// (deterministic results, no atomics).
#include "ttamen_internal.h"
#include "kernels_gemm.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <unordered_map>

namespace ttamen {

// ---------------------------------------------------------------- tile autotuner
// Shape buckets: element type, ceil(log2) of M, N, K and the batch count, and which operand axes
// have unit stride. The first launch of a bucket (autotune mode 1, outside cohorts, a workspace
// with room for a scratch output and no device skip flag) runs each of the four tile shapes, with
// its own split-K plan, once to warm up and three times under HIP events, writing a scratch
// output; the fastest is cached for the process and used for every later launch of the bucket.
// Neither the tile shape nor the split-K plan (a function of the shape alone) changes a result.
namespace {
std::atomic<int> g_tune_mode{0};
std::mutex g_tune_mu;
std::unordered_map<uint64_t, int> g_tune_cache;  // bucket -> tile index (0..3)
double g_tune_stats[4] = {0, 0, 0, 0};  // buckets tuned, launches served, seconds tuning, differ from rule

int lg2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v && l < 40) ++l;
  return l;
}
uint64_t bucket_key(const GemmDesc& g, int type) {
  uint64_t k = (uint64_t)type;
  k = k * 64 + lg2(g.M);
  k = k * 64 + lg2(g.N);
  k = k * 64 + lg2(std::max(g.K, 1));
  k = k * 64 + lg2((int64_t)g.B1 * g.B2);
  k = k * 4 + (g.a.c == 1 ? 2 : 0) + (g.b.c == 1 ? 1 : 0);
  return k;
}
}  // namespace

bool& gemm_tuning_off() {
  static thread_local bool off = false;
  return off;
}

int gemm_autotune_mode(int mode) {
  const int prev = g_tune_mode.load();
  if (mode >= 0) g_tune_mode.store(mode);
  return prev;
}
void gemm_autotune_stats(double* out4) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int i = 0; i < 4; ++i) out4[i] = g_tune_stats[i];
}
void gemm_autotune_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache.clear();
  for (double& v : g_tune_stats) v = 0;
}

template <typename T, int BM, int BN>
static void launch_tile(GemmParams<T>& p, int splits, hipStream_t st, bool bf16) {
  int tiles = ((p.M + BM - 1) / BM) * ((p.N + BN - 1) / BN);
  dim3 grid(tiles, p.B, splits);
  p.tile = (BM == 32 ? 0 : 2) + (BN == 32 ? 0 : 1) + (bf16 ? 4 : 0);
  launch<KGemm<T>>(grid, dim3(256), 0, st, p);
  TT_LAUNCH_CHECK();
}

template <typename T>
static void gemm_impl(const GemmDesc& g, Arena& ws, hipStream_t st, bool bf16 = false) {
  TT_CHECK(g.M >= 0 && g.N >= 0 && g.K >= 0 && g.B1 >= 1 && g.B2 >= 1, ERR_INVALID,
           "gemm: bad sizes");
  if (g.M == 0 || g.N == 0) return;
  GemmParams<T> p;
  p.M = g.M; p.N = g.N; p.K = g.K; p.B2 = g.B2; p.B = g.B1 * g.B2;
  p.a = (const T*)g.a.p; p.ar = g.a.r; p.ac = g.a.c; p.ab1 = g.a.b1; p.ab2 = g.a.b2;
  p.b = (const T*)g.b.p; p.br = g.b.r; p.bc = g.b.c; p.bb1 = g.b.b1; p.bb2 = g.b.b2;
  p.c = (T*)g.c.p; p.cr = g.c.r; p.cc = g.c.c; p.cb1 = g.c.b1; p.cb2 = g.c.b2;
  p.alpha = g.alpha; p.beta = g.beta; p.scale_ptr = g.scale_ptr; p.scale_mode = g.scale_mode;
  p.skip = g.skip; p.partial = nullptr;
  TT_CHECK(p.B <= 65535, ERR_INVALID, "gemm: batch too large");

  // 64x64 tiles unless the grid of 64x64 tiles would be under 4 workgroups per CU (1024):
  // the AMEn shapes (batched local-operator stages, Gram matrices, thin products) then run
  // with one wave per SIMD and expose the MFMA and load latencies, and 32x32 tiles (4x the
  // workgroups) measured 1.1-2x faster on them. Inside a cohort the GEMMs of many systems
  // share each launch and already fill the device, so the larger tiles stay.
  const int64_t tiles64 = (int64_t)((g.M + 63) / 64) * ((g.N + 63) / 64) * p.B;
  const bool small_grid = tiles64 < 1024 && current_recorder() == nullptr;
  int BM = (g.M <= 32 || small_grid) ? 32 : 64;
  int BN = (g.N <= 32 || small_grid) ? 32 : 64;
  // Split K when the output grid is too small to fill the 256 CUs and K is long, into slices at
  // least 64 deep (four steps of the K loop, which waits on each step's loads: a long slice on a
  // few workgroups is latency-bound). The plan
  // depends on the shape only (the tile count of the cohort's rule: 64, or 32 for an extent of
  // at most 32), never on the tile that runs, so a GEMM sums its K slices in the same order
  // alone, inside a cohort and under the autotuner: each result is bitwise the same.
  auto split_plan = [&](int, int, int64_t* kchunk_out) {
    const int bm = g.M <= 32 ? 32 : 64, bn = g.N <= 32 ? 32 : 64;
    const int tiles = ((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn) * p.B;
    int splits = 1;
    if (g.K == 0) {
      *kchunk_out = 1;
      return 1;
    }
    if (tiles < 256 && g.K >= 512) {
      splits = std::min<int64_t>((512 + tiles - 1) / tiles, g.K / 64);
      splits = std::max(1, std::min(splits, 64));
    }
    int64_t kchunk = (g.K + splits - 1) / splits;
    kchunk = (kchunk + 15) / 16 * 16;
    *kchunk_out = kchunk;
    return (int)((g.K + kchunk - 1) / kchunk);
  };
  const bool tunable = !ws.measuring() && current_recorder() == nullptr && g.skip == nullptr && g.K > 0 &&
                       g_tune_mode.load() == 1 && !gemm_tuning_off() && g.split_out == nullptr;
  if (tunable) {
    const uint64_t key = bucket_key(g, bf16 ? 2 : (sizeof(T) == 8 ? 0 : 1));
    int choice = -1;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) {
        choice = it->second;
        g_tune_stats[1] += 1;
      }
    }
    if (choice < 0) {
      // time the four tiles into a scratch output (row-major, batch-contiguous)
      const size_t out_elems = (size_t)p.B * g.M * g.N;
      size_t need = ((out_elems * sizeof(T) + 255) & ~size_t(255));
      for (int t = 0; t < 4; ++t) {
        int64_t kc;
        const int sp = split_plan(t & 2 ? 64 : 32, t & 1 ? 64 : 32, &kc);
        if (sp > 1) need = std::max(need, ((out_elems * sizeof(T) + 255) & ~size_t(255)) +
                                              (((size_t)sp * out_elems * sizeof(T) + 255) & ~size_t(255)));
      }
      if (ws.remaining() >= need + 1024) {
        const auto t0 = std::chrono::steady_clock::now();
        size_t mk0 = ws.mark();
        GemmParams<T> q = p;
        T* cs = ws.alloc<T>(out_elems);
        q.c = cs; q.cr = g.N; q.cc = 1; q.cb1 = (int64_t)g.B2 * g.M * g.N; q.cb2 = (int64_t)g.M * g.N;
        q.beta = 0.0;
        hipEvent_t e0, e1;
        TT_HIP(hipEventCreate(&e0));
        TT_HIP(hipEventCreate(&e1));
        float best_ms = 0;
        for (int t = 0; t < 4; ++t) {
          const int bm = t & 2 ? 64 : 32, bn = t & 1 ? 64 : 32;
          size_t mk1 = ws.mark();
          const int sp = split_plan(bm, bn, &q.kchunk);
          q.partial = sp > 1 ? ws.alloc<T>((size_t)sp * out_elems) : nullptr;
          auto go = [&] {
            if (bm == 32 && bn == 32) launch_tile<T, 32, 32>(q, sp, st, bf16);
            else if (bm == 32) launch_tile<T, 32, 64>(q, sp, st, bf16);
            else if (bn == 32) launch_tile<T, 64, 32>(q, sp, st, bf16);
            else launch_tile<T, 64, 64>(q, sp, st, bf16);
            if (sp > 1) {
              int64_t total = (int64_t)g.M * g.N * q.B;
              launch<KGemmReduce<T>>(dim3((int)std::min<int64_t>((total + 255) / 256, 4096)), dim3(256), 0, st,
                                     {q, sp});
              TT_LAUNCH_CHECK();
            }
          };
          go();
          TT_HIP(hipEventRecord(e0, st));
          for (int rep = 0; rep < 3; ++rep) go();
          TT_HIP(hipEventRecord(e1, st));
          TT_HIP(hipEventSynchronize(e1));
          float ms = 0;
          TT_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (choice < 0 || ms < best_ms) { choice = t; best_ms = ms; }
          ws.release(mk1);
        }
        TT_HIP(hipEventDestroy(e0));
        TT_HIP(hipEventDestroy(e1));
        ws.release(mk0);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = choice;
        g_tune_stats[0] += 1;
        g_tune_stats[2] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
    }
    if (choice >= 0) {
      const int tbm = choice & 2 ? 64 : 32, tbn = choice & 1 ? 64 : 32;
      if (tbm != BM || tbn != BN) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_stats[3] += 1;
      }
      BM = tbm;
      BN = tbn;
    }
  }
  int splits = split_plan(BM, BN, &p.kchunk);
  if (g.split_out) {
    TT_CHECK(g.alpha == 1.0 && g.beta == 0.0 && g.scale_ptr == nullptr && g.K > 0, ERR_INVALID,
             "gemm: kept split-K partials need alpha 1, beta 0, no scale");
    g.split_out->partial = nullptr;
    g.split_out->splits = splits;
  }
  size_t mk = ws.mark();
  const size_t partial_elems = splits > 1 ? (size_t)splits * p.B * g.M * g.N : 0;
  if (partial_elems) p.partial = ws.alloc<T>(partial_elems);
  if (g.split_out && partial_elems) {  // the consumer adds them: they stay allocated
    g.split_out->partial = p.partial;
    mk = ws.mark();
  }
  if (!ws.measuring()) {
    if (g.K == 0) {
      // C = beta * C: run the kernel with an empty K range.
      splits = 1;
    }
    if (BM == 32 && BN == 32) launch_tile<T, 32, 32>(p, splits, st, bf16);
    else if (BM == 32) launch_tile<T, 32, 64>(p, splits, st, bf16);
    else if (BN == 32) launch_tile<T, 64, 32>(p, splits, st, bf16);
    else launch_tile<T, 64, 64>(p, splits, st, bf16);
    if (splits > 1 && !g.split_out) {
      int64_t total = (int64_t)g.M * g.N * p.B;
      int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
      launch<KGemmReduce<T>>(dim3(blocks), dim3(256), 0, st, {p, splits});
      TT_LAUNCH_CHECK();
    }
  }
  ws.release(mk);
}

void gemm_f64(const GemmDesc& g, Arena& ws, hipStream_t st) { gemm_impl<double>(g, ws, st); }
void gemm_f32(const GemmDesc& g, Arena& ws, hipStream_t st) { gemm_impl<float>(g, ws, st); }
// FP32 operands and result, BF16 products (operands rounded to BF16), FP32 accumulation.
void gemm_bf16(const GemmDesc& g, Arena& ws, hipStream_t st) { gemm_impl<float>(g, ws, st, true); }

}  // namespace ttamen
