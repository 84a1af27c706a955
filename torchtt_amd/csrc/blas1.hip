// Vector kernels of the AMEn sweep: deterministic norms, scaling by device
// scalars, copies/casts, the operator-core permutation and the dense local
// operator assembly used by the direct local solver.
//
// All of these are HBM-bound streaming kernels; they use grid-stride loops with
// a fixed grid so that reductions combine partials in a fixed order.
#include "ttamen_internal.h"
#include "kernels_blas1.h"
#include <algorithm>

namespace ttamen {

static inline int grid_for(int64_t n, int per = 256 * 4, int cap = 2048) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}


void norm2(const double* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st, bool guard,
           double* mul_slot, const int* skip) {
  launch<KSumsq<double>>(dim3(grid_for(n)), dim3(256), 0, st,
                         {x, nullptr, n, rs.partial, rs.counter, out_slot, nullptr, guard ? 1 : 0,
                          mul_slot, skip});
  TT_LAUNCH_CHECK();
}

void norm2_scale_inv(double* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st, bool guard) {
  launch<KSumsq<double>>(dim3(grid_for(n)), dim3(256), 0, st,
                         {x, nullptr, n, rs.partial, rs.counter, out_slot, nullptr, guard ? 1 : 0, nullptr,
                          nullptr, x});
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

void axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t st,
           const int* skip) {
  if (n <= 0) return;
  launch<KAxpby>(dim3(grid_for(n)), dim3(256), 0, st, {a, x, b, y, n, skip});
  TT_LAUNCH_CHECK();
}

void axpby_f32(double a, const float* x, double b, float* y, int64_t n, hipStream_t st,
               const int* skip) {
  if (n <= 0) return;
  launch<KAxpbyF32>(dim3(grid_for(n)), dim3(256), 0, st, {a, x, b, y, n, skip});
  TT_LAUNCH_CHECK();
}

void add_f32_to_f64(const float* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  launch<KAddF32F64>(dim3(grid_for(n)), dim3(256), 0, st, {x, nullptr, y, n});
  TT_LAUNCH_CHECK();
}

void add_f32_to_f64_into(const float* x, const double* base, double* y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  launch<KAddF32F64>(dim3(grid_for(n)), dim3(256), 0, st, {x, base, y, n});
  TT_LAUNCH_CHECK();
}

void sub(const double* x, const double* y, double* z, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  launch<KSub>(dim3(grid_for(n)), dim3(256), 0, st, {x, y, z, n});
  TT_LAUNCH_CHECK();
}

void copy(const double* x, double* y, int64_t n, hipStream_t st) {
  if (n <= 0 || x == y) return;
  launch<KCopy>(dim3(grid_for(n)), dim3(256), 0, st, {x, y, n});
  TT_LAUNCH_CHECK();
}

void fill(double* x, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return;
  launch<KFill>(dim3(grid_for(n)), dim3(256), 0, st, {x, n, v});
  TT_LAUNCH_CHECK();
}

void zero_words(void* x, int64_t nwords, hipStream_t st) {
  if (nwords <= 0) return;
  launch<KZero32>(dim3(grid_for(nwords)), dim3(256), 0, st, {(uint32_t*)x, nwords});
  TT_LAUNCH_CHECK();
}

void eye(double* x, int64_t m, int64_t n, hipStream_t st) {
  if (m * n <= 0) return;
  launch<KEye>(dim3(grid_for(m * n)), dim3(256), 0, st, {x, m, n});
  TT_LAUNCH_CHECK();
}

void copy2d(const double* src, int64_t sr, int64_t sc, double* dst, int64_t ldd, int64_t m,
            int64_t n, hipStream_t st) {
  if (m * n <= 0) return;
  launch<KCopy2d>(dim3(grid_for(m * n)), dim3(256), 0, st, {src, sr, sc, dst, ldd, m, n});
  TT_LAUNCH_CHECK();
}

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

void permute_core_A_rect(const double* A, double* P, int s, int nm, int nk, int S, hipStream_t st) {
  int64_t total = (int64_t)s * nm * nk * S;
  launch<KPermA<double>>(dim3(grid_for(total)), dim3(256), 0, st, {A, P, s, nm, nk, S});
  TT_LAUNCH_CHECK();
}
void permute_core_A(const double* A, double* P, int s, int n, int S, hipStream_t st) {
  permute_core_A_rect(A, P, s, n, n, S, st);
}
void permute_core_A_f32(const double* A, float* P, int s, int n, int S, hipStream_t st) {
  int64_t total = (int64_t)s * n * n * S;
  launch<KPermA<float>>(dim3(grid_for(total)), dim3(256), 0, st, {A, P, s, n, n, S});
  TT_LAUNCH_CHECK();
}

void assemble_local(const double* PhiL, const double* A, const double* PhiR, double* B, int64_t ld,
                    int r1, int s, int n, int S, int r2, hipStream_t st) {
  int64_t N = (int64_t)r1 * n * r2;
  launch<KAssemble>(dim3(grid_for(N * N, 256, 4096)), dim3(256), 0, st,
                    {PhiL, A, PhiR, B, ld, r1, s, n, S, r2});
  TT_LAUNCH_CHECK();
}

void nrmsc_update(double* S, int i_nrmsc, int i_a, int i_x, int i_b, int mode, hipStream_t st) {
  launch<KNrmsc>(dim3(1), dim3(1), 0, st, {S, i_nrmsc, i_a, i_x, i_b, mode});
  TT_LAUNCH_CHECK();
}
void scalar_maxdiv(double* S, int dst, int src, int den, hipStream_t st) {
  launch<KMaxdiv>(dim3(1), dim3(1), 0, st, {S, dst, src, den});
  TT_LAUNCH_CHECK();
}
void scalar_set(double* S, int dst, double v, hipStream_t st) {
  launch<KSset>(dim3(1), dim3(1), 0, st, {S, dst, v});
  TT_LAUNCH_CHECK();
}
void scalar_div(double* S, int dst, int a, int b, hipStream_t st) {
  launch<KSdiv>(dim3(1), dim3(1), 0, st, {S, dst, a, b});
  TT_LAUNCH_CHECK();
}


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
