// The merged launch kind of a cohort: one 256-thread kernel whose items are launches of
// different kernels (each item carries its kernel index), so that systems at different
// steps of the sweep are served by one launch. The kernels listed here are the frequent,
// short ones; the LDS-resident factorisations (Cholesky, Jacobi, LU panel, ...) keep
// their own launch kinds.
#include "ttamen_internal.h"
#include "kernels_blas1.h"
#include "kernels_gemm.h"
#include "kernels_gmres.h"

namespace ttamen {

namespace {
typedef KGemm<double> M0; typedef KGemm<float> M1;
typedef KGemmReduce<double> M2; typedef KGemmReduce<float> M3;
typedef KCgsDots<double> M4; typedef KCgsDots<float> M5;
typedef KCgsUpdate<double> M6; typedef KCgsUpdate<float> M7;
typedef KGmresClose<double> M8; typedef KGmresClose<float> M9;
typedef KGmresBegin<double> M10; typedef KGmresBegin<float> M11;
typedef KGmresXupd<double> M12; typedef KGmresXupd<float> M13;
typedef KGmresSetup M14; typedef KGmresY M15; typedef KGmresRearm M16;
typedef KSumsq<double> M17; typedef KSumsq<float> M18;
typedef KScaleSlot M19; typedef KAxpby M20; typedef KAxpbyF32 M21; typedef KAddF32F64 M22;
typedef KSub M23; typedef KCopy M24; typedef KFill M25; typedef KZero32 M26; typedef KEye M27;
typedef KCopy2d M28; typedef KCast<double, float> M29; typedef KCast<float, double> M30;
typedef KNrmsc M31; typedef KMaxdiv M32; typedef KSset M33; typedef KSdiv M34;
typedef KStoreHost M35; typedef KScaleCols M36; typedef KRandn M37; typedef KTauScale M38;
}  // namespace

#define TT_MEGA_CASE(N)                                                             \
  case N:                                                                           \
    static_assert(M##N::kMega == N, "mega index");                                  \
    static_assert(M##N::kThreads <= kMegaThreads, "mega block size");               \
    {                                                                               \
      const typename M##N::Args a = load_const<typename M##N::Args>(args);         \
      M##N::run(a, bk);                                                             \
    }                                                                               \
    break;

// Locate the item of block blockIdx.x; returns its kernel index, arguments and block.
__device__ __forceinline__ unsigned mega_item(const GroupHdr& h, const char** args, Blk* bk) {
  const unsigned b = blockIdx.x;
  int lo = 0, hi = h.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (h.start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const char* it = h.items + (size_t)lo * h.stride;
  const uint4 g = load_const<uint4>(it);
  unsigned l = b - h.start[lo];
  bk->gx = g.x; bk->gy = g.y; bk->gz = g.z;
  bk->x = l % g.x;
  l /= g.x;
  bk->y = l % g.y;
  bk->z = l / g.y;
  *args = it + 16;
  return g.w;
}

__global__ __launch_bounds__(kMegaThreads) void k_mega(GroupHdr h) {
  const char* args;
  Blk bk;
  const unsigned kind = mega_item(h, &args, &bk);
  switch (kind) {
    TT_MEGA_CASE(4)
    TT_MEGA_CASE(5) TT_MEGA_CASE(6) TT_MEGA_CASE(7) TT_MEGA_CASE(8) TT_MEGA_CASE(9)
    TT_MEGA_CASE(10) TT_MEGA_CASE(11) TT_MEGA_CASE(12) TT_MEGA_CASE(13) TT_MEGA_CASE(14)
    TT_MEGA_CASE(15) TT_MEGA_CASE(16) TT_MEGA_CASE(17) TT_MEGA_CASE(18) TT_MEGA_CASE(19)
    TT_MEGA_CASE(20) TT_MEGA_CASE(21) TT_MEGA_CASE(22) TT_MEGA_CASE(23) TT_MEGA_CASE(24)
    TT_MEGA_CASE(25) TT_MEGA_CASE(26) TT_MEGA_CASE(27) TT_MEGA_CASE(28) TT_MEGA_CASE(29)
    TT_MEGA_CASE(30) TT_MEGA_CASE(31) TT_MEGA_CASE(32) TT_MEGA_CASE(33) TT_MEGA_CASE(34)
    TT_MEGA_CASE(35) TT_MEGA_CASE(36) TT_MEGA_CASE(37) TT_MEGA_CASE(38)
    default:
      break;
  }
}

// The GEMMs (both precisions, every tile shape) and their split-K reductions.
__global__ __launch_bounds__(kMegaThreads) void k_mega_gemm(GroupHdr h) {
  const char* args;
  Blk bk;
  const unsigned kind = mega_item(h, &args, &bk);
  switch (kind) {
    TT_MEGA_CASE(0) TT_MEGA_CASE(1) TT_MEGA_CASE(2) TT_MEGA_CASE(3)
    default:
      break;
  }
}

void mega_exec(const GroupLaunch& gl) {
  hipLaunchKernelGGL(k_mega, dim3(gl.total_blocks), dim3(kMegaThreads), 0, gl.stream, gl.hdr);
}
void mega_gemm_exec(const GroupLaunch& gl) {
  hipLaunchKernelGGL(k_mega_gemm, dim3(gl.total_blocks), dim3(kMegaThreads), 0, gl.stream, gl.hdr);
}

}  // namespace ttamen
