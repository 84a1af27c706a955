// Kernel launch plumbing: every kernel of the solver is a struct K with
//   struct Args { ... };                      plain-old-data launch arguments
//   static constexpr int kThreads;            block size upper bound
//   static __device__ void run(const Args&, const Blk&);
// and is launched through launch<K>(grid, block, lds, stream, args).
//
// Outside a cohort the launch is an ordinary kernel launch of k_one<K> on the
// given stream. Inside a cohort (a host thread that is running one system of a
// batched solve) the launch is recorded instead; the cohort later runs the
// recorded launches of many systems together as one k_grp<K> launch whose grid
// is the concatenation of their grids. Blk carries the block coordinates of the
// launch the block belongs to, so a kernel body computes exactly the same thing
// in either form.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <typeinfo>

namespace ttamen {

// Address space of read-only device memory (scalar loads on AMDGPU).
#define TT_CONST __attribute__((address_space(4)))

// Copy a plain struct out of read-only device memory with scalar loads.
template <class A>
__device__ __forceinline__ A load_const(const char* p) {
  static_assert(sizeof(A) % 4 == 0, "launch arguments are made of 4-byte words");
  union U {
    A a;
    unsigned w[sizeof(A) / 4];
    __device__ U() {}
  } u;
  const TT_CONST unsigned* q = (const TT_CONST unsigned*)p;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(A) / 4); ++i) u.w[i] = q[i];
  return u.a;
}

struct Blk {
  unsigned x, y, z;     // block index within its own launch
  unsigned gx, gy, gz;  // grid of its own launch
};

constexpr int kMaxGroup = 256;  // launches merged into one grouped launch, at most

// Items of a grouped launch live in device memory, `stride` bytes apart:
//   unsigned grid[4] (gx, gy, gz, kind), then the Args of that launch.
struct GroupHdr {
  const char* items;
  unsigned stride;
  int n;
  unsigned start[kMaxGroup + 1];  // first block of item i; start[n] = total blocks
};

template <class K>
__global__ __launch_bounds__(K::kThreads) void k_one(typename K::Args a) {
  K::run(a, Blk{blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z});
}

template <class K>
__global__ __launch_bounds__(K::kThreads) void k_grp(GroupHdr h) {
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
  Blk bk;
  bk.gx = g.x; bk.gy = g.y; bk.gz = g.z;
  bk.x = l % g.x;
  l /= g.x;
  bk.y = l % g.y;
  bk.z = l / g.y;
  // The item table is read-only for the kernel's lifetime: load it through the constant
  // address space so the arguments land in scalar registers, as kernel arguments do.
  const typename K::Args a = load_const<typename K::Args>(it + 16);
  K::run(a, bk);
}

// One grouped launch as prepared by the cohort: items already copied to the device.
struct GroupLaunch {
  GroupHdr hdr;
  unsigned total_blocks;
  dim3 block;
  size_t lds;
  hipStream_t stream;
};
typedef void (*GroupFn)(const GroupLaunch&);

template <class K>
void group_exec(const GroupLaunch& gl) {
  hipLaunchKernelGGL((k_grp<K>), dim3(gl.total_blocks), gl.block, gl.lds, gl.stream, gl.hdr);
}

// Receives launches made while a cohort member runs on this thread.
class Recorder {
 public:
  // kind: the kernel index within a merged launch kind (mega_exec), else 0.
  virtual void record(GroupFn fn, dim3 grid, dim3 block, size_t lds, const void* args,
                      size_t size, unsigned kind) = 0;
  // Run everything recorded so far and wait for it (a host synchronisation point).
  virtual void sync() = 0;
  virtual ~Recorder() = default;
};

// The recorder of the cohort member running on this host thread, or null.
Recorder*& current_recorder();

// Names of launch kinds, for the cohort's optional launch histogram.
void register_kind_name(GroupFn fn, const char* name);
template <class K>
inline void register_kind() {
  static const bool once = (register_kind_name(&group_exec<K>, typeid(K).name()), true);
  (void)once;
}

// Kernels that declare `static constexpr int kMega` (their index in mega.hip) are, inside
// a cohort, recorded as items of one merged launch kind: a 256-thread kernel that runs
// each item's own body, so that the cohort can serve systems that are at different
// steps of the sweep with one launch. They use no dynamic LDS and work for any block
// size up to 256 threads.
// Kernels that also declare `kMegaGroup = 1` (the GEMMs) form a second merged kind, so
// that the light kernels of the first run at the occupancy their own registers allow.
constexpr int kMegaThreads = 256;
void mega_exec(const GroupLaunch& gl);
void mega_gemm_exec(const GroupLaunch& gl);

template <class K, class = void>
struct has_mega : std::false_type {};
template <class K>
struct has_mega<K, std::void_t<decltype(K::kMega)>> : std::true_type {};
template <class K, class = void>
struct mega_group : std::integral_constant<int, 0> {};
template <class K>
struct mega_group<K, std::void_t<decltype(K::kMegaGroup)>>
    : std::integral_constant<int, K::kMegaGroup> {};

template <class K>
inline void launch(dim3 grid, dim3 block, size_t lds, hipStream_t st, const typename K::Args& a) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  if (Recorder* r = current_recorder()) {
    if constexpr (has_mega<K>::value) {
      r->record(mega_group<K>::value ? &mega_gemm_exec : &mega_exec, grid, dim3(kMegaThreads), 0,
                &a, sizeof(a), (unsigned)K::kMega);
    } else {
      register_kind<K>();
      r->record(&group_exec<K>, grid, block, lds, &a, sizeof(a), 0);
    }
    return;
  }
  hipLaunchKernelGGL((k_one<K>), grid, block, lds, st, a);
}

// Allow `bytes` of dynamic LDS for both launch forms of K.
template <class K>
inline hipError_t allow_lds(int bytes) {
  hipError_t e1 = hipFuncSetAttribute((const void*)k_one<K>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  hipError_t e2 = hipFuncSetAttribute((const void*)k_grp<K>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  return e1 != hipSuccess ? e1 : e2;
}

}  // namespace ttamen
