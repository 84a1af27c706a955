// Allocation-free local operator: the FP64 local matvec of ttamen_local_matvec with its scratch
// (the permuted operator core, its zero-tile masks, the stage intermediates and any split-K partials) in a file-scope
// device array instead of a caller workspace, for callers that cannot allocate device memory
// (code under graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` exists once per GPU the library runs on, from the code object's load until
// the process ends. Host code never passes its name to a kernel (in host code the name is a host
// stand-in, not a device address): the device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and kept per device.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind: what must start at
// zero (the fused kernel's counters) is cleared by a kernel at the start of each call, and every
// other byte a call reads it has written first. Calls that overlap on one GPU
// (two host threads, or two streams without ordering between them) race on the array.
#include "ttamen_internal.h"
#include "../../include/ttamen.h"

#include <mutex>
#include <string>
#include <vector>

namespace ttamen {

constexpr size_t kNoAllocScratchBytes = 64ull << 20;  // 64 MiB per GPU

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

// Device address of `scratch` on the current device (looked up on the first call per device).
static char* scratch_on_current_device() {
  static std::mutex mu;
  static std::vector<char*> ptrs;
  int dev = 0;
  TT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void* p = nullptr;
    TT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TT_CHECK(p != nullptr, ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    ptrs[dev] = (char*)p;
  }
  return ptrs[dev];
}

// Scratch the FP64 local matvec needs at these sizes: the same allocation sequence as the
// call, replayed on a measuring arena (no device work).
static size_t local_matvec_scratch(const LocalOp& shape, int nvec) {
  Arena meas;
  meas.alloc<double>((size_t)shape.s * shape.n * shape.n * shape.S);
  meas.alloc<double>(frag_a_elems(shape.n * shape.S, shape.s * shape.n));
  if ((int64_t)shape.s * shape.n <= kTileMaskMaxK) meas.alloc<uint64_t>(tile_mask_words(shape.n, shape.S));
  local_matvec(shape, nullptr, nullptr, nvec, meas, nullptr);
  return meas.peak();
}

}  // namespace ttamen

using namespace ttamen;

extern "C" {

size_t ttamen_local_matvec_noalloc_limit(void) { return kNoAllocScratchBytes; }

int ttamen_local_matvec_noalloc(const double* PhiL, const double* A, const double* PhiR, const double* x,
                                double* y, int64_t l1, int64_t s, int64_t r1, int64_t n, int64_t S,
                                int64_t l2, int64_t r2, int64_t nvec, void* stream) {
  return abi_call([&] {
    TT_CHECK(PhiL && A && PhiR && x && y, ERR_INVALID, "null argument");
    TT_CHECK(l1 > 0 && s > 0 && r1 > 0 && n > 0 && S > 0 && l2 > 0 && r2 > 0 && nvec > 0,
             ERR_INVALID, "sizes must be positive");
    TT_CHECK(l1 <= INT32_MAX && s <= INT32_MAX && r1 <= INT32_MAX && n <= INT32_MAX && S <= INT32_MAX &&
                 l2 <= INT32_MAX && r2 <= INT32_MAX && nvec <= INT32_MAX,
             ERR_INVALID, "sizes out of range");
    LocalOp op;
    op.l1 = (int)l1; op.s = (int)s; op.r1 = (int)r1; op.n = (int)n; op.S = (int)S;
    op.l2 = (int)l2; op.r2 = (int)r2;
    const size_t need = local_matvec_scratch(op, (int)nvec);
    TT_CHECK(need <= kNoAllocScratchBytes, ERR_WORKSPACE,
             "ttamen_local_matvec_noalloc: these sizes need " + std::to_string(need) +
                 " bytes of scratch; the allocation-free entry point has " +
                 std::to_string(kNoAllocScratchBytes) + " (use ttamen_local_matvec with a workspace)");
    char* base = scratch_on_current_device();
    hipStream_t st = (hipStream_t)stream;
    Arena ws(base, kNoAllocScratchBytes);
    struct TuningOff {  // the tile autotuner creates events and scratch outputs: keep it out
      bool prev;
      TuningOff() : prev(gemm_tuning_off()) { gemm_tuning_off() = true; }
      ~TuningOff() { gemm_tuning_off() = prev; }
    } no_tune;
    double* Apm = ws.alloc<double>((size_t)s * n * n * S);
    permute_core_A(A, Apm, (int)s, (int)n, (int)S, st);
    double* Apk = ws.alloc<double>(frag_a_elems((int)(n * S), (int)(s * n)));
    pack_frag_a(Apm, Apk, (int)(n * S), (int)(s * n), st);
    // the fused kernel's counters come from the scratch and are cleared (by a kernel) first
    op.PhiL = PhiL; op.Apm = Apm; op.Apk = Apk; op.PhiR = PhiR;
    if (s * n <= kTileMaskMaxK) {  // zero-tile masks, written in full by their kernel
      uint64_t* tm = ws.alloc<uint64_t>(tile_mask_words(n, S));
      op_tile_masks(Apm, (int)n, (int)S, (int)(s * n), tm, st);
      op.tmask = tm;
    }
    local_matvec(op, x, y, (int)nvec, ws, st);
  });
}

}  // extern "C"
