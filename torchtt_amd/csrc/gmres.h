// Device-resident restarted GMRES: state layout and host driver interface.
#pragma once
#include "ttamen_internal.h"
#include <functional>

namespace ttamen {

constexpr int kGmresMaxCols = 128;    // (local_iterations + 1) upper bound
constexpr int kGmresMaxBlocks = 256;   // blocks of the streaming kernels

enum { GM_DONE = 0, GM_FLAG = 1, GM_KCOLS = 2, GM_ITERS = 3, GM_RSKIP = 4, GM_REORTHS = 5,  // istate
       GM_EXACT = 6,    // Arnoldi steps whose ||w''|| was reduced exactly (cancellation)
       GM_CNT = 8 };    // arrival counter of KGmresClose's exact-norm path (kept at zero)
enum { GM_BNORM = 0, GM_THR = 1, GM_RNORM = 2, GM_ERR = 3, GM_WN2 = 4, GM_WNORM = 5 };      // dstate

// Device buffers of one GMRES instance (all inside the caller's arena).
struct GmresDev {
  int m = 0;            // iterations per cycle
  void* Q = nullptr;    // (m+1) x N basis, working precision
  void* w = nullptr;    // N
  double* H = nullptr;  // (m+1) x m, column k at H + k*(m+1)
  double* beta = nullptr;
  double* cs = nullptr;
  double* sn = nullptr;
  double* y = nullptr;
  double* dstate = nullptr;
  int* istate = nullptr;      // 16 ints, then the 4 doubles of xtra
  double* xtra = nullptr;     // caller values read back with the cycle state (GmresOp::post)
  double* partial = nullptr;  // per-block partials, row-major [row][block]
};

template <typename T>
struct GmresOp {
  std::function<void(const T*, T*, const int*)> apply;          // y = Op x (skip flag)
  // Optional: y = Op x with the first orthogonalisation pass fused in (MvDots: the shares of Q_i . y,
  // i < kc, and y . y in P); returns the number of shares per row, 0 when it did not fuse them (y is
  // computed either way).
  std::function<int(const T*, T*, const int*, const T*, int, double*)> apply_dots;
  // Optional: Arnoldi-step application, tagged (cycle, step) so a caller can attribute
  // timing to the steps that actually ran (later steps of a converged cycle are skipped).
  // Receives the basis (Q, kc rows) and P1 as apply_dots does and returns the same count.
  std::function<int(const T*, T*, const int*, int, int, const T*, int, double*)> apply_step;
  std::function<void(double, const T*, double, T*, int64_t)> axpby;  // y = a x + b y
  std::function<void(const T*, int64_t, double*)> norm;         // *slot = ||x||
  // Optional: work enqueued after each cycle's solution update and before the cycle's state is read
  // back; it may store 4 doubles into `xtra`, which come back with the state (GmresResult::xtra).
  // It must give the same result whether or not another cycle follows (it can run once per cycle).
  std::function<void(double* xtra)> post;
};

struct GmresResult {
  int flag = 0, iters = 0, cycles = 0, reorths = 0;
  double xtra[4] = {0, 0, 0, 0};  // what GmresOp::post stored, read with the last cycle's state
  std::vector<int> cycle_steps;  // Arnoldi steps executed in each cycle
};

// Small pinned-host mailbox used to poll device flags without stalling the queue.
class HostSync {
 public:
  static constexpr int NS = 4;
  void init() {
    if (host_) return;
    TT_HIP(hipHostMalloc((void**)&host_, sizeof(int) * (NS + 32), hipHostMallocDefault));
    TT_HIP(hipHostMalloc((void**)&hostd_, sizeof(double) * 4096, hipHostMallocDefault));
    for (int i = 0; i < NS; ++i) TT_HIP(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
  }
  void post(const int* dflag, int idx, hipStream_t st) {
    int s = idx % NS;
    post_to_host(host_ + s, dflag, sizeof(int), st);
    if (!current_recorder()) TT_HIP(hipEventRecord(ev_[s], st));
  }
  // Slot of post idx for a kernel to store the flag into itself (pinned host memory the device writes);
  // mark(idx) then records the event that wait(idx) waits for.
  int* slot(int idx) { return host_ + idx % NS; }
  void mark(int idx, hipStream_t st) {
    if (!current_recorder()) TT_HIP(hipEventRecord(ev_[idx % NS], st));
  }
  bool wait(int idx, hipStream_t st) {
    if (idx < 0) return false;
    int s = idx % NS;
    if (current_recorder()) sync_point(st);
    else TT_HIP(hipEventSynchronize(ev_[s]));
    return host_[s] != 0;
  }
  int* host_int() { return host_ + NS; }
  double* host_dbl() { return hostd_; }
  ~HostSync() {
    // Pinned buffers live for the process; freeing at exit can race the runtime teardown.
  }

 private:
  int* host_ = nullptr;
  double* hostd_ = nullptr;
  hipEvent_t ev_[NS];
};

GmresDev gmres_alloc(Arena& ws, int m, int64_t N, int esize);

// Orthogonalisation timing of single-system solves: enable clears the counters; read gives
// Arnoldi steps timed, their total ms and their algorithmic bytes.
void ortho_profile_enable(bool on);
void ortho_profile_read(double* out3);

// Solve Op dx = rhs from dx = 0 (x must be zeroed by the caller), restarted
// `resets` times, stopping when |beta_{k+1}| / ||rhs|| <= real_tol*norm_rhs/dnorm.
template <typename T>
GmresResult gmres_run(GmresDev g, int64_t N, T* x, const T* rhs, const double* norm_rhs,
                      const double* dnorm, double real_tol, int resets, const GmresOp<T>& op,
                      RedScratch rs, HostSync& hs, hipStream_t st);

}  // namespace ttamen
