// Cohort execution of many independent solves on one stream.
//
// Each member of a cohort is one solve, written as ordinary sequential host code
// (the same code as a single solve). It runs as a coroutine on one of a few host
// worker threads. Its kernel launches are recorded (launch.h) instead of issued;
// when it needs a value from the device (a host synchronisation point) it yields.
//
// One round: every runnable member (one whose recorded work has all run) runs until
// its next synchronisation point; then up to `window` grouped launches are built
// from the heads of the members' queues, launch kind by launch kind (one grid covering
// the same step of many members), issued on the cohort's stream, and waited for. A
// member whose queue is not yet empty stays suspended, so members with short segments
// (GMRES iterations) are not held back by members with long ones (a right-to-left
// half-sweep). Per member the launches keep their order, so each member computes
// exactly what it computes when run alone.
#pragma once

#include "ttamen_internal.h"
#include <functional>
#include <map>
#include <memory>
#include <vector>

namespace ttamen {

class HostSync;

// Scheduling parameters (ttamen_cohort_set_params): grouped launches per round, merge
// steps a member may wait before it is served, GMRES iterations between polls, lanes,
// and the share of the members' weight that, once idle, ends a round early.
struct CohortParams {
  int window = 128;
  int max_age = 64;
  int gmres_poll = 8;
  int lanes = 3;       // independent cohorts, each on its own stream, run concurrently
  int idle_pct = 15;   // a round also ends once this % of the members' weight waits to resume
};
CohortParams& cohort_params();

class Cohort {
 public:
  // dstage: device staging for the grouped launches' argument tables (stage_bytes).
  Cohort(hipStream_t st, void* dstage, size_t stage_bytes, int workers);
  ~Cohort();
  Cohort(const Cohort&) = delete;
  Cohort& operator=(const Cohort&) = delete;

  // Add a member; it starts at the next round. `on_done` runs on the calling thread of
  // run() once the member's last recorded work has completed on the device.
  // `weight` (> 0) ranks members when launch kinds compete: the kind whose waiting
  // members weigh most is launched first (heavier systems are served first).
  void add(std::function<void()> body, std::function<void()> on_done, double weight = 1.0);
  int members() const { return (int)members_.size(); }

  // Rounds until no member is left and `admit` (called before every round, may add
  // members) adds none. `admit` returns false when nothing more will ever be added.
  void run(const std::function<bool(Cohort&)>& admit);

  struct Counters {
    double rounds = 0, launches = 0, items = 0, resumes = 0;
    double t_resume = 0, t_build = 0, t_gpu = 0, t_total = 0;  // seconds
    double head_kinds = 0;  // sum over merge steps of the distinct launch kinds at the heads
  };
  const Counters& counters() const { return cnt_; }

  struct Member;

 private:
  void resume_all(int worker);
  void dispatch(int max_steps);
  void flush();
  void worker_loop(int w);

  hipStream_t st_;
  char* dstage_;
  size_t stage_cap_;
  char* hstage_ = nullptr;
  size_t stage_off_ = 0;
  int workers_;
  std::vector<std::unique_ptr<Member>> members_;
  struct Pending {
    GroupFn fn;
    GroupLaunch gl;
    double flops[3];  // merged GEMM launches: algorithmic flops of the items (FP64, FP32, BF16)
  };
  std::vector<hipEvent_t> ev_pool_;  // GEMM launch timing (cohort_gemm_profile)
  std::vector<Pending> pending_;
  Counters cnt_;
  bool trace_ = false;  // TTAMEN_COHORT_TRACE set: print a launch histogram at the end
  std::map<GroupFn, std::pair<double, double>> hist_;
  struct Sync;
  std::unique_ptr<Sync> sync_;
};

// Profile of the merged GEMM launches of every cohort (k_mega_gemm): enable 1 clears and starts
// it, 0 stops it, -1 leaves it; out6 (optional): launches timed, their total ms (HIP events on
// the launch's stream), algorithmic flops of their items in FP64, FP32 and BF16, items.
void cohort_gemm_profile(int enable, double* out6);

// The pinned-host mailbox of the member running on this thread, or of the thread.
HostSync& member_host_sync();

}  // namespace ttamen
