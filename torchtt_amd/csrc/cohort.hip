// Cohort engine (see cohort.h): coroutine members on host worker threads, recorded
// launches merged into grouped launches, one stream.
#include "cohort.h"
#include "gmres.h"
#include "kernels_gemm.h"

#include <ucontext.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>

namespace ttamen {

static thread_local Recorder* tl_recorder = nullptr;
Recorder*& current_recorder() { return tl_recorder; }

static std::mutex g_kind_mu;
static std::map<GroupFn, std::string> g_kind_names;
void register_kind_name(GroupFn fn, const char* name) {
  std::lock_guard<std::mutex> lk(g_kind_mu);
  g_kind_names[fn] = name;
}

CohortParams& cohort_params() {
  static CohortParams p;
  return p;
}

// ---------------------------------------------------------------- mailboxes
// Pinned host mailboxes are created once and reused by later members.
static std::mutex g_hs_mu;
static std::vector<HostSync*> g_hs_free;

static HostSync* hs_acquire() {
  {
    std::lock_guard<std::mutex> lk(g_hs_mu);
    if (!g_hs_free.empty()) {
      HostSync* h = g_hs_free.back();
      g_hs_free.pop_back();
      return h;
    }
  }
  HostSync* h = new HostSync();
  h->init();
  return h;
}
static void hs_release(HostSync* h) {
  std::lock_guard<std::mutex> lk(g_hs_mu);
  g_hs_free.push_back(h);
}

// ---------------------------------------------------------------- members
static constexpr size_t kMemberStack = 1u << 20;

struct Cohort::Member final : Recorder {
  struct Op {
    GroupFn fn;
    dim3 grid, block;
    uint32_t lds, off, size, kind;
  };
  std::vector<Op> ops;
  std::vector<char> args;
  ucontext_t ctx;
  ucontext_t* back = nullptr;
  char* stack = nullptr;
  std::function<void()> body, on_done;
  bool started = false, finished = false;
  int worker = 0;
  size_t head = 0;  // first op not yet issued
  int age = 0;      // merge steps waited with a non-empty queue
  double weight = 1.0;
  HostSync* hs = nullptr;
  bool idle() const { return head == ops.size(); }

  void record(GroupFn fn, dim3 grid, dim3 block, size_t lds, const void* a, size_t size,
              unsigned kind) override {
    const size_t off = (args.size() + 15) & ~size_t(15);
    args.resize(off + size);
    std::memcpy(args.data() + off, a, size);
    ops.push_back({fn, grid, block, (uint32_t)lds, (uint32_t)off, (uint32_t)size, kind});
  }
  // Yield to the worker; the cohort runs the recorded work before resuming us.
  void sync() override {
    if (!ops.empty()) swapcontext(&ctx, back);
  }
  ~Member() override {
    std::free(stack);
    if (hs) hs_release(hs);
  }
};

HostSync& member_host_sync() {
  if (Recorder* r = tl_recorder) return *static_cast<Cohort::Member*>(r)->hs;
  static thread_local HostSync hs;
  hs.init();
  return hs;
}

static void member_entry(unsigned lo, unsigned hi) {
  Cohort::Member* m = (Cohort::Member*)(((uintptr_t)hi << 32) | (uintptr_t)lo);
  try {
    m->body();
  } catch (...) {
    // bodies report their own errors; nothing may unwind past the coroutine
  }
  m->finished = true;
  swapcontext(&m->ctx, m->back);
  std::abort();  // a finished member is never resumed
}

static void resume(Cohort::Member* m, ucontext_t* here) {
  m->back = here;
  tl_recorder = m;
  if (!m->started) {
    m->started = true;
    TT_CHECK(getcontext(&m->ctx) == 0, ERR_INVALID, "cohort: getcontext failed");
    m->stack = (char*)std::malloc(kMemberStack);
    TT_CHECK(m->stack != nullptr, ERR_INVALID, "cohort: out of host memory");
    m->ctx.uc_stack.ss_sp = m->stack;
    m->ctx.uc_stack.ss_size = kMemberStack;
    m->ctx.uc_link = nullptr;
    const uintptr_t p = (uintptr_t)m;
    makecontext(&m->ctx, (void (*)())member_entry, 2, (unsigned)(p & 0xffffffffu),
                (unsigned)(p >> 32));
  }
  swapcontext(here, &m->ctx);
  tl_recorder = nullptr;
}

// ---------------------------------------------------------------- worker threads
struct Cohort::Sync {
  std::atomic<unsigned> gen{0};
  std::atomic<int> pending{0};
  std::atomic<bool> stop{false};
  std::vector<std::thread> threads;
  std::vector<std::vector<Member*>> lists;
};

template <class F>
static void spin_until(F f) {
  for (int i = 0; !f(); ++i) {
    if (i < 4096) __builtin_ia32_pause();
    else std::this_thread::yield();
  }
}

Cohort::Cohort(hipStream_t st, void* dstage, size_t stage_bytes, int workers)
    : st_(st), dstage_((char*)dstage), stage_cap_(stage_bytes), workers_(std::max(1, workers)),
      sync_(new Sync) {
  TT_CHECK(dstage && stage_bytes >= 65536, ERR_INVALID, "cohort: staging buffer too small");
  TT_HIP(hipHostMalloc((void**)&hstage_, stage_cap_, hipHostMallocDefault));
  trace_ = std::getenv("TTAMEN_COHORT_TRACE") != nullptr;
  register_kind_name(&mega_exec, "merged kind: GMRES / streaming / scalar kernels");
  register_kind_name(&mega_gemm_exec, "merged kind: GEMMs");
  sync_->lists.resize(workers_);
}

Cohort::~Cohort() {
  sync_->stop.store(true, std::memory_order_release);
  for (auto& t : sync_->threads) t.join();
  members_.clear();
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
  if (hstage_) (void)hipHostFree(hstage_);
}

void Cohort::add(std::function<void()> body, std::function<void()> on_done, double weight) {
  TT_CHECK((int)members_.size() < kMaxGroup, ERR_INVALID, "cohort: too many members");
  std::unique_ptr<Member> m(new Member());
  m->body = std::move(body);
  m->on_done = std::move(on_done);
  m->weight = weight > 0 ? weight : 1.0;
  m->hs = hs_acquire();
  // balance: the worker with the fewest members
  std::vector<int> load(workers_, 0);
  for (auto& x : members_) load[x->worker] += 1;
  int w = 0;
  for (int i = 1; i < workers_; ++i)
    if (load[i] < load[w]) w = i;
  m->worker = w;
  members_.push_back(std::move(m));
}

void Cohort::resume_all(int worker) {
  ucontext_t here;
  for (Member* m : sync_->lists[worker]) resume(m, &here);
}

void Cohort::worker_loop(int w) {
  unsigned seen = 0;
  for (;;) {
    spin_until([&] {
      return sync_->stop.load(std::memory_order_acquire) ||
             sync_->gen.load(std::memory_order_acquire) != seen;
    });
    if (sync_->stop.load(std::memory_order_acquire)) return;
    seen = sync_->gen.load(std::memory_order_acquire);
    resume_all(w);
    sync_->pending.fetch_sub(1, std::memory_order_acq_rel);
  }
}

namespace {
struct GemmGroupProfile {
  std::atomic<bool> on{false};
  std::mutex mu;
  double v[6] = {0, 0, 0, 0, 0, 0};  // launches, ms, flops FP64, FP32, BF16, items
} g_gprof;
}  // namespace

void cohort_gemm_profile(int enable, double* out6) {
  std::lock_guard<std::mutex> lk(g_gprof.mu);
  if (enable == 1)
    for (double& x : g_gprof.v) x = 0;
  if (enable >= 0) g_gprof.on.store(enable == 1);
  if (out6)
    for (int i = 0; i < 6; ++i) out6[i] = g_gprof.v[i];
}

void Cohort::flush() {
  if (pending_.empty()) return;
  const auto t0 = std::chrono::steady_clock::now();
  TT_HIP(hipMemcpyAsync(dstage_, hstage_, stage_off_, hipMemcpyHostToDevice, st_));
  const bool prof = g_gprof.on.load(std::memory_order_relaxed);
  size_t nev = 0;
  double fl[3] = {0, 0, 0}, items = 0;
  for (auto& p : pending_) {
    const bool timed = prof && p.fn == &mega_gemm_exec;
    if (timed) {
      if (ev_pool_.size() < nev + 2) {
        hipEvent_t e;
        for (int i = 0; i < 2; ++i) {
          TT_HIP(hipEventCreate(&e));
          ev_pool_.push_back(e);
        }
      }
      TT_HIP(hipEventRecord(ev_pool_[nev], st_));
    }
    p.fn(p.gl);
    if (timed) {
      TT_HIP(hipEventRecord(ev_pool_[nev + 1], st_));
      nev += 2;
      for (int i = 0; i < 3; ++i) fl[i] += p.flops[i];
      items += p.gl.hdr.n;
    }
  }
  TT_HIP(hipGetLastError());
  TT_HIP(hipStreamSynchronize(st_));
  if (nev) {
    double ms = 0;
    for (size_t i = 0; i < nev; i += 2) {
      float t = 0;
      TT_HIP(hipEventElapsedTime(&t, ev_pool_[i], ev_pool_[i + 1]));
      ms += t;
    }
    std::lock_guard<std::mutex> lk(g_gprof.mu);
    g_gprof.v[0] += (double)(nev / 2);
    g_gprof.v[1] += ms;
    for (int i = 0; i < 3; ++i) g_gprof.v[2 + i] += fl[i];
    g_gprof.v[5] += items;
  }
  cnt_.t_gpu += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  pending_.clear();
  stage_off_ = 0;
}

void Cohort::dispatch(int max_steps) {
  std::vector<Member*> act;
  for (auto& m : members_)
    if (!m->idle()) act.push_back(m.get());
  if (act.empty()) return;
  struct Key {
    GroupFn fn;
    unsigned bx, by, bz;
    double score;
  };
  std::vector<Key> keys;
  std::vector<int> sel;
  double total_w = 0.0, busy_w = 0.0;
  for (auto& m : members_) {
    total_w += m->weight;
    if (!m->idle()) busy_w += m->weight;
  }
  const double idle_stop = total_w * cohort_params().idle_pct / 100.0;
  for (int step = 0; step < max_steps; ++step) {
    // end the round early once enough of the cohort waits to be resumed
    if (step > 0 && total_w - busy_w >= idle_stop && idle_stop < total_w) break;
    // Launch kinds at the heads of the queues: take the one whose waiting members weigh
    // most, unless a member has waited max_age steps (then its kind, so none starves).
    keys.clear();
    int oldest = -1;
    for (size_t i = 0; i < act.size(); ++i) {
      Member* m = act[i];
      if (m->idle()) continue;
      if (oldest < 0 || m->age > act[oldest]->age) oldest = (int)i;
      const Member::Op& op = m->ops[m->head];
      bool found = false;
      for (auto& k : keys)
        if (k.fn == op.fn && k.bx == op.block.x && k.by == op.block.y && k.bz == op.block.z) {
          k.score += m->weight;
          found = true;
          break;
        }
      if (!found) keys.push_back({op.fn, op.block.x, op.block.y, op.block.z, m->weight});
    }
    if (keys.empty()) break;
    cnt_.head_kinds += (double)keys.size();
    Key best = keys[0];
    for (const auto& k : keys)
      if (k.score > best.score) best = k;
    if (act[oldest]->age >= cohort_params().max_age) {
      const Member::Op& op = act[oldest]->ops[act[oldest]->head];
      best = Key{op.fn, op.block.x, op.block.y, op.block.z, 0};
    }
    sel.clear();
    for (size_t i = 0; i < act.size(); ++i) {
      Member* m = act[i];
      if (m->idle()) continue;
      const Member::Op& op = m->ops[m->head];
      if (op.fn == best.fn && op.block.x == best.bx && op.block.y == best.by &&
          op.block.z == best.bz)
        sel.push_back((int)i);
      else
        m->age += 1;
    }
    uint32_t size = 0;
    for (int i : sel) size = std::max(size, act[i]->ops[act[i]->head].size);
    const size_t stride = 16 + ((size + 15) & ~size_t(15));
    const size_t bytes = stride * sel.size();
    TT_CHECK(bytes <= stage_cap_, ERR_WORKSPACE, "cohort: staging buffer too small");
    if (stage_off_ + bytes > stage_cap_) flush();
    Pending p;
    p.fn = best.fn;
    p.flops[0] = p.flops[1] = p.flops[2] = 0.0;
    GroupLaunch& gl = p.gl;
    gl.hdr.items = dstage_ + stage_off_;
    gl.hdr.stride = (unsigned)stride;
    gl.hdr.n = (int)sel.size();
    gl.block = dim3(best.bx, best.by, best.bz);
    gl.lds = 0;
    gl.stream = st_;
    uint64_t total = 0;
    char* hp = hstage_ + stage_off_;
    for (size_t j = 0; j < sel.size(); ++j) {
      Member* m = act[sel[j]];
      const Member::Op& op = m->ops[m->head];
      gl.hdr.start[j] = (unsigned)total;
      total += (uint64_t)op.grid.x * op.grid.y * op.grid.z;
      unsigned* g = (unsigned*)(hp + j * stride);
      g[0] = op.grid.x;
      g[1] = op.grid.y;
      g[2] = op.grid.z;
      g[3] = op.kind;
      std::memcpy(hp + j * stride + 16, m->args.data() + op.off, op.size);
      if (best.fn == &mega_gemm_exec && (op.kind == 0 || op.kind == 1)) {
        // algorithmic flops of a GEMM item: 2 M N K per batch entry (split-K slices included once)
        int dims[5];  // M, N, K, B2, B: the leading ints of GemmParams<double> and <float>
        std::memcpy(dims, m->args.data() + op.off, sizeof(dims));
        const double f = 2.0 * dims[0] * dims[1] * dims[2] * dims[4];
        if (op.kind == 0) p.flops[0] += f;
        else {
          GemmParams<float> q;
          std::memcpy(&q, m->args.data() + op.off, sizeof(q));
          p.flops[q.tile >= 4 ? 2 : 1] += f;
        }
      }
      gl.lds = std::max<size_t>(gl.lds, op.lds);
      m->head += 1;
      m->age = 0;
      if (m->idle()) busy_w -= m->weight;
    }
    TT_CHECK(total < (1ull << 31), ERR_INVALID, "cohort: grouped grid too large");
    gl.hdr.start[sel.size()] = (unsigned)total;
    gl.total_blocks = (unsigned)total;
    stage_off_ = (stage_off_ + bytes + 255) & ~size_t(255);
    pending_.push_back(p);
    cnt_.launches += 1;
    cnt_.items += (double)sel.size();
    if (trace_) {
      auto& h = hist_[best.fn];
      h.first += 1;
      h.second += (double)sel.size();
    }
  }
  flush();
  for (Member* m : act)
    if (m->idle()) {
      m->ops.clear();
      m->args.clear();
      m->head = 0;
      m->age = 0;
    }
}

void Cohort::run(const std::function<bool(Cohort&)>& admit) {
  typedef std::chrono::steady_clock clk;
  auto secs = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  const auto t_start = clk::now();
  for (int w = 1; w < workers_; ++w)
    sync_->threads.emplace_back([this, w] { worker_loop(w); });
  bool more = true;
  for (;;) {
    if (more) more = admit(*this);
    if (members_.empty()) {
      if (!more) break;
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      continue;
    }
    // Resume the members whose recorded work has all run.
    const auto t0 = clk::now();
    for (auto& l : sync_->lists) l.clear();
    int nrun = 0;
    for (auto& m : members_)
      if (!m->finished && m->idle()) {
        sync_->lists[m->worker].push_back(m.get());
        ++nrun;
      }
    if (nrun > 0) {
      sync_->pending.store(workers_ - 1, std::memory_order_release);
      sync_->gen.fetch_add(1, std::memory_order_acq_rel);
      resume_all(0);
      spin_until([&] { return sync_->pending.load(std::memory_order_acquire) == 0; });
    }
    const auto t1 = clk::now();
    cnt_.resumes += nrun;
    cnt_.t_resume += secs(t0, t1);
    const double g0 = cnt_.t_gpu;
    dispatch(cohort_params().window);
    cnt_.t_build += secs(t1, clk::now()) - (cnt_.t_gpu - g0);
    cnt_.rounds += 1;
    // Retire members whose work has all run.
    for (size_t i = 0; i < members_.size();) {
      if (members_[i]->finished && members_[i]->idle()) {
        if (members_[i]->on_done) members_[i]->on_done();
        members_.erase(members_.begin() + i);
      } else {
        ++i;
      }
    }
  }
  cnt_.t_total = secs(t_start, clk::now());
  if (trace_) {
    std::vector<std::pair<double, GroupFn>> order;
    for (auto& h : hist_) order.push_back({h.second.first, h.first});
    std::sort(order.rbegin(), order.rend());
    std::lock_guard<std::mutex> lk(g_kind_mu);
    fprintf(stderr, "cohort: %.0f rounds, %.0f launches, %.0f items, %.2f kinds at the heads per step\n",
            cnt_.rounds, cnt_.launches, cnt_.items, cnt_.head_kinds / std::max(1.0, cnt_.launches));
    for (auto& o : order) {
      auto& h = hist_[o.second];
      fprintf(stderr, "  %9.0f launches %10.0f items %7.1f per launch  %s\n", h.first, h.second,
              h.second / h.first, g_kind_names[o.second].c_str());
    }
  }
}

}  // namespace ttamen
