/*
 * ttamen.h -- C ABI of libttamen, the MI355X-native AMEn linear solver for
 * tensor-train systems A x = b.
 *
 * The drop-in boundary replaces the reference's native entry point
 *   torchttcpp.amen_solve(A_cores, b_cores, x0_cores, N, rA, rb, r_x0, nswp, eps, rmax,
 *                         max_full, kickrank, kick2, local_iterations, resets, verbose,
 *                         preconditioner)
 * (reference: cpp/cpp_ext.cpp:9, cpp/amen_solve.h:122-139, called from
 * torchtt/solvers.py:283-285), and exposes the kernels it is built from so that each
 * can be checked on its own.
 *
 * Conventions
 *  - All array arguments named *_cores / device pointers point to GPU memory of the
 *    current HIP device; the caller owns every buffer. No function allocates device
 *    memory: scratch comes from the `workspace` argument, sized by the matching
 *    *_workspace_size query.
 *  - Cores are row-major: a vector core is r x n x r', an operator core is
 *    r x m x n x r' (row index m, column index n), as torch/the reference store them.
 *  - Every function returns 0 on success and a positive status otherwise
 *    (1 invalid argument, 2 workspace too small, 3 HIP error, 4 unsupported,
 *    5 non-finite result); ttamen_last_error() returns a message for the calling thread.
 *  - `stream` is a hipStream_t (0 = the null stream). Calls are asynchronous unless
 *    stated otherwise; amen_solve synchronises internally where its control flow
 *    needs device results (rank decisions, convergence).
 */
#ifndef TTAMEN_H
#define TTAMEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Library identification. */
const char* ttamen_version(void);
const char* ttamen_last_error(void);

/* Statistics of one solve (host struct, filled by ttamen_amen_solve when non-NULL). */
typedef struct ttamen_stats {
  int64_t sweeps;            /* sweeps performed */
  int64_t gmres_iters;       /* total GMRES iterations */
  int64_t local_solves;      /* local problems solved */
  int64_t direct_solves;     /* of which by the dense direct solver */
  int64_t matvecs;           /* local-operator applications */
  double matvec_flops;       /* flops of those applications as dense contractions */
  double matvec_bytes;       /* compulsory bytes of those applications */
  double max_res_last;       /* max local residual of the last sweep */
  double elapsed_s;          /* wall time inside the call */
  int64_t max_rank;          /* largest TT rank reached */
  int64_t refinements;       /* FP64 refinement steps (precision 1 or 2) */
  int64_t matvecs_f32;       /* FP32 / BF16 local-operator applications (precision 1 or 2) */
  double matvec_flops_f32;   /* their flops as dense contractions */
} ttamen_stats;

/*
 * AMEn solve (FP64 plain method).
 * Replaces: torchttcpp.amen_solve, cpp/amen_solve.h:122-549.
 *
 *  d                 number of cores
 *  N[d]              mode sizes; rA[d+1], rb[d+1] ranks of A and b (ends are 1)
 *  A_cores[d]        operator cores rA[k] x N[k] x N[k] x rA[k+1] (double, device)
 *  b_cores[d]        rhs cores rb[k] x N[k] x rb[k+1] (double, device)
 *  x0_cores, rx0     initial guess cores/ranks, or NULL (then x0 = all-ones, rank 1)
 *  z0_cores          initial residual-proxy cores rz[k] x N[k] x rz[k+1] with
 *                    rz = (1, kickrank+kick2, ..., 1), normal random (device), or NULL
 *                    to draw them on the device from `seed`
 *  nswp, eps, rmax, max_full, kickrank, kick2, local_iterations, resets, verbose,
 *  preconditioner    as in the reference: 0 none, 1 'c' (per-(l,L) n x n block inverses),
 *                    2 'r' (per-l (n r2) x (n r2) block inverses, order <= 90), right
 *                    preconditioning of the iterative local solve (FP64 only); the workspace
 *                    must then also hold ttamen_amen_solve_prec_workspace_size(...) bytes
 *  precision         0 = FP64 everywhere (plain); 1 = local GMRES in FP32 on the FP32
 *                    local operator (MFMA f32), inside FP64 iterative refinement of every
 *                    local solve (FP64 residual rhs - B x each step); interfaces, QR/SVD,
 *                    truncation and the direct solver stay FP64; 2 = as 1 with BF16
 *                    products (operands rounded to BF16 on entry to the matrix cores,
 *                    FP32 accumulation, FP32 Krylov vectors)
 *  workspace         device scratch of ttamen_amen_solve_workspace_size(...) bytes
 *  x_out             device buffer receiving the solution cores back to back
 *  x_out_elems       its capacity in doubles (ttamen_amen_solve_output_elems(...))
 *  rx_out[d+1]       host array receiving the solution ranks
 *  offsets_out[d]    host array receiving each core's offset (in doubles) in x_out
 */
int ttamen_amen_solve(int d, const int64_t* N, const int64_t* rA, const int64_t* rb,
                      const double* const* A_cores, const double* const* b_cores,
                      const double* const* x0_cores, const int64_t* rx0,
                      const double* const* z0_cores, uint64_t seed, int64_t nswp, double eps,
                      int64_t rmax, int64_t max_full, int64_t kickrank, int64_t kick2,
                      int64_t local_iterations, int64_t resets, int verbose, int preconditioner,
                      int precision, void* workspace, size_t workspace_bytes, double* x_out,
                      size_t x_out_elems, int64_t* rx_out, int64_t* offsets_out,
                      ttamen_stats* stats, void* stream);

/*
 * Batched solves of independent systems (no counterpart in the reference, which solves one
 * system per call). The systems run as one cohort on `stream`: each runs the same solver as
 * ttamen_amen_solve in its own region of `workspace`, and at every host synchronisation
 * point the launches that all running systems have queued are merged into grouped launches
 * (one grid covering the same kernel for every system), so each launch fills the device.
 * As many systems as fit in the workspace run at once (admitted in index order), spread
 * over a few concurrent cohorts ("lanes") of up to 256 systems, each on its own stream;
 * `workers` host threads run the systems' host code. Results are bitwise those of
 * ttamen_amen_solve with the same inputs. Shared options apply to all systems; rmax and
 * seed are per system. Each system's status is written to its `status` field; the call
 * returns the first failure's status. The first ttamen_cohort_staging_bytes() bytes of
 * `workspace` hold the grouped launches' argument tables.
 */
typedef struct ttamen_system {
  int d;
  const int64_t* N;             /* d */
  const int64_t* rA;            /* d+1 */
  const int64_t* rb;            /* d+1 */
  const double* const* A_cores; /* d device pointers */
  const double* const* b_cores; /* d device pointers */
  const double* const* x0_cores; /* d device pointers or NULL */
  const int64_t* rx0;           /* d+1 or NULL */
  const double* const* z0_cores; /* d device pointers or NULL */
  uint64_t seed;
  int64_t rmax;
  double* x_out;                /* device */
  size_t x_out_elems;
  int64_t* rx_out;              /* host, d+1 */
  int64_t* offsets_out;         /* host, d */
  ttamen_stats* stats;          /* host, may be NULL */
  int status;                   /* out */
} ttamen_system;

int ttamen_amen_solve_batch(ttamen_system* systems, int count, int64_t nswp, double eps,
                            int64_t max_full, int64_t kickrank, int64_t kick2,
                            int64_t local_iterations, int64_t resets, int precision, int workers,
                            void* workspace, size_t workspace_bytes, void* stream);
/* Workspace for ttamen_amen_solve_batch: staging plus room for the systems to run all at
 * once (the 256 largest needs). A smaller workspace runs fewer systems at a time. */
size_t ttamen_amen_solve_batch_workspace_size(const ttamen_system* systems, int count,
                                              int64_t kickrank, int64_t kick2,
                                              int64_t local_iterations, int64_t max_full,
                                              int precision, int workers);

/*
 * Serving loop (no counterpart in the reference): continuous batching of solve requests.
 * Request i becomes eligible at arrival_s[i] seconds after the call starts (sorted). Up to
 * `slots` (<= 256) requests run at once as one cohort on `stream` (see
 * ttamen_amen_solve_batch), so the same sweep steps of different requests execute as one
 * grouped launch; a request is admitted between two such rounds, in arrival order, when a
 * slot is idle and a contiguous region of its footprint (its solver workspace at the rank
 * cap) is free in `workspace` after the staging area, and keeps its cores and interfaces
 * resident there until it finishes. start_s[i] / done_s[i] receive the admission and
 * completion times (seconds since the call started); latency = done_s - arrival_s.
 * Results equal those of ttamen_amen_solve.
 */
int ttamen_serve(ttamen_system* systems, const double* arrival_s, int count, int64_t nswp,
                 double eps, int64_t max_full, int64_t kickrank, int64_t kick2,
                 int64_t local_iterations, int64_t resets, int precision, int slots,
                 void* workspace, size_t workspace_bytes, double* start_s, double* done_s,
                 void* stream);

/*
 * Online serving (no counterpart in the reference): the same continuous batching as ttamen_serve,
 * for requests that arrive while others run. ttamen_server_create starts a server that owns
 * `workspace` (the first ttamen_cohort_staging_bytes() bytes hold launch tables, the rest is the
 * serving memory) and runs cohort lanes on host threads and streams of its own. Shared solver
 * options are fixed at creation; rank cap, seed and x0/z0 come with each request.
 * ttamen_server_submit queues one request (its descriptor, arrays and device buffers stay the
 * caller's and must stay valid until the request is done); its inputs may still be in flight on
 * `stream`: the request starts after the work queued there so far. Requests are admitted in
 * submission order when a slot is idle and a contiguous region of their footprint is free; a
 * request larger than the serving memory is refused (status 2). ttamen_server_poll reports the
 * state (0 queued, 1 running, 2 done) and the submission, admission and completion times (seconds
 * since creation); ttamen_server_wait blocks until the request is done and returns its status
 * (the solution is then complete on the device). ttamen_server_destroy lets every submitted request
 * finish, then stops the server. Each result equals ttamen_amen_solve's on the same inputs.
 */
typedef struct ttamen_server ttamen_server;
int ttamen_server_create(ttamen_server** out, int64_t nswp, double eps, int64_t max_full, int64_t kickrank,
                         int64_t kick2, int64_t local_iterations, int64_t resets, int precision, int slots,
                         void* workspace, size_t workspace_bytes);
int ttamen_server_submit(ttamen_server* server, ttamen_system* system, void* stream, int64_t* id_out);
int ttamen_server_poll(ttamen_server* server, int64_t id, int* state_out, double* times_out);
int ttamen_server_wait(ttamen_server* server, int64_t id);
int ttamen_server_destroy(ttamen_server* server);

/*
 * TT rounding (replaces the reference's native round_this, cpp/compression.h:3-41, bound as
 * torchttcpp.round_this at cpp/cpp_ext.cpp:10, which rounds in place): left-to-right QR, then
 * right-to-left truncated SVD with relative accuracy eps / sqrt(d-1) per bond and rank cap
 * rmax (rank rule of cpp/ortho.h:16-44). Input cores (device, FP64) are not modified; the
 * rounded cores are written to `out` at offsets_out[k] with ranks r_out (d+1). N[k] is the
 * mode size of core k (m*n for an operator core). Ranks up to 140. Synchronous at the end.
 */
/*
 * Elementwise division x = b ./ a of two TT tensors by AMEn. Replaces the reference's Python
 * solver torchtt._division.amen_divide (torchtt/_division.py:76-398; no native binding exists
 * in the reference), reached from TT.__truediv__ / __rtruediv__ (torchtt/_tt_base.py:898-970)
 * and torchtt.elementwise_divide (torchtt/_extras.py:588-612).
 *  a_cores: ra[k] x N[k] x ra[k+1] (TT tensor a), b_cores: rb[k] x N[k] x rb[k+1]; FP64 device
 *  x0_cores/rx0: initial guess (the reference's default is all-ones cores of ranks 1,2,..,2,1,
 *               _division.py:86-88, which the caller passes); z0_cores as in ttamen_amen_solve
 *  rmax: d+1 rank caps, one per interface (_division.py:94-95)
 *  local_iterations, resets: each GMRES restart runs up to local_iterations + 1 steps, as the
 *               reference's call does (_division.py:253)
 *  preconditioner: 0 none, 1 'c'
 * The operator diag(a) is formed in the workspace as TT-matrix cores a[s,m,S] delta_mn; the
 * sweep is ttamen_amen_solve's with the reference Python solver's truncation rule (rank >= 2
 * when every candidate passes, _division.py:283-296). Status codes as ttamen_amen_solve.
 */
int ttamen_amen_divide(int d, const int64_t* N, const int64_t* ra, const int64_t* rb,
                       const double* const* a_cores, const double* const* b_cores,
                       const double* const* x0_cores, const int64_t* rx0,
                       const double* const* z0_cores, int64_t nswp, double eps, const int64_t* rmax,
                       int64_t max_full, int64_t kickrank, int64_t kick2, int64_t local_iterations,
                       int64_t resets, int verbose, int preconditioner, void* workspace,
                       size_t workspace_bytes, double* x_out, size_t x_out_elems, int64_t* rx_out,
                       int64_t* offsets_out, ttamen_stats* stats, void* stream);
size_t ttamen_amen_divide_workspace_size(int d, const int64_t* N, const int64_t* ra,
                                         const int64_t* rb, const int64_t* rx0, const int64_t* rmax,
                                         int64_t kickrank, int64_t kick2, int64_t local_iterations,
                                         int64_t max_full, int preconditioner);
size_t ttamen_amen_divide_output_elems(int d, const int64_t* N, const int64_t* rx0,
                                       const int64_t* rmax, int64_t kickrank, int64_t kick2);

/*
 * AMEn approximation of TT products (reference torchtt/_amen.py; Python only, no native binding):
 * y ~= A b (amen_mv, :55-101) for a TT-matrix A with cores rA[k] x M[k] x K[k] x rA[k+1] and a TT
 * vector b with cores rb[k] x K[k] x rb[k+1]; the sweep of _amen_mm_python (:126-413): projected
 * product per core, rank_chop truncation at ||w|| eps / d^1.5 (d^0.5 in the last sweep), residual
 * proxy z of kickrank (+ kick2 random) columns, stop one sweep after max dx < eps.
 *  x0_cores/rx0: initial guess (the reference's default: zero cores of rank 1, when NULL)
 *  z0_cores: residual-proxy cores rz x M[k] x rz' (rz = kickrank + kick2; the reference draws a
 *            random TT, _amen.py:146-151); rmax: d+1 per-interface rank caps (:142-143)
 *  seed: random enrichment columns when kick2 > 0
 * Output as ttamen_amen_solve: cores rx[k] x M[k] x rx[k+1] at offsets_out in x_out.
 */
int ttamen_amen_mv(int d, const int64_t* M, const int64_t* K, const int64_t* rA, const int64_t* rb,
                   const double* const* A_cores, const double* const* b_cores,
                   const double* const* x0_cores, const int64_t* rx0, const double* const* z0_cores,
                   int64_t nswp, double eps, const int64_t* rmax, int64_t kickrank, int64_t kick2,
                   uint64_t seed, int verbose, void* workspace, size_t workspace_bytes, double* x_out,
                   size_t x_out_elems, int64_t* rx_out, int64_t* offsets_out, ttamen_stats* stats,
                   void* stream);
size_t ttamen_amen_mv_workspace_size(int d, const int64_t* M, const int64_t* K, const int64_t* rA,
                                     const int64_t* rb, const int64_t* rx0, const int64_t* rmax,
                                     int64_t kickrank, int64_t kick2);
size_t ttamen_amen_mv_output_elems(int d, const int64_t* M, const int64_t* K, const int64_t* rx0,
                                   const int64_t* rmax, int64_t kickrank, int64_t kick2);
/*
 * C ~= A B for TT-matrices A (rA[k] x M[k] x K[k] x rA[k+1]) and B (rB[k] x K[k] x N[k] x rB[k+1])
 * (reference amen_mm, torchtt/_amen.py:104-123): the ttamen_amen_mv sweep on the operator
 * A (x) I_N (formed in the workspace) and B read as a vector of modes K[k] N[k]; the result cores
 * rx[k] x (M[k] N[k]) x rx[k+1] are those of C, [r, m, n, r'] as the reference keeps them.
 * z0_cores: rz x (M[k] N[k]) x rz'. Output sizes: ttamen_amen_mv_output_elems with M[k] N[k].
 */
int ttamen_amen_mm(int d, const int64_t* M, const int64_t* K, const int64_t* N, const int64_t* rA,
                   const int64_t* rB, const double* const* A_cores, const double* const* B_cores,
                   const double* const* X0_cores, const int64_t* rx0, const double* const* z0_cores,
                   int64_t nswp, double eps, const int64_t* rmax, int64_t kickrank, int64_t kick2,
                   uint64_t seed, int verbose, void* workspace, size_t workspace_bytes, double* x_out,
                   size_t x_out_elems, int64_t* rx_out, int64_t* offsets_out, ttamen_stats* stats,
                   void* stream);
size_t ttamen_amen_mm_workspace_size(int d, const int64_t* M, const int64_t* K, const int64_t* N,
                                     const int64_t* rA, const int64_t* rB, const int64_t* rx0,
                                     const int64_t* rmax, int64_t kickrank, int64_t kick2);

/*
 * DMRG approximation of y = A x (reference torchttcpp.dmrg_mv, cpp/dmrg_mv.h:23-230, bound at
 * cpp/cpp_ext.cpp:11 and called from torchtt/_dmrg.py:42): A has cores rA[k] x M[k] x N[k] x
 * rA[k+1], x cores rx[k] x N[k] x rx[k+1]. Sweeps of right-to-left orthogonalisation and
 * left-to-right two-core updates: the supercore projection W of A x, delta = ||W - y_k y_k+1|| /
 * ||W|| steering a per-core rank enlargement, rank_chop truncation at ||W|| eps / d^1.5
 * (d^0.5 in the last sweep), kickrank random columns in every sweep but the final one; the
 * sweep after max delta < eps is the last.
 *  y0_cores/ry0: initial guess (NULL: rank-2 normal cores drawn on the device from `seed`, as
 *                the reference draws torch::randn cores, dmrg_mv.h:43-51)
 *  rmax: d+1 per-interface rank caps (the reference's scalar rmax on every interface)
 *  seed: the device's normal stream for the kick columns (and the default guess)
 *  d >= 2; ranks above 140 (the kick included) are refused with status 4 (unsupported).
 * Output as ttamen_amen_solve: cores ry[k] x M[k] x ry[k+1] at offsets_out in y_out; stats
 * report sweeps, the largest rank and the last sweep's max delta (max_res_last).
 */
int ttamen_dmrg_mv(int d, const int64_t* M, const int64_t* N, const int64_t* rA, const int64_t* rx,
                   const double* const* A_cores, const double* const* x_cores,
                   const double* const* y0_cores, const int64_t* ry0, int64_t nswp, double eps,
                   const int64_t* rmax, int64_t kickrank, uint64_t seed, int verbose, void* workspace,
                   size_t workspace_bytes, double* y_out, size_t y_out_elems, int64_t* ry_out,
                   int64_t* offsets_out, ttamen_stats* stats, void* stream);
size_t ttamen_dmrg_mv_workspace_size(int d, const int64_t* M, const int64_t* N, const int64_t* rA,
                                     const int64_t* rx, const int64_t* ry0, const int64_t* rmax,
                                     int64_t kickrank);
size_t ttamen_dmrg_mv_output_elems(int d, const int64_t* M, const int64_t* ry0, const int64_t* rmax,
                                   int64_t kickrank);

/*
 * Full tensor of a TT vector with cores R[k] x N[k] x R[k+1] (reference torchttcpp.tt_full,
 * cpp/full.h:3-12, bound at cpp/cpp_ext.cpp:8): out receives the N[0] x ... x N[d-1] tensor
 * (row-major, out_elems >= its size). One GEMM per core; the products must stay below 2^31 elements.
 */
int ttamen_tt_full(int d, const int64_t* N, const int64_t* R, const double* const* cores, void* workspace,
                   size_t workspace_bytes, double* out, size_t out_elems, void* stream);
size_t ttamen_tt_full_workspace_size(int d, const int64_t* N, const int64_t* R);

/*
 * Global residual check in FP64 on the GPU: out2[0] = ||A x - b||, out2[1] = ||b|| (Frobenius).
 * Replaces the TT algebra the reference's callers use for it, (A @ x - b).norm() / b.norm()
 * (torchtt/_tt_base.py:780-846, :545, :990-1052): a left-to-right sweep over the residual TT
 * (ranks rA rx + rb, its cores never formed) with Householder R factors, so that the norm keeps
 * its accuracy when A x and b cancel to a small residual. Cores row-major on the device:
 * A[k] rA_k x N_k x N_k x rA_{k+1}, x[k] rx_k x N_k x rx_{k+1}, b[k] rb_k x N_k x rb_{k+1}.
 * Synchronous (out2 is host memory).
 */
int ttamen_tt_residual_norm(int d, const int64_t* N, const int64_t* rA, const int64_t* rx, const int64_t* rb,
                            const double* const* A_cores, const double* const* x_cores,
                            const double* const* b_cores, void* workspace, size_t workspace_bytes,
                            double* out2, void* stream);
size_t ttamen_tt_residual_norm_workspace_size(int d, const int64_t* N, const int64_t* rA, const int64_t* rx,
                                              const int64_t* rb);
size_t ttamen_tt_round_workspace_size(int d, const int64_t* N, const int64_t* R);
int ttamen_tt_round(int d, const int64_t* N, const int64_t* R, const double* const* cores,
                    double eps, int64_t rmax, void* workspace, size_t workspace_bytes, double* out,
                    size_t out_elems, int64_t* r_out, int64_t* offsets_out, void* stream);

/*
 * Profile of the merged GEMM launches of batched / served solves (no counterpart in the reference):
 * enable 1 clears and starts it (each merged GEMM launch is bracketed by HIP events on its stream),
 * 0 stops it, -1 leaves it as it is. out6 (optional): launches timed, their total duration in ms,
 * the algorithmic flops (2 M N K per batch entry) of their GEMM items in FP64, FP32 and BF16, and
 * the number of items.
 */
int ttamen_cohort_gemm_profile(int enable, double* out6);
/* Bytes at the front of a batch / serving workspace that hold grouped-launch tables. */
size_t ttamen_cohort_staging_bytes(void);
/* Counters of the last batch / serving call (no counterpart in the reference): rounds,
 * grouped launches, launch items (member launches merged into them), member resumptions,
 * and seconds spent resuming members, building grouped launches, waiting for the device,
 * in total. */
int ttamen_cohort_stats(double* out8);
/* Scheduling parameters of later batch / serving calls (process-wide; results do not
 * depend on them): grouped launches per round (128), merge steps a system may wait before
 * its next launch is served (64), GMRES iterations between convergence polls (8), and the
 * number of lanes (3, at most 8): independent cohorts on their own streams that run
 * concurrently, the systems being spread over them as they are admitted; a round also ends
 * early once idle_pct % (15; 100: never) of the members' weight waits to be resumed. */
int ttamen_cohort_set_params(int window, int max_age, int gmres_poll, int lanes, int idle_pct);

/* Extra workspace bytes for a local preconditioner (1 'c', 2 'r'; 0 for none). */
size_t ttamen_amen_solve_prec_workspace_size(int d, const int64_t* N, const int64_t* rA,
                                             const int64_t* rb, const int64_t* rx0, int64_t rmax,
                                             int64_t kickrank, int64_t kick2, int preconditioner);
/*
 * One local preconditioner application y = P x (reference AMENsolveMV::setter/apply_prec,
 * cpp/matvecs.h:153-194): mode 1 'c', 2 'r'; PhiL r1 x s x r1, A s x n x n x S, PhiR
 * r2 x S x r2, x and y r1 x n x r2 (device, FP64). Synchronous.
 */
int ttamen_local_prec(int mode, const double* PhiL, const double* A, const double* PhiR,
                      const double* x, double* y, int64_t r1, int64_t s, int64_t n, int64_t S,
                      int64_t r2, void* workspace, size_t workspace_bytes, void* stream);

/* Bytes of device workspace ttamen_amen_solve needs for these sizes (0 on bad input). */
size_t ttamen_amen_solve_workspace_size(int d, const int64_t* N, const int64_t* rA,
                                        const int64_t* rb, const int64_t* rx0, int64_t rmax,
                                        int64_t kickrank, int64_t kick2,
                                        int64_t local_iterations, int64_t max_full,
                                        int precision);
/* Doubles needed in x_out (solution cores at their rank bounds). */
size_t ttamen_amen_solve_output_elems(int d, const int64_t* N, const int64_t* rx0, int64_t rmax,
                                      int64_t kickrank, int64_t kick2);

/*
 * Local operator y[v] = Phi_L * A_k * Phi_R * x[v], v < nvec (FP64).
 * Replaces: AMENsolveMV<T>::matvec, cpp/matvecs.h:196-212 and local_product,
 * cpp/amen_solve.h:17-27 ('lsr,smnS,LSR,rnR->lmL').
 * PhiL: l1 x s x r1, A: s x n x n x S, PhiR: l2 x S x r2, x: nvec x r1 x n x r2,
 * y: nvec x l1 x n x l2.
 */
int ttamen_local_matvec(const double* PhiL, const double* A, const double* PhiR,
                        const double* x, double* y, int64_t l1, int64_t s, int64_t r1,
                        int64_t n, int64_t S, int64_t l2, int64_t r2, int64_t nvec,
                        void* workspace, size_t workspace_bytes, void* stream);
/* Same in FP32 (inputs/outputs float, FP32 accumulation on MFMA). */
int ttamen_local_matvec_f32(const float* PhiL, const float* A, const float* PhiR, const float* x,
                            float* y, int64_t l1, int64_t s, int64_t r1, int64_t n, int64_t S,
                            int64_t l2, int64_t r2, int64_t nvec, void* workspace,
                            size_t workspace_bytes, void* stream);
/* Same with BF16 products: FP32 inputs/outputs, every operand rounded to BF16 (round to nearest
 * even) as it enters the matrix cores, FP32 accumulation. The operator of the BF16 local solves
 * (precision 2 of ttamen_amen_solve); no counterpart in the reference, whose only reduced
 * precision is FP32 (torchtt/solvers.py:469-496). */
int ttamen_local_matvec_bf16(const float* PhiL, const float* A, const float* PhiR, const float* x,
                             float* y, int64_t l1, int64_t s, int64_t r1, int64_t n, int64_t S,
                             int64_t l2, int64_t r2, int64_t nvec, void* workspace,
                             size_t workspace_bytes, void* stream);
size_t ttamen_local_matvec_workspace_size(int64_t l1, int64_t s, int64_t r1, int64_t n,
                                          int64_t S, int64_t l2, int64_t r2, int64_t nvec);
/*
 * Allocation-free variant of ttamen_local_matvec (same arguments and results, no workspace).
 * No counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid hipMalloc). The scratch the
 * kernels need (permuted operator core, stage intermediates, split-K partials) lives in a
 * file-scope device array of ttamen_local_matvec_noalloc_limit() bytes (64 MiB) that each GPU
 * holds from the library's load until the process ends; sizes needing more are refused with
 * status 2 before any device work. The call allocates nothing (no hipMalloc, hipFree or
 * allocating library routine). All calls on one GPU share its array: calls that overlap on one
 * GPU (two host threads, or two streams without ordering between them) race.
 */
int ttamen_local_matvec_noalloc(const double* PhiL, const double* A, const double* PhiR,
                                const double* x, double* y, int64_t l1, int64_t s, int64_t r1,
                                int64_t n, int64_t S, int64_t l2, int64_t r2, int64_t nvec,
                                void* stream);
size_t ttamen_local_matvec_noalloc_limit(void);

/*
 * Interface updates (FP64). Replace compute_phi_fwd_A / compute_phi_bck_A /
 * compute_phi_fwd_rhs / compute_phi_bck_rhs, cpp/amen_solve.h:38-98.
 *  phi_fwd_A:  Pn[L,S,R] = sum Phi[l,s,r] xL[l,M,L] A[s,M,N,S] xR[r,N,R]
 *  phi_bck_A:  Pp[l,s,r] = sum xL[l,M,L] Phi[L,S,R] A[s,M,N,S] xR[r,N,R]
 *  phi_fwd_rhs: Pn[B,R] = sum Pb[b,r] bc[b,n,B] x[r,n,R]
 *  phi_bck_rhs: Pp[b,r] = sum bc[b,n,B] Pb[B,R] x[r,n,R]
 */
int ttamen_phi_fwd_A(const double* Phi, const double* xL, const double* A, const double* xR,
                     double* Pn, int64_t l, int64_t s, int64_t r, int64_t n, int64_t S,
                     int64_t L, int64_t R, void* workspace, size_t workspace_bytes, void* stream);
int ttamen_phi_bck_A(const double* Phi, const double* xL, const double* A, const double* xR,
                     double* Pp, int64_t l, int64_t s, int64_t r, int64_t n, int64_t S,
                     int64_t L, int64_t R, void* workspace, size_t workspace_bytes, void* stream);
int ttamen_phi_fwd_rhs(const double* Pb, const double* bc, const double* x, double* Pn,
                       int64_t b, int64_t r, int64_t n, int64_t B, int64_t R, void* workspace,
                       size_t workspace_bytes, void* stream);
int ttamen_phi_bck_rhs(const double* Pb, const double* bc, const double* x, double* Pp,
                       int64_t b, int64_t r, int64_t n, int64_t B, int64_t R, void* workspace,
                       size_t workspace_bytes, void* stream);
size_t ttamen_phi_workspace_size(int64_t l, int64_t s, int64_t r, int64_t n, int64_t S,
                                 int64_t L, int64_t R);

/*
 * Local right-hand side rhs[r,n,R] = scale * sum Pb_L[b,r] bc[b,n,B] Pb_R[B,R] (reference
 * amen_solve.h:303-304, scale = the sweep's nrmsc). Workspace: ttamen_phi_workspace_size(b, 1, r,
 * n, 1, B, R) suffices.
 */
int ttamen_rhs_assemble(const double* PL, const double* bc, const double* PR, double* out, int64_t b,
                        int64_t r, int64_t n, int64_t B, int64_t R, double scale, void* workspace,
                        size_t workspace_bytes, void* stream);
/*
 * Restarted GMRES on the local operator B = PhiL x A x PhiR of ttamen_local_matvec (square: l = r1,
 * L = r2). Replaces gmres<T> (cpp/gmres.h:15-161) as amen_solve.h:349 calls it: x = B^-1 rhs from
 * x = 0, local_iterations (<= 127) Arnoldi steps per cycle, at most `resets` cycles, stopping once
 * |beta_k+1| <= tol ||rhs||. Classical Gram-Schmidt with selective re-orthogonalisation, Givens
 * rotations on the device; iters_out / flag_out (1 = converged) as the reference's gmres returns
 * them.
 */
int ttamen_gmres_solve(const double* PhiL, const double* A, const double* PhiR, const double* rhs, double* x,
                       int64_t r1, int64_t s, int64_t n, int64_t S, int64_t r2, double tol,
                       int64_t local_iterations, int64_t resets, void* workspace, size_t workspace_bytes,
                       int64_t* iters_out, int* flag_out, void* stream);
size_t ttamen_gmres_solve_workspace_size(int64_t r1, int64_t s, int64_t n, int64_t S, int64_t r2,
                                         int64_t local_iterations);
/*
 * GEMM tile autotuner (process-wide; no counterpart in the reference, whose contractions are ATen
 * tensordots). Mode 0 (default): the tile of a single-system GEMM follows a fixed rule (32x32 when a
 * 64x64 grid would give under 4 workgroups per CU, else 64x64). Mode 1: the first launch of each
 * shape bucket (element type, ceil(log2) of M, N, K and batch, unit-stride operand axes) times the
 * four tile shapes on its own operands into a scratch output of the workspace (skipped when the
 * workspace lacks the room) and the fastest serves every later launch of the bucket. Batched and
 * served solves keep their per-item rule. Results can differ in the last bits between modes (the
 * split-K plan follows the tile). ttamen_gemm_autotune(-1) returns the mode without changing it;
 * otherwise the previous mode is returned. stats: buckets tuned, launches served from the cache,
 * seconds spent tuning, served launches whose tile differs from the rule.
 */
/*
 * Form of the local operator (no counterpart in the reference, whose local operator is three ATen
 * tensordots): 0 (default) stage 1 (Phi_L x) as a strided-batched MFMA GEMM and stages 2-3 (the
 * operator core, then Phi_R) as one kernel whose stage-2 MFMA results feed the stage-3 MFMAs in
 * registers (FP64 and FP32, right rank l2 <= 144, else GEMMs; BF16 always uses GEMMs); 1 one
 * fused kernel per application (stage intermediates in LDS, partial sums over blocks of the right
 * rank added by the last workgroup in a fixed order) wherever its tiles fit the LDS, else the
 * GEMMs; 2 three strided-batched MFMA GEMMs. All give the local product to rounding; mode 1
 * measured slower on the solver's shapes and is there to be measured. mode -1 only queries.
 * Returns the previous mode (-1 on a bad argument).
 */
int ttamen_matvec_mode(int mode);
/*
 * Rank capacity of the solves planned on the calling thread (no counterpart in the reference, whose
 * C++ solver grows its tensors as the ranks grow): every interface's buffers are planned at
 * min(rmax + kickrank + kick2, the product bounds, capacity). Default 140, the largest square block
 * the LDS-resident Cholesky-QR and Jacobi SVD hold; up to 1024, QR and SVD of more columns run as
 * Householder QR and one-sided Jacobi in device memory. A sweep whose ranks outgrow the capacity
 * stops with status 4; solving again at a larger capacity gives the result the solve would have
 * given (the capacity only sizes buffers). cap <= 0 only queries. Returns the previous capacity, or
 * -1 for cap above 1024. Workspace sizes must be queried at the capacity the solve runs with.
 */
int ttamen_rank_capacity(int cap);
int ttamen_gemm_autotune(int mode);
int ttamen_gemm_autotune_stats(double* out4);
void ttamen_gemm_autotune_clear(void);
/* One line describing the current device (name, architecture, CUs, LDS per block, HBM) into buf. */
int ttamen_device_info(char* buf, size_t len);

/*
 * Thin QR of a row-major m x n matrix (FP64): Q m x k (k = min(m,n)), R k x n, A = Q R,
 * Q^T Q = I. Replaces at::linalg_qr in cpp/amen_solve.h:242,254,407,437,454 and
 * cpp/ortho.h:59. R is not triangular in general (any orthonormal Q is accepted by the
 * AMEn sweep); synchronous.
 */
int ttamen_qr(const double* A, int64_t m, int64_t n, double* Q, double* R, void* workspace,
              size_t workspace_bytes, void* stream);
/*
 * QR of a row-major m x (r + p) matrix A = [U | Z] whose first r columns are orthonormal and whose
 * last p (1..4) columns are new (FP64): Q = [U | Qz] m x (r + p), R = Q^T A, Q^T Q = I (dead columns
 * of Qz completed against all of Q). No counterpart in the reference as an entry point: it is how the
 * solver performs the enrichment QR of cpp/amen_solve.h:454 (torch QR of [u, z]), exposed for tests.
 * Requires r + p <= m and r + p <= 140; synchronous.
 */
int ttamen_qr_extend(const double* A, int64_t m, int64_t r, int64_t p, double* Q, double* R, void* workspace,
                     size_t workspace_bytes, void* stream);
/*
 * Thin SVD of a row-major m x n matrix (FP64): U m x k, s k (descending), V n x k,
 * A = U diag(s) V^T. Replaces at::linalg_svd in cpp/amen_solve.h:231,383,430 and
 * cpp/compression.h:23.
 */
int ttamen_svd(const double* A, int64_t m, int64_t n, double* U, double* s, double* V,
               void* workspace, size_t workspace_bytes, void* stream);
size_t ttamen_dense_workspace_size(int64_t m, int64_t n);

/* Dense solve B x = rhs (FP64, B n x n row-major) by blocked LU with partial pivoting.
 * Replaces at::linalg_solve of the direct local solver, cpp/amen_solve.h:326. */
int ttamen_dense_solve(const double* B, const double* rhs, double* x, int64_t n, void* workspace,
                       size_t workspace_bytes, void* stream);
size_t ttamen_dense_solve_workspace_size(int64_t n);

/*
 * Strided batched GEMM on the matrix cores (FP64), the building block of every
 * contraction above: C[b](i,j) = alpha * sum_k A[b](i,k) B[b](k,j) + beta * C[b](i,j),
 * element (i,k) of batch b at A + b*sab + i*sar + k*sac (likewise B, C).
 */
int ttamen_gemm_f64(int64_t M, int64_t N, int64_t K, int64_t batch, double alpha,
                    const double* A, int64_t sar, int64_t sac, int64_t sab, const double* B,
                    int64_t sbr, int64_t sbc, int64_t sbb, double beta, double* C, int64_t scr,
                    int64_t scc, int64_t scb, void* workspace, size_t workspace_bytes,
                    void* stream);

/*
 * Measurement hooks (no counterpart in the reference). When enabled, every full-size
 * local-operator application of ttamen_amen_solve that is not speculative is bracketed
 * by HIP events on the solver's stream; so are the FP32 applications inside the local GMRES
 * of precision 1 (only the Arnoldi steps that ran are counted). ttamen_profile_read fills
 * out4[0..7] with {FP64 applications, their event time in ms, algorithmic flops,
 * compulsory bytes, then the same four for FP32} accumulated since ttamen_profile_enable.
 * The flops count the operator-core contraction over the core's nonzero 16 x 16 tiles only
 * (the tiles the local operator multiplies), the other two contractions in full.
 * Process-global and lock-protected (batched solves serialise their timed launches).
 */
int ttamen_profile_enable(int enable);
/* Orthogonalisation of the local GMRES while profiling is enabled (no counterpart in the
 * reference): out3 = Arnoldi steps timed (HIP events around the CGS dots, update and close
 * launches of a step; with an FP64 basis the first pass's dots run inside the local operator and
 * are not timed), their total ms, and their algorithmic bytes ((2k+7) N w for step k; (k+5) N w
 * without the first pass's dots). */
int ttamen_profile_read_ortho(double* out3);
int ttamen_profile_read(double* out4);

#ifdef __cplusplus
}
#endif

#endif /* TTAMEN_H */
