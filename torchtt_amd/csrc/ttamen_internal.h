// Internal declarations shared by the HIP translation units of libttamen.
//
// Everything here is device-side plumbing for the AMEn linear solver: a bump
// arena over a caller-provided workspace, the strided batched GEMM descriptor
// that every tensor-train contraction is lowered to, and the launchers of the
// reduction / QR / SVD / GMRES kernels. Nothing in this header crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include "launch.h"
#include <algorithm>
#include <cstdint>
#include <cstddef>
#include <string>
#include <stdexcept>
#include <vector>

namespace ttamen {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  OK = 0,
  ERR_INVALID = 1,      // bad argument (shape, size, null pointer)
  ERR_WORKSPACE = 2,    // workspace too small
  ERR_HIP = 3,          // HIP runtime error
  ERR_UNSUPPORTED = 4,  // feature not built (e.g. preconditioner, rank above LDS limit)
  ERR_NUMERIC = 5,      // non-finite value detected
};

void set_last_error(const std::string& msg);

// Body of a C ABI entry point: 0 when f returns; a thrown Error gives its status code, any other
// exception ERR_INVALID; the message is kept for ttamen_last_error().
template <class F>
int abi_call(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ERR_INVALID;
  }
}

// Blocks of 256 threads for a grid-stride loop over n elements: one element per thread, at most cap blocks.
inline int grid_1d(int64_t n, int64_t cap = 4096) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap));
}

#define TT_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::ttamen::Error(::ttamen::ERR_HIP, std::string(#call) + ": " +         \
                                                   hipGetErrorString(e_));         \
  } while (0)

#define TT_CHECK(cond, code, msg)                                                  \
  do {                                                                             \
    if (!(cond)) throw ::ttamen::Error((code), (msg));                             \
  } while (0)

#define TT_LAUNCH_CHECK() TT_HIP(hipGetLastError())

// ---------------------------------------------------------------- arena
// Bump allocator over one caller-owned device buffer. Reset() rewinds; Mark /
// Release give scoped temporaries. Alignment is 256 bytes.
class Arena {
 public:
  Arena() = default;
  Arena(void* base, size_t bytes) : base_((char*)base), cap_(bytes), off_(0), peak_(0) {}
  // When base is null the arena only measures: alloc() returns fake addresses and
  // peak() reports the bytes a real run needs.
  template <typename T>
  T* alloc(size_t n) {
    size_t b = (n * sizeof(T) + 255) & ~size_t(255);
    if (b == 0) b = 256;
    if (base_ && off_ + b > cap_)
      throw Error(ERR_WORKSPACE, "workspace too small: need more than " +
                                     std::to_string(off_ + b) + " bytes, have " +
                                     std::to_string(cap_));
    char* p = (base_ ? base_ : (char*)nullptr) + off_;
    off_ += b;
    if (off_ > peak_) peak_ = off_;
    return (T*)p;
  }
  size_t mark() const { return off_; }
  size_t remaining() const { return base_ ? cap_ - off_ : ~size_t(0); }
  void release(size_t m) { off_ = m; }
  size_t peak() const { return peak_; }
  bool measuring() const { return base_ == nullptr; }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, off_ = 0, peak_ = 0;
};

// ---------------------------------------------------------------- GEMM
// C[b1,b2](i,j) = alpha * scale * sum_k A[b1,b2](i,k) B[b1,b2](k,j) + beta * C[b1,b2](i,j)
// Element (i,k) of A for batch (b1,b2) lives at a + b1*a_b1 + b2*a_b2 + i*a_r + k*a_c.
// `scale` is 1, or *scale_ptr, or 1 / *scale_ptr (scale_mode 0/1/2), read on the device.
// When skip != nullptr and *skip != 0 the launch does nothing (used by GMRES to make
// the iterations that follow convergence free without a host round trip).
struct MatRef {
  const void* p = nullptr;
  int64_t r = 0, c = 0, b1 = 0, b2 = 0;
};
struct OutRef {
  void* p = nullptr;
  int64_t r = 0, c = 0, b1 = 0, b2 = 0;
};

enum class DType : int { F64 = 0, F32 = 1, BF16 = 2 };

// Partial products of a split-K GEMM left to its consumer (GemmDesc::split_out).
struct SplitOut {
  void* partial = nullptr;  // splits x B x M x N, row-major per slice, in the GEMM's type
  int splits = 1;           // 1: K was not split and C holds the product
};

struct GemmDesc {
  int M = 0, N = 0, K = 0;
  int B1 = 1, B2 = 1;
  MatRef a, b;
  OutRef c;
  double alpha = 1.0, beta = 0.0;
  const double* scale_ptr = nullptr;
  int scale_mode = 0;
  const int* skip = nullptr;
  // Optional: when the GEMM splits K, leave the slices' partial products (allocated from the workspace
  // and kept there) for the consumer to add in slice order instead of reducing them into C; alpha = 1,
  // beta = 0 and no scale. Receives them and their count (1: C was written).
  SplitOut* split_out = nullptr;
};

// Tile autotuner of the single-system GEMMs (outside cohorts): mode 0 uses the fixed rule, mode 1
// times the four tile shapes on the first launch of each shape bucket and keeps the fastest.
int gemm_autotune_mode(int mode);  // sets the mode (if mode >= 0), returns the previous one
// While set on the calling thread, GEMMs use the fixed tile rule whatever the autotune mode (the
// tuner creates events and a scratch output; the allocation-free entry point must do neither).
bool& gemm_tuning_off();
void gemm_autotune_stats(double* out4);
void gemm_autotune_clear();

// Launch a GEMM in FP64 (A, B, C all double). `ws` provides split-K scratch.
void gemm_f64(const GemmDesc& g, Arena& ws, hipStream_t st);
// FP32 inputs and outputs, FP32 accumulation on MFMA.
void gemm_f32(const GemmDesc& g, Arena& ws, hipStream_t st);
void gemm_bf16(const GemmDesc& g, Arena& ws, hipStream_t st);

// Row-major helpers for building descriptors.
inline MatRef rm(const void* p, int64_t ld) { MatRef m; m.p = p; m.r = ld; m.c = 1; return m; }
inline MatRef rmT(const void* p, int64_t ld) { MatRef m; m.p = p; m.r = 1; m.c = ld; return m; }
inline OutRef orm(void* p, int64_t ld) { OutRef m; m.p = p; m.r = ld; m.c = 1; return m; }

// ---------------------------------------------------------------- BLAS-1 style
// All reductions are deterministic: fixed block partition, per-block partials,
// the last block to finish sums the partials in block order.
struct RedScratch {
  double* partial = nullptr;  // >= kRedBlocks * 64 doubles
  unsigned* counter = nullptr;  // zero-initialised, restored to zero by the kernels
};
constexpr int kRedBlocks = 256;

// out_slot = ||x||_2 (n doubles). guard: a non-positive result is stored as 1.0.
// mul_slot (optional): *mul_slot *= result (after the guard).
void norm2(const double* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st,
           bool guard = false, double* mul_slot = nullptr, const int* skip = nullptr);
// out_slot = ||x||_2 as norm2 does, then x *= 1 / out_slot (by the kernel's last block: for small x).
void norm2_scale_inv(double* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st, bool guard);
void norm2_f32(const float* x, int64_t n, double* out_slot, RedScratch rs, hipStream_t st,
               const int* skip = nullptr);
// out_slot = ||x - y||_2 / (den_slot ? *den_slot : 1)
void diffnorm(const double* x, const double* y, int64_t n, const double* den_slot, double* out_slot,
              RedScratch rs, hipStream_t st);
// x *= 1 / *slot  (slot holds an already-guarded positive number), optional extra scalar factor
void scale_inv(double* x, int64_t n, const double* slot, hipStream_t st);
void scale_by(double* x, int64_t n, const double* slot, hipStream_t st);
// y = a*x + b*y with host scalars
void axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t st,
           const int* skip = nullptr);
void axpby_f32(double a, const float* x, double b, float* y, int64_t n, hipStream_t st,
               const int* skip = nullptr);
// y (FP64) += x (FP32)
void add_f32_to_f64(const float* x, double* y, int64_t n, hipStream_t st);
// y (FP64) = base (FP64) + x (FP32)
void add_f32_to_f64_into(const float* x, const double* base, double* y, int64_t n, hipStream_t st);
// z = x - y
void sub(const double* x, const double* y, double* z, int64_t n, hipStream_t st);
void copy(const double* x, double* y, int64_t n, hipStream_t st);
void fill(double* x, int64_t n, double v, hipStream_t st);
void zero_words(void* x, int64_t nwords, hipStream_t st);  // nwords 32-bit words set to 0
void eye(double* x, int64_t m, int64_t n, hipStream_t st);  // row-major m x n identity
// Strided 2-D copy: dst(i,j) = src(i*sr + j*sc), dst row-major with leading dim ldd
void copy2d(const double* src, int64_t sr, int64_t sc, double* dst, int64_t ldd, int64_t m,
            int64_t n, hipStream_t st);
void cast_f64_f32(const double* x, float* y, int64_t n, hipStream_t st);
void cast_f32_f64(const float* x, double* y, int64_t n, hipStream_t st);
// Permute an operator core A[s][m][n][S] -> P[m][S][s][n] (the layout the local
// matvec's middle contraction reads as a row-major (n*S) x (s*n) matrix).
void permute_core_A(const double* A, double* P, int s, int n, int S, hipStream_t st);
// Rectangular operator core A[s][nm][nk][S] -> P[nm][S][s][nk].
void permute_core_A_rect(const double* A, double* P, int s, int nm, int nk, int S, hipStream_t st);
void permute_core_A_f32(const double* A, float* P, int s, int n, int S, hipStream_t st);
// Dense local operator B[(l,m,L),(r,n,R)] = sum_{s,S} PhiL[l,s,r] A[s,m,n,S] PhiR[L,S,R]
// (row-major with leading dimension ld >= r1*n*r2)
void assemble_local(const double* PhiL, const double* A, const double* PhiR, double* B, int64_t ld,
                    int r1, int s, int n, int S, int r2, hipStream_t st);

// ---------------------------------------------------------------- scalar program
// Tiny single-thread kernels for the sweep's norm bookkeeping.
// nrmsc = nrmsc * a * x / b   (mode 0)   or   nrmsc = nrmsc * b / (a * x)   (mode 1)
void nrmsc_update(double* S, int i_nrmsc, int i_a, int i_x, int i_b, int mode, hipStream_t st);
// S[dst] = max(S[dst], S[src] / S[den])
void scalar_maxdiv(double* S, int dst, int src, int den, hipStream_t st);
void scalar_set(double* S, int dst, double v, hipStream_t st);
// S[dst] = S[a] / S[b]
void scalar_div(double* S, int dst, int a, int b, hipStream_t st);

// ---------------------------------------------------------------- host transfers
// Queue a copy of `bytes` (a multiple of 4) of device memory into pinned host memory.
void post_to_host(void* host, const void* dev, size_t bytes, hipStream_t st);
// Wait until everything queued so far has run (inside a cohort: until the cohort has
// run this system's recorded work).
void sync_point(hipStream_t st);
// post_to_host + sync_point.
void read_to_host(void* host, const void* dev, size_t bytes, hipStream_t st);

// ---------------------------------------------------------------- dense linear algebra
constexpr int kMaxSmall = 140;   // largest square block kept in LDS (FP64)
constexpr int kMaxRank = 1024;   // largest rank of the device-memory QR / SVD (linalg_large.hip)

// The rank every interface of a solve is planned at when rmax allows more (workspace sizing only:
// results do not depend on it). kMaxSmall unless raised on the calling thread, at most kMaxRank;
// QR and SVD above kMaxSmall columns run in device memory (qr_large, svd_large).
int rank_capacity();
int set_rank_capacity(int cap);  // returns the previous value; cap <= 0 only queries
// Householder QR (m >= n) and one-sided Jacobi SVD (k x k) in device memory, for n, k > kMaxSmall;
// same outputs as qr() and the SVD of its square factor.
int qr_large(const double* A, int64_t sr, int64_t sc, int64_t m, int64_t n, double* Q, double* R, Arena& ws,
             hipStream_t st);
void svd_large(const double* M, int64_t sr, int64_t sc, int k, double* s, double* Ur, double* V, int64_t ldv,
               Arena& ws, hipStream_t st);

// Thin QR of a tall m x n (m >= n) matrix given by strides, via shifted
// Cholesky-QR with up to four passes and triangular solves; Q is m x n row-major
// (ld n), R = Q^T A is n x n row-major. If m < n: Q = I_m, R = A (m x n).
// With r_split, R's split-K partials may be left to the caller (SplitOut: R is then
// their sum in slice order, not written; splits = 1 when R was written).
// Returns the number of columns of Q (min(m, n)).
int qr(const double* A, int64_t sr, int64_t sc, int64_t m, int64_t n, double* Q, double* R,
       Arena& ws, hipStream_t st, SplitOut* r_split = nullptr);
// QR of A = [U | Z] (m x (r + p), row-major, ld lda) whose first r columns U are orthonormal and whose
// last p columns Z are few (p <= 4): Q = [U | Qz] (m x (r + p), ld r + p), Qz from a block Gram-Schmidt
// pass of Z against U, a QR of the remainder, a second pass of those columns and a second QR, any dead column
// of Qz completed against all of Q's live columns; R = Q^T A ((r + p) x (r + p)). Returns false, doing nothing, when the shape is outside that
// (p > 4, r + p > m or above the LDS-resident QR's 140 columns).
bool qr_extend(const double* A, int64_t lda, int64_t m, int64_t r, int64_t p, double* Q, double* R, Arena& ws,
               hipStream_t st);
// Thin SVD of an m x n matrix (strided): U (m x k row-major, ld k), s (k), V (n x k
// row-major, ld k), k = min(m, n). Singular values descending.
int svd(const double* M, int64_t sr, int64_t sc, int64_t m, int64_t n, double* U, double* s,
        double* V, Arena& ws, hipStream_t st);
// The reference's truncation rule (cpp/ortho.h:16-44, torchtt/_decomposition.py:289-324): the
// smallest rank whose discarded tail of the descending singular values s[0..n) has 2-norm below
// eps; 1 for a zero vector, n - 1 for eps <= 0. Host arrays.
int rank_chop(const double* s, int n, double eps);
// Solve B x = rhs for the augmented n x (n+1) row-major matrix Aug = [B | rhs]
// (overwritten) by blocked LU with partial pivoting; x receives the solution.
void lu_solve_aug(double* Aug, int n, double* x, Arena& ws, hipStream_t st);
void lu_init();
int lu_max_order();

// ---------------------------------------------------------------- local operator
// Local AMEn operator, y = PhiL x A x PhiR applied to nvec vectors:
//   y[v][l,m,L] = sum PhiL[l,s,r] A[s,m,n,S] PhiR[L,S,R] x[v][r,n,R]
// PhiL is l1 x s x r1, PhiR is l2 x S x r2, x is r1 x n x r2, y is l1 x n x l2.
// Apm is the permuted core from permute_core_A.
// With Apk (the permuted core packed by pack_frag_a) the operator runs as one fused kernel
// (kernels_matvec.h) where the shape allows, else as three GEMMs on Apm. cnt (optional): cnt_cap
// zeroed counters the fused kernel may use (it leaves them at zero); without them it takes
// counters from the workspace and clears them first.
// Tile masks of a permuted operator core Apm[(m, S)][k] (nm S rows of K = s n values): word mb S + S'
// has bit j set when rows 16 mb .. 16 mb + 15 (below nm) of S' hold a nonzero at some k in [16 j, 16 j + 16).
// Defined for K <= kTileMaskMaxK (64 steps); the local operator skips the steps whose bit is clear.
constexpr int kTileMaskMaxK = 1024;
inline size_t tile_mask_words(int64_t nm, int64_t S) { return (size_t)((nm + 15) / 16) * S; }
void op_tile_masks(const double* Apm, int nm, int S, int K, uint64_t* out, hipStream_t st);
void op_tile_masks(const float* Apm, int nm, int S, int K, uint64_t* out, hipStream_t st);
// Whether the local operator's current form skips zero operator tiles for this shape (the default form's
// stages-2/3 kernel with masks; not the BF16 products, not the other forms).
bool matvec_skips_zero_tiles(int l2, int K, bool bf16);

struct LocalOp {
  const double* PhiL = nullptr;
  const double* Apm = nullptr;
  const double* PhiR = nullptr;
  int l1 = 1, s = 1, r1 = 1, n = 1, S = 1, l2 = 1, r2 = 1;
  int n_out = 0;  // mode size of y for a rectangular operator core A[s][n_out][n][S] (0: n)
  const double* Apk = nullptr;
  unsigned* cnt = nullptr;
  int cnt_cap = 0;
  const uint64_t* tmask = nullptr;  // tile masks of Apm (op_tile_masks), or null (K above kTileMaskMaxK)
  double tdens = 1.0;  // share of the operator entries in nonzero tiles (host bookkeeping of profiled flops)
};
// Optional first pass of the GMRES orthogonalisation fused into the local operator (one vector, the
// default form, FP64/FP32): each workgroup of the stages-2/3 kernel writes its share of Q_i . y (i < kc,
// Q_i = Q + i N in the operator's precision) and of y . y (row kc) to P[i * kGmresMaxBlocks + blk]. On
// return nb is the number of shares per row, or 0 when the form cannot fuse them (the caller then
// computes the dots itself).
struct MvDots {
  const void* Q = nullptr;
  int64_t N = 0;
  int kc = 0;
  double* P = nullptr;
  int nb = 0;
};
void local_matvec(const LocalOp& op, const double* x, double* y, int nvec, Arena& ws,
                  hipStream_t st, const int* skip = nullptr, MvDots* dots = nullptr);
size_t local_matvec_ws(const LocalOp& op, int nvec);

struct LocalOpF32 {
  const float* PhiL = nullptr;
  const float* Apm = nullptr;
  const float* PhiR = nullptr;
  int l1 = 1, s = 1, r1 = 1, n = 1, S = 1, l2 = 1, r2 = 1;
  bool bf16 = false;  // BF16 products (operands rounded to BF16), FP32 accumulation
  const float* Apk = nullptr;
  unsigned* cnt = nullptr;
  int cnt_cap = 0;
  const uint64_t* tmask = nullptr;  // tile masks of Apm, or null
};
// Elements of the operator core Apm (MS x KS, MS = n S, KS = s n) packed in MFMA fragment order.
size_t frag_a_elems(int MS, int KS);
void pack_frag_a(const double* Apm, double* Apk, int MS, int KS, hipStream_t st);
void pack_frag_a_f32(const float* Apm, float* Apk, int MS, int KS, hipStream_t st);
// Local-operator form: 0 three GEMMs (default), 1 the fused kernel where its plan allows it; sets the
// mode if mode >= 0 and returns the previous one.
int matvec_mode(int mode);
// Workspace of the fused local operator's partial sums and counters at these sizes (0 if none).
size_t local_matvec_fused_bytes(int l1, int s, int r1, int n, int S, int l2, int r2, int nvec, size_t elem);
void local_matvec_f32(const LocalOpF32& op, const float* x, float* y, int nvec, Arena& ws,
                      hipStream_t st, const int* skip = nullptr, MvDots* dots = nullptr);

// Interface updates (FP64). With inv_scale, the result is multiplied by 1 / *inv_scale (a device
// scalar) in the last contraction's epilogue, exactly as a separate scale_inv would round it.
// fwd: Pn[L,S,R] = sum Phi[l,s,r] xL[l,M,L] A[s,M,N,S] xR[r,N,R]
void phi_fwd_A(const double* Phi, const double* xL, const double* Apm, const double* xR,
               double* Pn, int l, int s, int r, int n, int S, int L, int R, Arena& ws,
               hipStream_t st, const double* inv_scale = nullptr);
// bck: Pp[l,s,r] = sum xL[l,M,L] Phi[L,S,R] A[s,M,N,S] xR[r,N,R]
void phi_bck_A(const double* Phi, const double* xL, const double* A, const double* xR,
               double* Pp, int l, int s, int r, int n, int S, int L, int R, Arena& ws,
               hipStream_t st, const double* inv_scale = nullptr);
// The same for a rectangular operator core (M of size nm on the xL side, N of size nk on the
// xR side; Apm from permute_core_A_rect, A unpermuted).
void phi_fwd_A_rect(const double* Phi, const double* xL, const double* Apm, const double* xR,
                    double* Pn, int l, int s, int r, int nm, int nk, int S, int L, int R,
                    Arena& ws, hipStream_t st, const double* inv_scale = nullptr);
void phi_bck_A_rect(const double* Phi, const double* xL, const double* A, const double* xR,
                    double* Pp, int l, int s, int r, int nm, int nk, int S, int L, int R,
                    Arena& ws, hipStream_t st, const double* inv_scale = nullptr);
// The A-interface updates with FP32 operands and result (FP32 accumulation on the matrix cores).
void phi_fwd_A_f32(const float* Phi, const float* xL, const float* Apm, const float* xR, float* Pn, int l, int s,
                   int r, int n, int S, int L, int R, Arena& ws, hipStream_t st);
void phi_bck_A_f32(const float* Phi, const float* xL, const float* A, const float* xR, float* Pp, int l, int s,
                   int r, int n, int S, int L, int R, Arena& ws, hipStream_t st);
// The forward update before its contraction with xL: t2[l][M][S][R] = sum Phi[l,s,r]
// A[s,M,N,S] xR[r,N,R] (l x nm x S x R; Apm from permute_core_A_rect).
void phi_fwd_half_rect(const double* Phi, const double* Apm, const double* xR, double* t2, int l, int s,
                       int r, int nm, int nk, int S, int R, Arena& ws, hipStream_t st);
// fwd rhs: Pn[B,R] = sum Pb[b,r] bc[b,n,B] x[r,n,R]   (times 1 / *inv_scale, if given)
void phi_fwd_rhs(const double* Pb, const double* bc, const double* x, double* Pn, int b, int r,
                 int n, int B, int R, Arena& ws, hipStream_t st, const double* inv_scale = nullptr);
// bck rhs: Pp[b,r] = sum bc[b,n,B] Pb[B,R] x[r,n,R]   (times 1 / *inv_scale, if given)
void phi_bck_rhs(const double* Pb, const double* bc, const double* x, double* Pp, int b, int r,
                 int n, int B, int R, Arena& ws, hipStream_t st, const double* inv_scale = nullptr);
// rhs[r,n,R] = scale * sum PL[b,r] bc[b,n,B] PR[B,R]   (scale = *scale_ptr, if given)
void assemble_rhs(const double* PL, const double* bc, const double* PR, double* out, int b,
                  int r, int n, int B, int R, const double* scale_ptr, Arena& ws,
                  hipStream_t st);

}  // namespace ttamen
