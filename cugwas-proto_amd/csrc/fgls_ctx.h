// fgls_ctx.h -- the device context shared by the translation units of the
// HIP layer (internal; callers see only the opaque fgls_hip_ctx).
#ifndef FGLS_CTX_H
#define FGLS_CTX_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fgls_hip.h"

struct fgls_adjoint_state;
// Releases the adjoint buffers (fgls_adjoint.hip); NULL is allowed.
void fgls_adjoint_free(fgls_adjoint_state* a);
// Marks the adjoint state stale (a new trait was set); NULL is allowed.
void fgls_adjoint_invalidate(fgls_adjoint_state* a);

struct fgls_hip_ctx {
  int device;
  int n, wXL, wXR, p;
  int ldn;        // padded rows (multiple of 128)
  int nblk;       // 64-row diagonal blocks
  int RT;         // 128-row TRMM tiles
  int max_block;
  int reduced;    // SNP count of the last block_reduce (finalize must match)
  int last_rt;    // row tiles of the partials written by the last block_reduce
  bool have_trait;
  int mode;       // FGLS_MODE_FP64 or FGLS_MODE_MIXED_I8
  int RT8;        // 32-row tiles of the int8 product
  int part_rt;    // row-tile capacity of `part`
  int8_t* A8;     // slices of L^-1 (mixed mode)
  int8_t* X8t;    // tiled genotypes of the current block (mixed mode)
  size_t X8t_cap;
  double* invsc;  // per-row 1 / sc_i of the slices (ldn + 1 doubles: the last is build_slices' loss cell)
  int nslice;     // digits of the slices in A8: FGLS_NSLICE or FGLS_TOP_NSLICE (five-digit form)
  int slice_form; // FGLS_SLICES_AUTO or FGLS_SLICES_SIX
  double slice_loss;  // error bound of the five-digit form over the six-digit one (-1: not tried)
  unsigned int* bad;  // count of genotypes that are not int8-exact
  double* L;
  double* Linv;
  double* Dinv;
  double* W;
  double* Wtmp;   // [X_L | y] of the trait, zero padded (input of W = L^-1 Wtmp)
  double* trait;
  double* part;
  double* z;      // lazily allocated (generic per-SNP path, adjoint)
  fgls_adjoint_state* adj;  // lazily allocated (adjoint)
  unsigned long long* err;
  int* status;
  // the factor recursion's second stream (the small nodes' X = L21 T11 runs
  // there beside the node's right child) and its fork / join events;
  // created on first use
  hipStream_t side;
  hipEvent_t ev_fork, ev_join;
};

// Calls FN<p>(...) for p = 1..16.
#define FGLS_P_SWITCH(p, FN, ...)                    \
  switch (p) {                                  \
    case 1: FN<1>(__VA_ARGS__); break;          \
    case 2: FN<2>(__VA_ARGS__); break;          \
    case 3: FN<3>(__VA_ARGS__); break;          \
    case 4: FN<4>(__VA_ARGS__); break;          \
    case 5: FN<5>(__VA_ARGS__); break;          \
    case 6: FN<6>(__VA_ARGS__); break;          \
    case 7: FN<7>(__VA_ARGS__); break;          \
    case 8: FN<8>(__VA_ARGS__); break;          \
    case 9: FN<9>(__VA_ARGS__); break;          \
    case 10: FN<10>(__VA_ARGS__); break;        \
    case 11: FN<11>(__VA_ARGS__); break;        \
    case 12: FN<12>(__VA_ARGS__); break;        \
    case 13: FN<13>(__VA_ARGS__); break;        \
    case 14: FN<14>(__VA_ARGS__); break;        \
    case 15: FN<15>(__VA_ARGS__); break;        \
    case 16: FN<16>(__VA_ARGS__); break;        \
    default: break;                             \
  }

namespace fgls {

constexpr int TM = 128;   // FP64 TRMM tile rows
constexpr int TN = 128;   // FP64 TRMM tile columns
constexpr int PMAX = 16;  // largest p of the per-SNP kernels

// Sets the calling thread's error message; returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int set_device(const fgls_hip_ctx* c);
// Offset of res_sigma in the per-trait scalars (ctx->trait).
__host__ __device__ inline int trait_res_off(int wXL) { return wXL + wXL * wXL; }

// Z = A X, column-major out (ld ldz): A = ctx->Linv (lower) or, if dense, a
// full row-major ldn x ldn matrix.
void launch_product_store(const fgls_hip_ctx* c, const double* A, bool dense, const double* X,
                          int64_t ldx, int ncols, double* Z, int64_t ldz, hipStream_t st);
// Z = L^-1 X (n columns) for a symmetric result: only the 64 x 64 blocks on
// or below the diagonal of Z are written.
void launch_product_store_lower_half(const fgls_hip_ctx* c, const double* X, int64_t ldx,
                                     double* Z, int64_t ldz, hipStream_t st);
// Z = A X for an upper triangular row-major ldn x ldn A (zeros below the
// diagonal are skipped).
void launch_product_store_upper(const fgls_hip_ctx* c, const double* A, const double* X,
                                int64_t ldx, int ncols, double* Z, int64_t ldz, hipStream_t st);
// Y = A X for a dense row-major A; per 128-row tile and column, the sums of
// y * X[row, col] and y * ctx->W[:, k] go to ctx->part (p + 1 per column).
void launch_product_dot(const fgls_hip_ctx* c, const double* A, const double* X, int64_t ldx,
                        int count, hipStream_t st);
// Dense Z = A X (A: full row-major ldn x ldn, rows >= n zero), column-major
// out (ld ldz >= ldn).
void launch_dense_store(int n, int ldn, const double* A, const double* X, int64_t ldx, int ncols,
                        double* Z, int64_t ldz, hipStream_t st);
// Dense FP64 product of a block with ntr groups of p + 1 weight columns
// (group k at W + k (p+1) ldw: column 0 weighs z*z, the others z); per
// 128-row tile, trait and SNP the p + 1 sums go to
// part[((rt ntr + k)(p+1) + q) ldp + col].
void launch_eig_reduce(int p, int n, int ldn, const double* A, const double* X, int64_t ldx,
                       int count, const double* W, int64_t ldw, int ntr, double* part, int64_t ldp,
                       hipStream_t st);
// b and V of every SNP and trait from partials laid out as launch_eig_reduce's
// (RT row tiles; trait k: scalars at trait + k trait_ld, outputs at dB + k ldb
// and dV + k ldv, failure record err[k]).
void launch_finalize_multi(int p, int count, int RT, const double* part, int64_t ldp, int ntr,
                           const double* trait, int64_t trait_ld, double* dB, int64_t ldb,
                           double* dV, int64_t ldv, unsigned long long* err, int64_t base,
                           hipStream_t st);
// B_t, V_tl, b_L and res_sigma of one trait from W = [X~_L | y~] (n x (wXL+1),
// ld ldw) into trait (TRAIT_LD doubles); status[0] = 0 or the LAPACK-style
// info of the Cholesky of V_tl.
void launch_trait_stats(int n, int wXL, const double* W, int64_t ldw, double* trait, int* status,
                        hipStream_t st);
// int8 digit slices and row scales of a row-major ldn x ldn matrix.
void launch_slices(int ldn, const double* A, int8_t* A8, double* invsc, hipStream_t st);
// Grows a device buffer to at least `need` bytes (synchronizes the device
// when it reallocates).
int ensure_buffer(int8_t** buf, size_t* cap, size_t need);
// Doubles per trait of the per-trait scalars.
constexpr int TRAIT_LD = 2 * PMAX + PMAX * PMAX + 8;

// b and V of every SNP of a block from its transformed columns Z.
void launch_snp_stage_any(const fgls_hip_ctx* c, const double* Z, int64_t ldz, int count,
                          double* dB, double* dV, int64_t base, hipStream_t st);
// k_snp_stage with row weights dwt of the Z'Z sums (the eigen variant with
// wXR > 1): p = wXL + wXR, W the weighted [Z'X_L, Z'y] of one trait.
void launch_snp_stage_weighted_any(int p, int n, int count, int wXL, int wXR, const double* Z, int64_t ldz,
                                   const double* W, int64_t ldw, const double* dwt, const double* trait,
                                   double* dB, double* dV, unsigned long long* err, int64_t base,
                                   hipStream_t st);

// T = L^-1 (row-major, ctx->Linv) by recursive inversion on MFMA products;
// with `factor`, first the Cholesky factor of the matrix in ctx->L (both
// triangles), written over it (info of a failed pivot in ctx->status[1]).
// Without it, ctx->L holds the factor in its column-major lower triangle;
// its strict upper triangle is used as workspace.
int chol_inv(fgls_hip_ctx* c, bool factor, hipStream_t st);

// The per-trait precompute from the factor already in ctx->L (L^-1, W,
// B_t, V_tl, res_sigma, slices); synchronizes `st`.  With potrf_info set,
// ctx->status[1] holds the device factorisation's LAPACK-style info.
int finish_trait(fgls_hip_ctx* c, const double* dXL, int64_t ldxl, const double* dy, hipStream_t st,
                 const int* potrf_info);

}  // namespace fgls

#endif
