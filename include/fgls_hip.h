/*
 * fgls_hip.h -- C-ABI of the MI355X (gfx950) device layer of the Cholesky
 * FGLS solver.
 *
 * Everything here takes plain pointers and sizes.  Matrices are column-major
 * FP64 unless stated.  Pointers named d* are device (HBM) pointers; `stream`
 * is a hipStream_t passed as void* (NULL = the legacy default stream).
 *
 * Each entry point replaces a host BLAS/LAPACK call or a cuBLAS call of the
 * reference's GPU solver (fgls_chol_gpu.c), as cited per function.
 *
 * Error convention: every function returns FGLS_OK (0) or a negative code;
 * fgls_hip_last_error() gives a message for the calling thread.  The host
 * solver (fgls_chol_gpu) turns codes into the reference's
 * "[Error]: <msg>" + exit(EXIT_FAILURE) behaviour (common.c:11-17).
 */
#ifndef FGLS_HIP_H
#define FGLS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGLS_OK            0
#define FGLS_ERR_ARG      -1   /* invalid argument (sizes, strides, NULL)      */
#define FGLS_ERR_HIP      -2   /* a HIP runtime call failed                    */
#define FGLS_ERR_NOMEM    -3   /* device or pinned allocation failed           */
#define FGLS_ERR_NOTSPD   -4   /* a Cholesky factorisation met a non-positive pivot */
#define FGLS_ERR_IO       -5   /* file open / read / write failure (host layer) */
#define FGLS_ERR_STATE    -6   /* call order violated (e.g. no trait set)      */

/* Arithmetic of the per-SNP hot path (fgls_hip_ctx_set_mode). */
#define FGLS_MODE_FP64     0   /* FP64 genotypes, FP64 MFMA                    */
#define FGLS_MODE_MIXED_I8 1   /* int8 genotypes; L^-1 as 5 or 6 int8 digit slices
                                  (balanced radix 254); exact int32 accumulation
                                  on the int8 MFMA; FP64 per-row recombination
                                  and per-SNP solve; n <= 133,000 */

typedef struct fgls_hip_ctx fgls_hip_ctx;

/* Message describing the last failure on the calling thread. */
const char *fgls_hip_last_error(void);

/* Number of visible HIP devices (0 when there is none); never fails. */
int fgls_hip_device_count(void);

/*
 * Create a device context for problems of n individuals, wXL covariate
 * columns and wXR SNP columns per SNP (p = wXL + wXR), processing at most
 * max_block SNPs per solve call.  Allocates all device state up front
 * (replaces the cudaMalloc of L and the X_R buffers, fgls_chol_gpu.c:230-265,
 * and the stream creation, :212-228).
 */
int fgls_hip_ctx_create(fgls_hip_ctx **out, int n, int wXL, int wXR,
                        int max_block, int device);
int fgls_hip_ctx_destroy(fgls_hip_ctx *ctx);

/* Leading dimension (rows, a multiple of 128) that the context pads n to. */
int fgls_hip_ctx_ldn(const fgls_hip_ctx *ctx);

/*
 * Per-trait precompute from a factor already in device memory.
 * dL: n x n lower Cholesky factor of M (upper triangle ignored), ld ldl.
 * dXL: n x wXL covariates (ld ldxl), dy: the trait vector (n).
 * Computes on the device, in FP64:
 *   - L^-1 (recursive inversion: 64x64 diagonal blocks inverted in one
 *     wave each, the off-diagonal blocks as FP64 MFMA products),
 *   - the inverses of the 64x64 diagonal blocks of L (for fgls_hip_trsm),
 *   - X~_L = L^-1 X_L and y~ = L^-1 y,
 *   - B_t = X~_L' y~, V_tl = lower(X~_L' X~_L) and the null-model residual
 *     variance res_sigma = |L^-1 (y - X_L b_L)|^2 / (n - wXL).
 * Replaces: cublasSetVector(L) fgls_chol_gpu.c:434-445; dtrsm(XL) :447-449;
 * dtrsv(y), dgemv, dsyrk :461-468; the res_sigma block :469-493.
 * Synchronises `stream`; returns FGLS_ERR_NOTSPD when X~_L' X~_L is not SPD.
 */
int fgls_hip_set_trait(fgls_hip_ctx *ctx, const double *dL, int64_t ldl,
                       const double *dXL, int64_t ldxl, const double *dy,
                       void *stream);

/* Copy the per-trait scalars back: B_t (wXL), V_tl (wXL*wXL, col-major,
 * upper triangle 0) and res_sigma.  Any output pointer may be NULL. */
int fgls_hip_get_trait(fgls_hip_ctx *ctx, double *hBt, double *hVtl,
                       double *hres_sigma);

/*
 * Fused per-block solve (FP64): for `count` SNPs whose columns are in
 * dXR (column-major, ld ldxr >= n, ldxr % 4 == 0, SNP i's wXR columns at
 * column i*wXR), compute b_i (p values) and V_i (p*p, column-major, lower
 * triangle = res_sigma * S_i^-1 with sqrt on the diagonal, upper = 0) into
 * dB (count*p) and dV (count*p*p), the layout of B.out/V.out
 * (fgls_chol.c:458-474).
 * Replaces cublasDtrsm fgls_chol_gpu.c:531-557, the D2H of L^-1 X_R
 * :564-576 and the CPU OpenMP per-SNP stage :619-687.
 * Asynchronous on `stream`.  A non-SPD per-SNP system is recorded in the
 * context (first failing SNP, counted from snp_base) and reported by
 * fgls_hip_check().
 */
int fgls_hip_solve_block(fgls_hip_ctx *ctx, const double *dXR, int64_t ldxr,
                         int count, double *dB, double *dV, int64_t snp_base,
                         void *stream);

/*
 * The two stages of fgls_hip_solve_block, for callers that overlap or time
 * them separately.  block_reduce runs the triangular MFMA product of the
 * block with the per-row-tile reductions (X~_L' z, z'z, z'y~) kept in the
 * context (for wXR > 1: z itself); block_finalize turns them into b_i and
 * V_i.  A finalize must follow the reduce of the same `count` on the same
 * stream before the next reduce.
 */
int fgls_hip_block_reduce(fgls_hip_ctx *ctx, const double *dXR, int64_t ldxr,
                          int count, void *stream);
int fgls_hip_block_finalize(fgls_hip_ctx *ctx, int count, double *dB,
                            double *dV, int64_t snp_base, void *stream);

/*
 * Mixed mode (no counterpart in the reference, whose arithmetic is FP64 only).
 * set_mode(FGLS_MODE_MIXED_I8) allocates ldn*ldn*6 bytes of slices of L^-1
 * (built at set_trait, or at once if a trait is set).
 * The genotypes of a block are int8, column-major, ld ldx8 >=
 * ceil(n/64)*64 with ldx8 % 16 == 0 (rows n..ld-1 are ignored), SNP i's
 * wXR columns at column i*wXR as for fgls_hip_solve_block.  wXR = 1 reduces
 * in the int8 product's epilogue; wXR > 1 stores L^-1 X_R of the block
 * (ldn x count*wXR FP64, allocated at set_mode) for the per-SNP stage, and
 * dosage digits need wXR = 1.  b and V
 * have the layout of fgls_hip_solve_block; their error against the FP64
 * path is set by the slice count (b within 1e-9 of max(|b|, se), see the tests).
 */
int fgls_hip_ctx_set_mode(fgls_hip_ctx *ctx, int mode);
/*
 * Digit form of the slices of L^-1 in the mixed mode (no counterpart in the
 * reference).  FGLS_SLICES_AUTO (the default): row i of L^-1 is scaled by
 * its largest entry outside its diagonal 64-column block, and five digits
 * per entry plus one top digit on that diagonal block are kept whenever the
 * resulting error bound is at most 16 times the six-digit form's (the
 * "loss": 254 max_i (off-block maximum / row maximum) of row i) and
 * n <= 116,000; five-sixths of the int8 work of six digits.  Otherwise, and
 * always with FGLS_SLICES_SIX, six digits of the row scaled by its maximum.
 * set_slice_form takes effect at the next set_trait (and rebuilds the slices
 * at once if a trait is set and the context is in the mixed mode).
 * slice_info gives the digits in use (5 or 6; 0 before any slices exist)
 * and the loss measured at the last build (-1 if the five-digit form was not
 * tried).
 */
#define FGLS_SLICES_AUTO 0
#define FGLS_SLICES_SIX  6
int fgls_hip_ctx_set_slice_form(fgls_hip_ctx *ctx, int form);
int fgls_hip_ctx_slice_info(const fgls_hip_ctx *ctx, int *nslice, double *loss);
int fgls_hip_block_reduce_i8(fgls_hip_ctx *ctx, const int8_t *dX8, int64_t ldx8,
                             int count, void *stream);
int fgls_hip_solve_block_i8(fgls_hip_ctx *ctx, const int8_t *dX8, int64_t ldx8,
                            int count, double *dB, double *dV, int64_t snp_base,
                            void *stream);
/* FP64 -> int8 genotypes (ld ldx8, rows n..ldx8-1 zeroed); values that are
 * not integers in [-127, 127] are counted into *dbad (device counter,
 * caller-cleared) and written as 0.  Asynchronous. */
int fgls_hip_to_i8(const double *dX, int64_t ldx, int n, int ncols,
                   int8_t *dX8, int64_t ldx8, unsigned int *dbad, void *stream);
/*
 * Mixed mode for dosages (no counterpart in the reference): values with a
 * fixed number of decimals, e.g. imputed dosages rounded to 0.001, as
 * v = x * scale = 128 hi + lo with two int8 digits lo in [0, 127] and
 * hi in [-127, 127] (|v| <= 16,256).  dX8 holds two planes of count columns
 * (ld ldx8): lo at dX8, hi at dX8 + count * ldx8.  The product costs twice
 * the int8 work of genotypes; b and V as fgls_hip_solve_block_i8.
 * to_dosage_i8 builds the planes from FP64 values (ncols columns, planes
 * ncols * ldx8 apart), counting into *dbad the values out of range and those
 * whose y = x * scale is farther than 8 eps max(1, |y|) (eps = 2^-52) from
 * an integer: exact multiples of 1/scale pass, anything else is an error,
 * never rounded.  Asynchronous.
 */
int fgls_hip_to_dosage_i8(const double *dX, int64_t ldx, int n, int ncols, double scale,
                          int8_t *dX8, int64_t ldx8, unsigned int *dbad, void *stream);
int fgls_hip_block_reduce_dosage(fgls_hip_ctx *ctx, const int8_t *dX8, int64_t ldx8,
                                 int count, double scale, void *stream);
int fgls_hip_solve_block_dosage(fgls_hip_ctx *ctx, const int8_t *dX8, int64_t ldx8,
                                int count, double scale, double *dB, double *dV,
                                int64_t snp_base, void *stream);

/*
 * Tiled genotypes: the layout the int8 product reads (1 KB blocks, one per 16
 * SNP columns and 64 individuals, in the MFMA operand's lane order).  The
 * column-major entry points above tile each block into the context's own
 * buffer first; callers that keep genotypes resident, or solve several
 * traits per block, tile once with fgls_hip_tile_i8 (dX8T of
 * fgls_hip_x8t_bytes(n, count, planes) bytes, 16-byte aligned; planes = 2 for
 * dosage digit planes laid out as for fgls_hip_solve_block_dosage) and pass
 * dX8T to the *_i8t entry points (planes = 1 genotypes, 2 dosages with
 * x = (128 hi + lo) / scale; scale is ignored for planes = 1).  For
 * wXR > 1 the tiles hold the block's count*wXR columns (tile them with
 * count*wXR columns) and `count` of the *_i8t entry points is SNPs.  No
 * counterpart in the reference.
 */
size_t fgls_hip_x8t_bytes(int n, int count, int planes);
int fgls_hip_tile_i8(const int8_t *dX8, int64_t ldx8, int n, int count, int planes,
                     int8_t *dX8T, void *stream);
int fgls_hip_block_reduce_i8t(fgls_hip_ctx *ctx, const int8_t *dX8T, int count, int planes,
                              double scale, void *stream);
int fgls_hip_solve_block_i8t(fgls_hip_ctx *ctx, const int8_t *dX8T, int count, int planes,
                             double scale, double *dB, double *dV, int64_t snp_base,
                             void *stream);

/*
 * 2-bit tiled genotypes (hard calls 0, 1, 2): the int8 product with its
 * genotype operand loaded 4 values per byte and unpacked in registers, a
 * quarter of the bytes of fgls_hip_tile_i8's layout per MFMA (bitwise the
 * same b and V).  Layout: one 256-byte block per (16 SNP columns, 64
 * individuals), fgls_hip_x2t_bytes(n, count) bytes, 4-byte aligned; values
 * 0..3 only.  tile_u2 builds it from column-major int8 genotypes (as for
 * fgls_hip_solve_block_i8), tile_u2_packed from the 2-bit file layout of
 * fgls_hip_u2_to_i8 (ceil(n/4) bytes per column at dP + c*ldp); a value
 * outside 0..3 (int8 input) or a code 3 (packed input) in rows < n is
 * counted into *dbad (device counter, caller-cleared) and written as 0.
 * reduce/solve take the tiles (count*wXR columns) in the mixed mode as the *_i8t entry
 * points do; eig_solve_block_u2t is the eigen variant's.  No counterpart in
 * the reference (its X_R is FP64, io.c:35-55).  Asynchronous.
 */
size_t fgls_hip_x2t_bytes(int n, int count);
int fgls_hip_tile_u2(const int8_t *dX8, int64_t ldx8, int n, int count, uint8_t *dX2T,
                     unsigned int *dbad, void *stream);
int fgls_hip_tile_u2_packed(const uint8_t *dP, int64_t ldp, int n, int count, uint8_t *dX2T,
                            unsigned int *dbad, void *stream);
int fgls_hip_block_reduce_u2t(fgls_hip_ctx *ctx, const uint8_t *dX2T, int count, void *stream);
int fgls_hip_solve_block_u2t(fgls_hip_ctx *ctx, const uint8_t *dX2T, int count, double *dB,
                             double *dV, int64_t snp_base, void *stream);

/* 2-bit genotypes -> int8 (column-major, ld ldx8 >= n, ldx8 % 4 == 0, rows
 * n..ldx8-1 zeroed): column c is ceil(n/4) bytes at dP + c*ldp, individual r
 * in bits 2(r%4)..2(r%4)+1 of byte r/4 (the first individual in the low
 * bits); codes 0, 1, 2 are the genotype, code 3 is counted into *dbad
 * (device counter, caller-cleared) and written as 0.  No counterpart in the
 * reference (its X_R is FP64 only, io.c:35-55).  Asynchronous. */
int fgls_hip_u2_to_i8(const uint8_t *dP, int64_t ldp, int n, int ncols,
                      int8_t *dX8, int64_t ldx8, unsigned int *dbad, void *stream);
/* 16-bit dosage values v = x * scale (column-major int16, ld ldv elements)
 * -> the two int8 digit planes of fgls_hip_solve_block_dosage (hi plane at
 * dX8 + ncols * ldx8; v outside [-16256, 16383] counted into *dbad), or ->
 * FP64 x = v / scale (ld ldx).  For compact dosage files; no counterpart in
 * the reference (its X_R is FP64 only, io.c:35-55).  Asynchronous. */
int fgls_hip_i16_to_dosage_i8(const int16_t *dV, int64_t ldv, int n, int ncols,
                              int8_t *dX8, int64_t ldx8, unsigned int *dbad, void *stream);
int fgls_hip_i16_to_f64(const int16_t *dV, int64_t ldv, int n, int ncols, double scale,
                        double *dX, int64_t ldx, void *stream);
/* Count into *dbad (device counter, caller-cleared) the values of -128 in
 * rows 0..n-1 of ncols int8 genotype columns (ld ldx8): the int8 product
 * takes |x| <= 127 only.  For int8 genotypes that did not come through
 * fgls_hip_to_i8 (e.g. an int8 X_R file).  No counterpart in the
 * reference.  Asynchronous. */
int fgls_hip_check_i8(const int8_t *dX8, int64_t ldx8, int n, int ncols,
                      unsigned int *dbad, void *stream);
/* int8 genotypes (ld ldx8) -> FP64 (ld ldx, rows n..ldx-1 zeroed), for
 * compact genotype input on the FP64 path.  No counterpart in the
 * reference (its X_R is FP64 only, io.c:35-55).  Asynchronous. */
int fgls_hip_i8_to_f64(const int8_t *dX8, int64_t ldx8, int n, int ncols,
                       double *dX, int64_t ldx, void *stream);

/*
 * In-place X <- L^-1 X for ncols columns (ld ldx), with the factor of the
 * current trait: the direct replacement of
 * cublasDtrsm(LEFT, LOWER, N, NON_UNIT, n, ncols, 1, L, n, X, n)
 * at fgls_chol_gpu.c:550 (blocked left-looking TRSM, diagonal blocks via
 * their LDS-computed inverses, FP64 MFMA updates).  Asynchronous.
 */
int fgls_hip_trsm(fgls_hip_ctx *ctx, double *dX, int64_t ldx, int ncols,
                  void *stream);

/*
 * Z <- L^-1 X through the explicit inverse (triangular MFMA product), for
 * ncols columns; dZ may not alias dX.  ldx % 4 == 0, ldz % 4 == 0,
 * ldz >= fgls_hip_ctx_ldn(ctx).  Asynchronous.
 */
int fgls_hip_trmm(fgls_hip_ctx *ctx, const double *dX, int64_t ldx,
                  int ncols, double *dZ, int64_t ldz, void *stream);

/*
 * Per-SNP stage from already transformed columns Z = L^-1 X_R (count SNPs,
 * wXR columns each, ld ldz): the batched replacement of the OpenMP loop at
 * fgls_chol_gpu.c:619-687 (= fgls_chol.c:386-439) for any wXR.  Asynchronous.
 */
int fgls_hip_snp_stage(fgls_hip_ctx *ctx, const double *dZ, int64_t ldz,
                       int count, double *dB, double *dV, int64_t snp_base,
                       void *stream);

/*
 * fgls_hip_solve_block for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, results
 * and failure records (fgls_hip_check) as fgls_hip_solve_block, FP64
 * arithmetic in either mode; it calls no hipMalloc, hipFree or allocating
 * library routine.  Its scratch -- the per-row-tile partial sums of the
 * fused product (wXR = 1) or the transformed columns L^-1 X_R (wXR > 1) --
 * is one file-scope device array of FGLS_NOALLOC_SCRATCH_BYTES that each
 * device holds for the life of the process, so count is limited to
 *   wXR = 1:  count * ceil(n/128) * (p + 1) * 8 <= FGLS_NOALLOC_SCRATCH_BYTES
 *             (21,236 SNPs at n = 10,000, p = 4)
 *   wXR > 1:  count * wXR * ceil(n/128)*128 * 8 <= FGLS_NOALLOC_SCRATCH_BYTES
 * and a larger count is FGLS_ERR_ARG.  Every call on a device shares that
 * device's array: calls that overlap in time on one device (other streams
 * or threads) race.  The context's own partial sums are untouched, so a
 * pending fgls_hip_block_reduce / fgls_hip_block_finalize pair stays valid.
 * Asynchronous on `stream`.
 */
#define FGLS_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int fgls_hip_solve_block_noalloc(fgls_hip_ctx *ctx, const double *dXR, int64_t ldxr,
                                 int count, double *dB, double *dV, int64_t snp_base,
                                 void *stream);

/*
 * Synchronise the context's work on `stream` and report a recorded per-SNP
 * factorisation failure: returns FGLS_ERR_NOTSPD with *bad_snp (0-based,
 * global index) and *info (LAPACK-style pivot index, 1-based), after which
 * the record is cleared; FGLS_OK otherwise.
 */
int fgls_hip_check(fgls_hip_ctx *ctx, void *stream, int64_t *bad_snp,
                   int *info);

/*
 * Per-trait precompute from Phi instead of a host factor: builds
 * M = sigma2 (h2 Phi + (1 - h2) I) and, on the device in FP64, its lower
 * Cholesky factor together with L^-1 (one recursion: 64 x 64 leaves, MFMA
 * products for the off-diagonal blocks) and then does what
 * fgls_hip_set_trait does after its inversion.  Replaces the host M construction and dpotrf of
 * fgls_chol_gpu.c:406-430 and the L upload of fgls_chol_gpu.c:434-445.
 * dPhi: device, column-major, ld >= n (only the lower triangle is read).
 * A non-SPD M gives FGLS_ERR_NOTSPD with "dpotrf(M) failed (info: i)",
 * i the order of the first non-positive leading minor as LAPACK reports it.
 * Synchronizes `stream`.
 */
int fgls_hip_set_trait_phi(fgls_hip_ctx *ctx, const double *dPhi, int64_t ldphi,
                           double h2, double sigma2, const double *dXL,
                           int64_t ldxl, const double *dy, void *stream);

/* Copy of the current trait's factor L (n x n, column-major, ld >= n; the
 * lower triangle is the factor, the strict upper triangle is unspecified). */
int fgls_hip_get_factor(fgls_hip_ctx *ctx, double *dL, int64_t ldl, void *stream);

/*
 * Adjoint of the loss l = sum_i w_i beta_i^2 over the SNPs of the current
 * trait, beta_i = b_i[p-1] (the SNP coefficient of the per-SNP solve,
 * fgls_chol.c:386-439), with respect to y and h^2, where
 * M = sigma2 (h^2 Phi + (1 - h^2) I) is the matrix whose factor L was given
 * to fgls_hip_set_trait (fgls_chol_gpu.c:406-414).  The reference has no
 * counterpart: this is an addition on top of its forward path.
 * wXR = 1.  In the mixed mode (fgls_hip_ctx_set_mode before begin) begin
 * also builds H = G L^-1 and its int8 digit slices (ldn*ldn*14 bytes more),
 * and fgls_hip_adjoint_block_i8 takes int8 genotypes (column-major, as for
 * fgls_hip_solve_block_i8): both per-block products then run on the int8
 * MFMA (L^-1 X_R and H X_R = G L^-1 X_R, exact int32 sums, the slices'
 * 3.7e-15 relative representation error); fgls_hip_adjoint_block works in
 * either mode.
 *
 * begin: Phi (device, column-major, ld ldphi >= n) and sigma2 of the trait;
 *        builds G = L^-1 (Phi - I) L^-T once and clears the accumulators.
 * block: one block of SNPs (device X_R as for fgls_hip_solve_block, count
 *        <= max_block) with their weights dw (device, count doubles); the
 *        block's L^-1 X_R is recomputed here (nothing is kept from a forward
 *        pass).  Per-SNP dpotrf failures are recorded as by solve_block.
 * end:   dgrad_y (device, n doubles) = dl/dy; *hloss = l and
 *        *hgrad_h2 = dl/dh^2 (host); synchronizes `stream`.
 */
int fgls_hip_adjoint_begin(fgls_hip_ctx *ctx, const double *dPhi, int64_t ldphi,
                           double sigma2, void *stream);
int fgls_hip_adjoint_block(fgls_hip_ctx *ctx, const double *dXR, int64_t ldxr,
                           int count, const double *dw, int64_t snp_base,
                           void *stream);
int fgls_hip_adjoint_block_i8(fgls_hip_ctx *ctx, const int8_t *dX8, int64_t ldx8,
                              int count, const double *dw, int64_t snp_base,
                              void *stream);
int fgls_hip_adjoint_end(fgls_hip_ctx *ctx, double *dgrad_y, double *hloss,
                         double *hgrad_h2, void *stream);

/*
 * Eigenbasis variant (src/fgls_eigen.c): Phi = Z diag(W) Z' once, then for
 * trait k M_k = sigma2_k (h2_k Phi + (1 - h2_k) I) = Z diag(a_k W + b_k) Z'
 * (fgls_eigen.c:455-468), so that one dense product x' = Z' x_R per SNP
 * serves every trait and only weighted sums over the rows of x' differ per
 * trait.  b and V are those of the Cholesky path (same layout, same
 * per-SNP stage, fgls_eigen.c:510-563).  wXR = 1.
 *
 * create:     n individuals, wXL covariates (p = wXL + 1), up to max_traits
 *             traits per set_traits, up to max_block SNPs per solve call.
 * set_mode:   FGLS_MODE_FP64 (default) or FGLS_MODE_MIXED_I8 (Z' as six
 *             int8 digit slices, int8 genotypes or dosage digit planes as in
 *             the Cholesky context; n <= 133,000).
 * set_basis:  dZ (n x n eigenvectors, column-major, ld ldz), dW (n
 *             eigenvalues), dXL (n x wXL, ld ldxl): keeps Z' and computes
 *             Z' X_L on the device.  Replaces the out-of-core gemms of XL
 *             (fgls_eigen.c:693-724) and keeps Z resident instead of
 *             writing Z'X to temporary files.  Synchronizes `stream`.
 * set_traits: ntr traits with h2[k], sigma2[k] (host arrays) and dY (n x ntr,
 *             column-major, ld ldy): Z' Y, the per-trait weights, B_t, V_tl
 *             and res_sigma (fgls_eigen.c:455-491 and residual_sigma,
 *             :997-1152).  FGLS_ERR_NOTSPD if some a_k W_l + b_k <= 0 or a
 *             V_tl is not SPD.  Synchronizes `stream`.
 * solve_block(_i8, _dosage): count SNPs (X_R as for fgls_hip_solve_block,
 *             fgls_hip_solve_block_i8, fgls_hip_solve_block_dosage); trait k
 *             writes b to dB + k*ldb (count*p) and V to dV + k*ldv
 *             (count*p*p); ldb, ldv are ignored for one trait.  Replaces the
 *             ooc_loops per-SNP code, fgls_eigen.c:494-563.  Asynchronous;
 *             per-SNP dpotrf failures are recorded per trait.
 * check:      synchronizes `stream`; FGLS_ERR_NOTSPD with the first recorded
 *             failure (*trait, *bad_snp counted from snp_base, *info), then
 *             clears the records.
 */
typedef struct fgls_hip_eig fgls_hip_eig;
/* Device eigendecomposition Phi = Z diag(W) Z' (dPhi column-major, lower
 * triangle read; dZ n x n column-major, ld ldz; dW n eigenvalues, ascending)
 * through rocSOLVER's dsyevd, loaded on first use: the device counterpart of
 * eigenDec's dsyevr (src/fgls_eigen.c:780-806).  FGLS_ERR_STATE when
 * rocSOLVER cannot be loaded.  Synchronizes `stream`. */
int fgls_hip_eigen_decompose(int n, const double *dPhi, int64_t ldphi, double *dZ,
                             int64_t ldz, double *dW, int device, void *stream);
int fgls_hip_eig_create(fgls_hip_eig **out, int n, int wXL, int max_traits,
                        int max_block, int device);
int fgls_hip_eig_destroy(fgls_hip_eig *e);
int fgls_hip_eig_ldn(const fgls_hip_eig *e);
int fgls_hip_eig_set_mode(fgls_hip_eig *e, int mode);
int fgls_hip_eig_set_basis(fgls_hip_eig *e, const double *dZ, int64_t ldz,
                           const double *dW, const double *dXL, int64_t ldxl,
                           void *stream);
int fgls_hip_eig_set_traits(fgls_hip_eig *e, int ntr, const double *h2,
                            const double *sigma2, const double *dY, int64_t ldy,
                            void *stream);
/* B_t (wXL), V_tl (wXL*wXL, column-major, upper 0) and res_sigma of trait k. */
int fgls_hip_eig_get_trait(fgls_hip_eig *e, int k, double *hBt, double *hVtl,
                           double *hres_sigma);
int fgls_hip_eig_solve_block(fgls_hip_eig *e, const double *dXR, int64_t ldxr,
                             int count, double *dB, int64_t ldb, double *dV,
                             int64_t ldv, int64_t snp_base, void *stream);
/* Several columns per SNP (wXR >= 1, FP64 mode): SNP i's columns at
 * dXR + (i*wXR + a)*ldxr, a < wXR; p = wXL + wXR <= 16 values of b and
 * p*p of V per SNP and trait (ldb >= count*p, ldv >= count*p*p for several
 * traits).  One dense product Z' X_R per block into a context buffer, then
 * per trait the per-SNP stage with the trait's weights d_k (the reference's
 * fgls_eigen handles any wXR: src/fgls_eigen.c:494-538).  Asynchronous. */
int fgls_hip_eig_solve_block_wide(fgls_hip_eig *e, const double *dXR, int64_t ldxr,
                                  int count, int wXR, double *dB, int64_t ldb,
                                  double *dV, int64_t ldv, int64_t snp_base,
                                  void *stream);
int fgls_hip_eig_solve_block_i8(fgls_hip_eig *e, const int8_t *dX8, int64_t ldx8,
                                int count, double *dB, int64_t ldb, double *dV,
                                int64_t ldv, int64_t snp_base, void *stream);
int fgls_hip_eig_solve_block_dosage(fgls_hip_eig *e, const int8_t *dX8,
                                    int64_t ldx8, int count, double scale,
                                    double *dB, int64_t ldb, double *dV,
                                    int64_t ldv, int64_t snp_base, void *stream);
/* The same from tiled genotypes (fgls_hip_tile_i8). */
int fgls_hip_eig_solve_block_i8t(fgls_hip_eig *e, const int8_t *dX8T, int count,
                                 int planes, double scale, double *dB, int64_t ldb,
                                 double *dV, int64_t ldv, int64_t snp_base, void *stream);
/* The same from 2-bit tiles (fgls_hip_tile_u2, fgls_hip_tile_u2_packed). */
int fgls_hip_eig_solve_block_u2t(fgls_hip_eig *e, const uint8_t *dX2T, int count, double *dB,
                                 int64_t ldb, double *dV, int64_t ldv, int64_t snp_base,
                                 void *stream);
int fgls_hip_eig_check(fgls_hip_eig *e, void *stream, int *trait,
                       int64_t *bad_snp, int *info);

#ifdef __cplusplus
}
#endif

#endif /* FGLS_HIP_H */
