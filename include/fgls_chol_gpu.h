/*
 * fgls_chol_gpu.h -- the solver entry point of the reference, rebuilt for one
 * MI355X.
 *
 * int fgls_chol_gpu(...) has the reference's signature and file contract
 * (src/fgls_chol_gpu.h:4-10, identical to src/fgls_chol.h:4-10):
 *   inputs  Phi_path (n*n doubles, column-major), h2_path and sigma2_path
 *           (>= t doubles), XL_path (wXL*n, column k at k*n), XR_path
 *           (m*wXR*n, SNP i at i*wXR*n), Y_path (t*n, trait j at j*n);
 *   outputs B_path (t*m*p doubles, (j*m+i)*p + k) and V_path (t*m*p*p,
 *           (j*m+i)*p*p + c*p + r, lower triangle, sqrt on the diagonal,
 *           strict upper exactly 0), truncated first ("wb").
 * All files are raw little-endian doubles (io.c:7-33).
 *
 * Differences from the reference, all deliberate:
 *   - t >= 1 traits are supported (the reference GPU path exits for t != 1,
 *     fgls_chol_gpu.c:169-174; its CPU twin loops over traits, re-reading
 *     X_R per trait, fgls_chol.c:223 and :336-337, which is what is done here);
 *   - missing/short input files, non-SPD matrices and device failures are
 *     reported as "[Error]: <message>" on stderr followed by
 *     exit(EXIT_FAILURE) (the reference's error_msg, common.c:11-17) instead
 *     of crashing; no LXR.out file is created (fgls_chol_gpu.c:338);
 *   - no per-block progress lines on stdout (fgls_chol_gpu.c:501-504);
 *   - p = wXL + wXR <= 16 (the per-SNP p x p stage is unrolled per p);
 *     larger p is an error ("[Error]: ..." + exit), which the reference
 *     does not have.  precision=mixed (fgls_chol_gpu_ex options) further
 *     needs n <= 133,000, and wXR = 1 for dosages.
 * y_b is ignored with the reference's warning (fgls_chol_gpu.c:185-186);
 * num_threads sets the host BLAS threads for the per-trait Cholesky of M.
 */
#ifndef FGLS_CHOL_GPU_H
#define FGLS_CHOL_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces fgls_chol_gpu (src/fgls_chol_gpu.h:4-10).  Returns 0; on failure
 * prints "[Error]: ..." to stderr and calls exit(EXIT_FAILURE). */
int fgls_chol_gpu(int n, int p, int m, int t, int wXL, int wXR,
                  int x_b, int y_b, int num_threads,
                  char *Phi_path, char *h2_path, char *sigma2_path,
                  char *XL_path, char *XR_path, char *Y_path,
                  char *B_path, char *V_path);

/*
 * Same solver returning a status instead of exiting: FGLS_OK (0) or a
 * negative FGLS_ERR_* code from fgls_hip.h, with the message available from
 * fgls_hip_last_error().  `options` (may be NULL) is a comma-separated list
 * of key=value settings; when NULL the FGLS_OPTIONS environment variable is
 * read instead:
 *   block=<SNPs>   SNPs per device launch (default: a multiple of x_b near
 *                  16384, near 65536 for xr_format=u2 with precision=mixed;
 *                  at most m)
 *   device=<id>    HIP device (default 0)
 *   serial=1       synchronise after every stage (debug aid, like the
 *                  reference's FGLS_GPU_SERIAL, fgls_chol_gpu.c:44-46)
 *   stats=1        print stage timings and rates to stderr
 *   roctx=0        no roctx ranges around the pipeline stages (by default
 *                  they are emitted when the ROCm roctx library loads; the
 *                  reference's VampirTrace regions, fgls_chol_gpu.c:57-59)
 *   io_threads=<k> threads reading each X_R block in 32 MB preads (default:
 *                  the host cores available, at most 16)
 *   dosage_scale=<s>  with precision=mixed (or with xr_format=i16): X_R values are dosages with a
 *                  fixed number of decimals (s = 1000 for 3 decimals), run as
 *                  two int8 digits per value (twice the int8 work of
 *                  genotypes; wXR = 1); a value that is not a multiple of 1/s
 *                  is an FGLS_ERR_ARG
 *   xr_format=i8   XR_path holds int8 values (m*wXR*n bytes, the same order
 *                  as the FP64 file): 8x fewer bytes from disk and over
 *                  PCIe; used as is by precision=mixed, converted to FP64 on
 *                  the device otherwise (default xr_format=f64, io.c:35-55)
 *   xr_format=u2   XR_path holds 2-bit genotypes, ceil(n/4) bytes per column
 *                  (m*wXR columns in the FP64 file's order; individual r in
 *                  bits 2(r%4)..2(r%4)+1 of byte r/4, codes 0/1/2 = the
 *                  genotype, code 3 is an FGLS_ERR_ARG): 32x fewer bytes than
 *                  FP64; with precision=mixed tiled on the device straight
 *                  into the product's 2-bit operand tiles
 *                  (fgls_hip_tile_u2_packed), else unpacked to int8
 *                  (fgls_hip_u2_to_i8)
 *   xr_format=i16  with dosage_scale=<s>: XR_path holds int16 values
 *                  v = x * s (2 bytes per value, the FP64 file's order),
 *                  4x fewer bytes than FP64; the mixed path splits them into
 *                  its two digit planes, the FP64 path uses x = v / s
 *   traits_per_pass=<k>  traits solved per read of each X_R block (default:
 *                  all t traits when their contexts fit in 60 % of the free
 *                  device memory; the reference re-reads X_R for every
 *                  trait, fgls_chol.c:223 and :336-337)
 *   potrf=host     build M and its Cholesky factor on the host (OpenBLAS,
 *                  as the reference) and upload the factor per trait; the
 *                  default potrf=gpu uploads Phi once and builds and factors
 *                  M on the device (fgls_hip_set_trait_phi)
 *   precision=mixed  int8 genotypes x int8 digit slices of L^-1 on the int8
 *                  MFMA (every X_R value must be an integer in [-127, 127],
 *                  else FGLS_ERR_ARG; wXR = 1 reduces in the product's
 *                  epilogue, wXR > 1 stores L^-1 X_R for the FP64 per-SNP
 *                  stage); precision=fp64 (default) is the reference's
 *                  arithmetic
 *   slices=6       with precision=mixed: six digits per entry of L^-1; the
 *                  default slices=auto keeps five (with a top digit on the
 *                  diagonal 64-blocks) wherever the error bound stays within
 *                  16 times the six-digit one (fgls_hip_ctx_set_slice_form)
 * No counterpart in the reference (its GPU entry only exits on error).
 */
int fgls_chol_gpu_ex(int n, int p, int m, int t, int wXL, int wXR,
                     int x_b, int y_b, int num_threads,
                     const char *Phi_path, const char *h2_path, const char *sigma2_path,
                     const char *XL_path, const char *XR_path, const char *Y_path,
                     const char *B_path, const char *V_path, const char *options);

/*
 * Timings of the calling thread's last successful fgls_chol_gpu_ex /
 * fgls_eigen_gpu_ex call (the reference reports wall time only, driver.c:52-100;
 * its VampirTrace regions, fgls_chol_gpu.c:57-71, are the counterpart).
 * Seconds from a monotonic clock unless stated.
 */
typedef struct {
  double total_s;       /* the whole call: file reads, setups, block loops, writes */
  double factor_s;      /* host dpotrf of M (potrf=host) or the eigendecomposition */
  double trait_s;       /* per-trait setups (device factor, L^-1, precompute) */
  double loop_s;        /* the block loops (read, upload, solve, write back) */
  double read_busy_s;   /* reader thread time inside pread of X_R */
  double h2d_busy_s;    /* copy-stream time of the X_R uploads (HIP events) */
  double xr_bytes;      /* bytes of X_R read (every pass) */
  double h2d_bytes;     /* bytes of X_R uploaded */
  int64_t trait_snps;   /* m * t */
  int64_t block;        /* SNPs per device launch */
  int passes;           /* reads of X_R (t / traits_per_pass, rounded up) */
  int traits_per_pass;
  int io_threads;
} fgls_run_stats;

/* FGLS_OK and *out filled (all zero before the first successful call). */
int fgls_last_run_stats(fgls_run_stats *out);

/*
 * Host factor of one trait's covariance: L (column-major, ld ldl) receives
 * M = sigma2 (h2 Phi + (1 - h2) I) and is overwritten by its lower Cholesky
 * factor (fgls_chol_gpu.c:406-430: M built in place, then dpotrf "L"; the
 * strict upper triangle keeps sigma2 h2 Phi).  threads < 1 uses every core.
 * FGLS_OK, or FGLS_ERR_NOTSPD with "dpotrf(M) failed (info: i)".
 */
int fgls_factor_M(int n, const double *Phi, int64_t ldphi, double h2, double sigma2,
                  double *L, int64_t ldl, int threads);

#ifdef __cplusplus
}
#endif

#endif /* FGLS_CHOL_GPU_H */
