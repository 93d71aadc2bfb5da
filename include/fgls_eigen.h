/*
 * fgls_eigen.h -- the eigenbasis variant of the reference's solver, rebuilt
 * for one MI355X.
 *
 * int fgls_eigen(...) has the reference's signature and file contract
 * (src/fgls_eigen.h:68-74, identical to fgls_chol_gpu): the same input files,
 * the same B.out / V.out layout (see fgls_chol_gpu.h), the same results up to
 * rounding.  It computes Phi = Z diag(W) Z' once (on the device by default;
 * host LAPACK dsyevr as eigenDec, src/fgls_eigen.c:780-806, with eig=host),
 * keeps Z resident in device memory,
 * and runs one dense product Z' x_R per SNP on the device that serves every
 * trait of a pass (the per-trait work is a weighted reduction over the rows
 * of Z' x_R, fused into that product).
 *
 * Differences from the reference, all deliberate:
 *   - no temporary files: the reference writes Z'X_L, Z'X_R and Z'Y to
 *     "<path>.tmp" files (common.c:98-100, fgls_eigen.c:693-784); here Z' x_R
 *     is formed per block on the device and never stored;
 *   - p = wXL + wXR <= 16, and precision=mixed needs wXR = 1: anything else
 *     is an error ("[Error]: ..." + exit), which the reference does not
 *     have.  With wXR > 1 (FP64) Z' X_R of a block is stored on the device
 *     and each trait runs the per-SNP stage on its weighted sums
 *     (fgls_hip_eig_solve_block_wide); with wXR = 1 the sums are fused into
 *     the product;
 *   - failures are reported as "[Error]: <message>" + exit(EXIT_FAILURE)
 *     (common.c:11-17), including M_k not positive definite
 *     (h2 W_l + (1 - h2) <= 0 for some eigenvalue W_l of Phi).
 * y_b is ignored; num_threads sets the host LAPACK threads of the
 * eigendecomposition with eig=host.
 */
#ifndef FGLS_EIGEN_H
#define FGLS_EIGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces fgls_eigen (src/fgls_eigen.h:68-74).  Returns 0; on failure prints
 * "[Error]: ..." to stderr and calls exit(EXIT_FAILURE). */
int fgls_eigen(int n, int p, int m, int t, int wXL, int wXR,
               int x_b, int y_b, int num_threads,
               char *Phi_path, char *h2_path, char *sigma2_path,
               char *XL_path, char *XR_path, char *Y_path,
               char *B_path, char *V_path);

/*
 * Same solver returning a status (FGLS_OK or a negative FGLS_ERR_* code of
 * fgls_hip.h, message from fgls_hip_last_error()).  `options` as for
 * fgls_chol_gpu_ex (block, device, serial, stats, io_threads, precision,
 * dosage_scale, xr_format, traits_per_pass); potrf is not used.  With
 * traits_per_pass (default: all traits that fit) the traits of a pass share
 * one product Z' x_R per block.  eig=host decomposes Phi with host LAPACK
 * dsyevr as the reference; the default eig=gpu uses the device
 * (fgls_hip_eigen_decompose).
 */
int fgls_eigen_gpu_ex(int n, int p, int m, int t, int wXL, int wXR,
                      int x_b, int y_b, int num_threads,
                      const char *Phi_path, const char *h2_path, const char *sigma2_path,
                      const char *XL_path, const char *XR_path, const char *Y_path,
                      const char *B_path, const char *V_path, const char *options);

/*
 * Host eigendecomposition Phi = Z diag(W) Z' (Phi column-major, lower
 * triangle read; Z column-major n x n, ld ldz; W ascending): LAPACK dsyevr
 * ("V", "A", "L"), as eigenDec (src/fgls_eigen.c:780-806).  threads < 1 uses
 * every core.  FGLS_OK, or FGLS_ERR_NOTSPD with "dsyevr failed (info: i)".
 */
int fgls_eigen_decompose(int n, const double *Phi, int64_t ldphi, double *Z, int64_t ldz,
                         double *W, int threads);

#ifdef __cplusplus
}
#endif

#endif /* FGLS_EIGEN_H */
