/* Internal helpers shared by the HIP layer and the C host solver. */
#ifndef FGLS_INTERNAL_H
#define FGLS_INTERNAL_H

#include "fgls_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device address of the context's per-SNP failure record: a 64-bit word,
 * all ones when clear, else (first failing SNP << 8) | info. */
unsigned long long *fgls_ctx_err_ptr(fgls_hip_ctx *ctx);
/* Device address of the eigen context's per-trait failure records
 * (max_traits words, same encoding; SNPs counted from snp_base). */
unsigned long long *fgls_eig_err_ptr(fgls_hip_eig *e);

/* Set the calling thread's message returned by fgls_hip_last_error(); returns code. */
int fgls_set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#ifdef __cplusplus
}
#endif

#endif
