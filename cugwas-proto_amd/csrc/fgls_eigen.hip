// fgls_eigen.hip -- the eigenbasis variant of the FGLS solve on the device
// (reference src/fgls_eigen.c) and its C-ABI (include/fgls_hip.h).
//
// With Phi = Z diag(W) Z' (one symmetric eigendecomposition, shared by every
// trait) the covariance of trait k is M_k = Z diag(a_k W + b_k) Z' with
// a_k = sigma2_k h2_k, b_k = sigma2_k (1 - h2_k) (fgls_eigen.c:455-468), so
//   X' M_k^-1 X = (Z'X)' D_k (Z'X),   D_k = diag(1 / (a_k W + b_k)).
// One dense product x' = Z' x_R per SNP serves all traits; per trait only
// weighted sums over the n rows of x' remain:
//   sum_l d_kl x'_l^2,  sum_l d_kl x'_l (Z'X_L)_lc,  sum_l d_kl x'_l (Z'y_k)_l
// (the reference scales by sqrt(d_k) and forms the same dot products,
// fgls_eigen.c:494-538).  They are fused into the epilogue of the dense
// product (FP64 MFMA, or the int8 digit-slice product of the mixed mode),
// and the p x p stage is the one of the Cholesky path.
//
// Device data (FP64, ldn = ceil(n/128)*128 rows, zero padding):
//   Zt    ldn x ldn  row-major Z' (row r = eigenvector r)
//   XLp   ldn x wXL  Z' X_L;  Yp ldn x ntr  Z' Y
//   Wg    ldn x ntr (p+1)  per trait: [d_k, d_k Z'X_L, d_k Z'y_k]
//   Wsc   ldn x ntr p      per trait: [sqrt(d_k) Z'X_L, sqrt(d_k) Z'y_k]
//   part  tiles x ntr x (p+1) x max_block  per-row-tile partial sums

#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <rocsolver/rocsolver.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "fgls_ctx.h"
#include "fgls_device.h"
#include "fgls_hip.h"
#include "fgls_internal.h"

struct fgls_hip_eig {
  int device;
  int n, wXL, p;
  int ldn;
  int RT, RT8;       // 128-row (FP64) and 32-row (int8) row tiles
  int max_traits, max_block;
  int ntr;           // traits of the last set_traits
  int mode;
  bool have_basis, have_traits;
  int reduced, last_rt, part_rt;
  double* Zt;
  double* evals;
  double* XLin;      // X_L, zero padded (operand of Z' X_L)
  double* XLp;
  double* Yin;       // Y, zero padded
  double* Yp;
  double* Wg;
  double* Wsc;
  double* hyp;       // a_k, b_k
  double* trait;     // ntr x TRAIT_LD
  double* part;
  int8_t* A8;        // slices of Z' (mixed mode)
  int8_t* X8t;       // tiled genotypes of the current block (mixed mode)
  size_t X8t_cap;
  double* xp;        // Z' X_R of the current block (wXR > 1), ldn rows
  size_t xp_cap;
  double* invsc;
  unsigned long long* err;   // per trait: (first failing SNP << 8) | info, all ones if none
  int* status;       // [0, max_traits): V_tl Cholesky info per trait; [max_traits]: bad weights
  std::vector<double> hyp_host;
};

namespace {

using fgls::fail;
using fgls::TRAIT_LD;

#define HIP_TRY(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(FGLS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Per-trait weights (fgls_eigen.c:455-468: Winv = sqrt(1 / (a W + b))):
// thread per row l, grid row per trait k.  Rows l >= n are written as 0.  A
// non-positive or non-finite a W_l + b (M_k not positive definite) is counted
// in *bad.
__global__ __launch_bounds__(256) void k_eig_weights(int n, int ldn, int wXL,
                                                     const double* __restrict__ evals,
                                                     const double* __restrict__ hyp,
                                                     const double* __restrict__ XLp,
                                                     const double* __restrict__ Yp,
                                                     double* __restrict__ Wg, double* __restrict__ Wsc,
                                                     int* __restrict__ bad) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (l >= ldn) return;
  const int p = wXL + 1;
  double* wg = Wg + (size_t)k * (p + 1) * ldn;
  double* ws = Wsc + (size_t)k * p * ldn;
  double d = 0.0, winv = 0.0;
  if (l < n) {
    const double m = hyp[2 * k] * evals[l] + hyp[2 * k + 1];
    if (m > 0.0 && isfinite(m)) {
      d = 1.0 / m;
      winv = sqrt(d);
    } else {
      atomicAdd(bad, 1);
    }
  }
  wg[l] = d;
  for (int c = 0; c < wXL; ++c) {
    const double x = (l < n) ? XLp[(size_t)c * ldn + l] : 0.0;
    wg[(size_t)(1 + c) * ldn + l] = d * x;
    ws[(size_t)c * ldn + l] = winv * x;
  }
  const double y = (l < n) ? Yp[(size_t)k * ldn + l] : 0.0;
  wg[(size_t)p * ldn + l] = d * y;
  ws[(size_t)wXL * ldn + l] = winv * y;
}

int check_eig(const fgls_hip_eig* e) {
  if (!e) return fail(FGLS_ERR_ARG, "NULL eigen context");
  HIP_TRY(hipSetDevice(e->device));
  return FGLS_OK;
}

bool aligned32(const void* p) { return ((uintptr_t)p & 31u) == 0; }

int check_outputs(const fgls_hip_eig* e, int count, const double* dB, int64_t ldb, const double* dV,
                  int64_t ldv) {
  if (!dB || !dV) return fail(FGLS_ERR_ARG, "NULL pointer");
  const int64_t p = e->p;
  if (e->ntr > 1 && (ldb < (int64_t)count * p || ldv < (int64_t)count * p * p))
    return fail(FGLS_ERR_ARG, "per-trait output strides ldb=%lld ldv=%lld below count*p, count*p*p",
                (long long)ldb, (long long)ldv);
  return FGLS_OK;
}

int ensure_part(fgls_hip_eig* e, int tiles) {
  if (e->part_rt >= tiles) return FGLS_OK;
  const size_t bytes =
      (size_t)tiles * e->max_traits * (e->p + 1) * (size_t)e->max_block * sizeof(double);
  double* np = nullptr;
  HIP_TRY(hipDeviceSynchronize());
  if (hipMalloc((void**)&np, bytes) != hipSuccess)
    return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
  if (e->part) HIP_TRY(hipFree(e->part));
  e->part = np;
  e->part_rt = tiles;
  return FGLS_OK;
}

// The int8 product of a block of tiled genotypes (planes = 1) or dosage
// digits (planes = 2).
int reduce_i8t(fgls_hip_eig* e, const int8_t* dX8T, int count, int planes, double scale,
               hipStream_t st) {
  if (e->mode != FGLS_MODE_MIXED_I8) return fail(FGLS_ERR_STATE, "eigen context is not in the mixed mode");
  if (planes != 1 && planes != 2) return fail(FGLS_ERR_ARG, "planes must be 1 or 2");
  if (!dX8T || ((uintptr_t)dX8T & 15)) return fail(FGLS_ERR_ARG, "tiled genotypes need 16-byte alignment");
  if (planes == 2 && fgls_x8t_plane(e->n, count) > ((int64_t)1 << 31) - 1)
    return fail(FGLS_ERR_ARG, "int8 X_R block of %d columns too large", count);
  if (planes == 2 && !(scale > 0.0)) return fail(FGLS_ERR_ARG, "dosage scale must be positive");
  fgls_launch_trmm_i8(e->p, e->n, count, e->RT8, e->ldn / 64, e->A8, e->invsc, dX8T,
                      round_up(e->n, 64), e->Wg, (int64_t)e->ldn, e->part, (int64_t)e->max_block,
                      planes == 2 ? fgls_x8t_plane(e->n, count) : 0, planes == 2 ? 1.0 / scale : 1.0,
                      e->ntr, FGLS_NSLICE, st);
  HIP_TRY(hipGetLastError());
  e->last_rt = e->RT8;
  return FGLS_OK;
}

// The int8 product of 2-bit tiled genotypes (fgls_x2t_bytes).
int reduce_u2t(fgls_hip_eig* e, const uint8_t* dX2T, int count, hipStream_t st) {
  if (e->mode != FGLS_MODE_MIXED_I8) return fail(FGLS_ERR_STATE, "eigen context is not in the mixed mode");
  if (!dX2T || ((uintptr_t)dX2T & 3)) return fail(FGLS_ERR_ARG, "2-bit tiles need 4-byte alignment");
  fgls_launch_trmm_u2(e->p, e->n, count, e->RT8, e->ldn / 64, e->A8, e->invsc, dX2T, e->Wg, (int64_t)e->ldn,
                      e->part, (int64_t)e->max_block, e->ntr, FGLS_NSLICE, st);
  HIP_TRY(hipGetLastError());
  e->last_rt = e->RT8;
  return FGLS_OK;
}

// Column-major genotypes: tiled into the context's buffer, then the product.
int reduce_i8(fgls_hip_eig* e, const int8_t* dX8, int64_t ldx8, int count, int planes, double scale,
              hipStream_t st) {
  if (e->mode != FGLS_MODE_MIXED_I8) return fail(FGLS_ERR_STATE, "eigen context is not in the mixed mode");
  if (!dX8) return fail(FGLS_ERR_ARG, "NULL pointer");
  if (ldx8 < round_up(e->n, 64) || (ldx8 & 15) || ((uintptr_t)dX8 & 15))
    return fail(FGLS_ERR_ARG, "int8 X_R needs ld >= ceil(n/64)*64, ld %% 16 == 0 and 16-byte alignment");
  int rc = fgls::ensure_buffer(&e->X8t, &e->X8t_cap, fgls_x8t_bytes(e->n, count, planes));
  if (rc) return rc;
  fgls_launch_tile_i8(e->n, count, planes, dX8, ldx8, e->X8t, st);
  HIP_TRY(hipGetLastError());
  return reduce_i8t(e, e->X8t, count, planes, scale, st);
}

int solve_common(fgls_hip_eig* e, int count, double* dB, int64_t ldb, double* dV, int64_t ldv,
                 int64_t snp_base, hipStream_t st) {
  fgls::launch_finalize_multi(e->p, count, e->last_rt, e->part, (int64_t)e->max_block, e->ntr,
                              e->trait, (int64_t)TRAIT_LD, dB, ldb, dV, ldv, e->err, snp_base, st);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int pre_solve(fgls_hip_eig* e, int count) {
  int rc = check_eig(e);
  if (rc) return rc;
  if (!e->have_traits) return fail(FGLS_ERR_STATE, "no traits set");
  if (count < 0 || count > e->max_block)
    return fail(FGLS_ERR_ARG, "count=%d outside [0, %d]", count, e->max_block);
  return FGLS_OK;
}

// rocSOLVER's dsyevd for the device eigendecomposition, loaded on first use
// (the library is large and only this one-off setup needs it).
struct RocSolver {
  decltype(&rocblas_create_handle) create = nullptr;
  decltype(&rocblas_destroy_handle) destroy = nullptr;
  decltype(&rocblas_set_stream) set_stream = nullptr;
  decltype(&rocsolver_dsyevd) dsyevd = nullptr;
  bool ok = false;
};

const RocSolver& rocsolver() {
  static RocSolver r;
  static std::once_flag once;
  std::call_once(once, [] {
    void* h = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("/opt/rocm/lib/librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    r.create = (decltype(r.create))dlsym(h, "rocblas_create_handle");
    r.destroy = (decltype(r.destroy))dlsym(h, "rocblas_destroy_handle");
    r.set_stream = (decltype(r.set_stream))dlsym(h, "rocblas_set_stream");
    r.dsyevd = (decltype(r.dsyevd))dlsym(h, "rocsolver_dsyevd");
    r.ok = r.create && r.destroy && r.set_stream && r.dsyevd;
  });
  return r;
}

}  // namespace

extern "C" {

int fgls_hip_eigen_decompose(int n, const double* dPhi, int64_t ldphi, double* dZ, int64_t ldz,
                             double* dW, int device, void* stream) {
  if (n < 1 || !dPhi || !dZ || !dW || ldphi < n || ldz < n || ldz > INT32_MAX)
    return fail(FGLS_ERR_ARG, "fgls_hip_eigen_decompose: bad arguments");
  const RocSolver& rs = rocsolver();
  if (!rs.ok) return fail(FGLS_ERR_STATE, "rocSOLVER (librocsolver.so.0) is not available");
  HIP_TRY(hipSetDevice(device));
  hipStream_t st = (hipStream_t)stream;
  // dsyevd overwrites its input with the eigenvectors: Z starts as Phi
  HIP_TRY(hipMemcpy2DAsync(dZ, (size_t)ldz * sizeof(double), dPhi, (size_t)ldphi * sizeof(double),
                           (size_t)n * sizeof(double), (size_t)n, hipMemcpyDeviceToDevice, st));
  double* E = nullptr;
  rocblas_int* dinfo = nullptr;
  if (hipMalloc((void**)&E, (size_t)n * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&dinfo, sizeof(rocblas_int)) != hipSuccess) {
    if (E) hipFree(E);
    return fail(FGLS_ERR_NOMEM, "hipMalloc of the dsyevd workspace failed");
  }
  rocblas_handle h = nullptr;
  int rc = FGLS_OK;
  rocblas_int info = 0;
  if (rs.create(&h) != rocblas_status_success || rs.set_stream(h, st) != rocblas_status_success) {
    rc = fail(FGLS_ERR_HIP, "rocblas handle creation failed");
  } else if (rs.dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, dZ, (rocblas_int)ldz, dW, E,
                       dinfo) != rocblas_status_success) {
    rc = fail(FGLS_ERR_HIP, "rocsolver_dsyevd failed to launch");
  } else if (hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess) {
    rc = fail(FGLS_ERR_HIP, "dsyevd: synchronisation failed");
  } else if (info != 0) {
    rc = fail(FGLS_ERR_NOTSPD, "dsyevd failed (info: %d)", (int)info);
  }
  if (h) rs.destroy(h);
  hipFree(E);
  hipFree(dinfo);
  return rc;
}

int fgls_hip_eig_create(fgls_hip_eig** out, int n, int wXL, int max_traits, int max_block,
                        int device) {
  if (!out) return fail(FGLS_ERR_ARG, "NULL output pointer");
  *out = nullptr;
  if (n < 1 || wXL < 0 || wXL + 1 > fgls::PMAX || max_traits < 1 || max_block < 1)
    return fail(FGLS_ERR_ARG,
                "invalid sizes n=%d wXL=%d max_traits=%d max_block=%d (need n>=1, p<=%d)", n, wXL,
                max_traits, max_block, fgls::PMAX);
  if (wXL >= n) return fail(FGLS_ERR_ARG, "wXL=%d must be smaller than n=%d", wXL, n);
  const int ndev = fgls_hip_device_count();
  if (device < 0 || device >= ndev)
    return fail(FGLS_ERR_ARG, "device %d not available (%d visible)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  fgls_hip_eig* e = new fgls_hip_eig();
  e->device = device;
  e->n = n;
  e->wXL = wXL;
  e->p = wXL + 1;
  e->ldn = (int)round_up(n, fgls::TM);
  e->RT = e->ldn / fgls::TM;
  e->RT8 = e->ldn / I8_TM;
  e->max_traits = max_traits;
  e->max_block = (int)round_up(max_block, 4);
  e->mode = FGLS_MODE_FP64;
  e->hyp_host.assign(2 * (size_t)max_traits, 0.0);
  const size_t ldn = (size_t)e->ldn, T = (size_t)max_traits, p = (size_t)e->p;
  const size_t wl = wXL > 0 ? (size_t)wXL : 1;
  struct { void** ptr; size_t bytes; } allocs[] = {
      {(void**)&e->Zt, ldn * ldn * sizeof(double)},
      {(void**)&e->evals, ldn * sizeof(double)},
      {(void**)&e->XLin, ldn * wl * sizeof(double)},
      {(void**)&e->XLp, ldn * wl * sizeof(double)},
      {(void**)&e->Yin, ldn * T * sizeof(double)},
      {(void**)&e->Yp, ldn * T * sizeof(double)},
      {(void**)&e->Wg, ldn * T * (p + 1) * sizeof(double)},
      {(void**)&e->Wsc, ldn * T * p * sizeof(double)},
      {(void**)&e->hyp, 2 * T * sizeof(double)},
      {(void**)&e->trait, T * TRAIT_LD * sizeof(double)},
      {(void**)&e->err, T * sizeof(unsigned long long)},
      {(void**)&e->status, (T + 4) * sizeof(int)},
  };
  for (auto& a : allocs) {
    if (hipMalloc(a.ptr, a.bytes) != hipSuccess) {
      fgls_hip_eig_destroy(e);
      return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", a.bytes);
    }
    if (hipMemset(*a.ptr, 0, a.bytes) != hipSuccess) {
      fgls_hip_eig_destroy(e);
      return fail(FGLS_ERR_HIP, "hipMemset failed");
    }
  }
  int rc = ensure_part(e, e->RT);
  if (rc) {
    fgls_hip_eig_destroy(e);
    return rc;
  }
  HIP_TRY(hipMemset(e->err, 0xff, T * sizeof(unsigned long long)));
  HIP_TRY(hipDeviceSynchronize());
  *out = e;
  return FGLS_OK;
}

int fgls_hip_eig_destroy(fgls_hip_eig* e) {
  if (!e) return FGLS_OK;
  hipSetDevice(e->device);
  hipDeviceSynchronize();
  void* ptrs[] = {e->Zt, e->evals, e->XLin, e->XLp, e->Yin, e->Yp, e->Wg, e->Wsc, e->hyp,
                  e->trait, e->part, e->A8, e->invsc, e->err, e->status, e->X8t, e->xp};
  for (void* q : ptrs)
    if (q) hipFree(q);
  delete e;
  return FGLS_OK;
}

int fgls_hip_eig_set_mode(fgls_hip_eig* e, int mode) {
  int rc = check_eig(e);
  if (rc) return rc;
  if (mode != FGLS_MODE_FP64 && mode != FGLS_MODE_MIXED_I8) return fail(FGLS_ERR_ARG, "unknown mode %d", mode);
  if (mode == FGLS_MODE_MIXED_I8 && e->n > FGLS_I8_MAX_N)
    return fail(FGLS_ERR_ARG, "the mixed mode is exact for n <= %d (got %d)", FGLS_I8_MAX_N, e->n);
  if (mode == FGLS_MODE_MIXED_I8) {
    if (!e->A8) {
      const size_t bytes = (size_t)e->ldn * e->ldn * FGLS_NSLICE;
      if (hipMalloc((void**)&e->A8, bytes) != hipSuccess)
        return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
      if (hipMalloc((void**)&e->invsc, (size_t)e->ldn * sizeof(double)) != hipSuccess)
        return fail(FGLS_ERR_NOMEM, "hipMalloc of the row scales failed");
    }
    if ((rc = ensure_part(e, e->RT8))) return rc;
    if ((rc = fgls::ensure_buffer(&e->X8t, &e->X8t_cap, fgls_x8t_bytes(e->n, e->max_block, 1))))
      return rc;
    if (e->have_basis) {
      fgls::launch_slices(e->ldn, e->Zt, e->A8, e->invsc, nullptr);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipDeviceSynchronize());
    }
  }
  e->mode = mode;
  return FGLS_OK;
}

int fgls_hip_eig_set_basis(fgls_hip_eig* e, const double* dZ, int64_t ldz, const double* dW,
                           const double* dXL, int64_t ldxl, void* stream) {
  int rc = check_eig(e);
  if (rc) return rc;
  if (!dZ || !dW || (e->wXL > 0 && !dXL)) return fail(FGLS_ERR_ARG, "NULL input pointer");
  if (ldz < e->n || (e->wXL > 0 && ldxl < e->n)) return fail(FGLS_ERR_ARG, "leading dimension below n");
  hipStream_t st = (hipStream_t)stream;
  const int n = e->n;
  const size_t ldn = (size_t)e->ldn;
  e->have_basis = e->have_traits = false;
  // column r of Z (column-major) is row r of Z' (row-major)
  HIP_TRY(hipMemcpy2DAsync(e->Zt, ldn * sizeof(double), dZ, (size_t)ldz * sizeof(double),
                           (size_t)n * sizeof(double), (size_t)n, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(e->evals, dW, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (e->wXL > 0) {
    HIP_TRY(hipMemcpy2DAsync(e->XLin, ldn * sizeof(double), dXL, (size_t)ldxl * sizeof(double),
                             (size_t)n * sizeof(double), (size_t)e->wXL, hipMemcpyDeviceToDevice, st));
    // Z' X_L (fgls_eigen.c:693-724, the out-of-core gemm of XL)
    fgls::launch_dense_store(n, e->ldn, e->Zt, e->XLin, (int64_t)ldn, e->wXL, e->XLp, (int64_t)ldn, st);
  }
  if (e->mode == FGLS_MODE_MIXED_I8) fgls::launch_slices(e->ldn, e->Zt, e->A8, e->invsc, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  e->have_basis = true;
  return FGLS_OK;
}

int fgls_hip_eig_set_traits(fgls_hip_eig* e, int ntr, const double* h2, const double* sigma2,
                            const double* dY, int64_t ldy, void* stream) {
  int rc = check_eig(e);
  if (rc) return rc;
  if (!e->have_basis) return fail(FGLS_ERR_STATE, "no eigenbasis set");
  if (ntr < 1 || ntr > e->max_traits)
    return fail(FGLS_ERR_ARG, "ntr=%d outside [1, %d]", ntr, e->max_traits);
  if (!h2 || !sigma2 || !dY) return fail(FGLS_ERR_ARG, "NULL input pointer");
  if (ldy < e->n) return fail(FGLS_ERR_ARG, "leading dimension of Y below n");
  hipStream_t st = (hipStream_t)stream;
  const int n = e->n, wXL = e->wXL;
  const size_t ldn = (size_t)e->ldn;
  e->have_traits = false;
  for (int k = 0; k < ntr; ++k) {   // fgls_eigen.c:455-459
    e->hyp_host[2 * k] = sigma2[k] * h2[k];
    e->hyp_host[2 * k + 1] = sigma2[k] * (1.0 - h2[k]);
  }
  HIP_TRY(hipMemcpyAsync(e->hyp, e->hyp_host.data(), 2 * (size_t)ntr * sizeof(double),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpy2DAsync(e->Yin, ldn * sizeof(double), dY, (size_t)ldy * sizeof(double),
                           (size_t)n * sizeof(double), (size_t)ntr, hipMemcpyDeviceToDevice, st));
  // Z' Y (fgls_eigen.c:759-784, the out-of-core gemm of Y)
  fgls::launch_dense_store(n, e->ldn, e->Zt, e->Yin, (int64_t)ldn, ntr, e->Yp, (int64_t)ldn, st);
  HIP_TRY(hipMemsetAsync(e->status, 0, (size_t)(e->max_traits + 4) * sizeof(int), st));
  hipLaunchKernelGGL(k_eig_weights, dim3((unsigned)(ldn / 256 + (ldn % 256 != 0)), ntr), dim3(256), 0,
                     st, n, e->ldn, wXL, e->evals, e->hyp, e->XLp, e->Yp, e->Wg, e->Wsc,
                     e->status + e->max_traits);
  // B_t, V_tl, b_L and res_sigma per trait (fgls_eigen.c:470-491 and the
  // residual variance of residual_sigma, :1119-1152)
  for (int k = 0; k < ntr; ++k)
    fgls::launch_trait_stats(n, wXL, e->Wsc + (size_t)k * e->p * ldn, (int64_t)ldn,
                             e->trait + (size_t)k * TRAIT_LD, e->status + k, st);
  HIP_TRY(hipGetLastError());
  std::vector<int> status((size_t)e->max_traits + 1, 0);
  HIP_TRY(hipMemcpyAsync(status.data(), e->status, status.size() * sizeof(int),
                         hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (status[e->max_traits] != 0)
    return fail(FGLS_ERR_NOTSPD, "M not positive definite: %d eigenvalues give h2 W + (1 - h2) <= 0",
                status[e->max_traits]);
  for (int k = 0; k < ntr; ++k)
    if (status[k] != 0)
      return fail(FGLS_ERR_NOTSPD, "sigma2.score: dpotrf(scoreV) failed (info: %d) - j: %d", status[k], k);
  e->ntr = ntr;
  e->have_traits = true;
  return FGLS_OK;
}

int fgls_hip_eig_get_trait(fgls_hip_eig* e, int k, double* hBt, double* hVtl, double* hres) {
  int rc = check_eig(e);
  if (rc) return rc;
  if (!e->have_traits) return fail(FGLS_ERR_STATE, "no traits set");
  if (k < 0 || k >= e->ntr) return fail(FGLS_ERR_ARG, "trait %d outside [0, %d)", k, e->ntr);
  const int wXL = e->wXL;
  double buf[TRAIT_LD];
  HIP_TRY(hipMemcpy(buf, e->trait + (size_t)k * TRAIT_LD, sizeof(buf), hipMemcpyDeviceToHost));
  if (hBt) memcpy(hBt, buf, (size_t)wXL * sizeof(double));
  if (hVtl) memcpy(hVtl, buf + wXL, (size_t)wXL * wXL * sizeof(double));
  if (hres) *hres = buf[fgls::trait_res_off(wXL)];
  return FGLS_OK;
}

int fgls_hip_eig_solve_block(fgls_hip_eig* e, const double* dXR, int64_t ldxr, int count,
                             double* dB, int64_t ldb, double* dV, int64_t ldv, int64_t snp_base,
                             void* stream) {
  int rc = pre_solve(e, count);
  if (rc || count == 0) return rc;
  if (e->mode != FGLS_MODE_FP64) return fail(FGLS_ERR_STATE, "eigen context is in the mixed mode");
  if ((rc = check_outputs(e, count, dB, ldb, dV, ldv))) return rc;
  if (!dXR) return fail(FGLS_ERR_ARG, "NULL pointer");
  if (ldxr < e->n || (ldxr & 3) || !aligned32(dXR))
    return fail(FGLS_ERR_ARG, "X_R needs ld >= n, ld %% 4 == 0 and 32-byte alignment");
  hipStream_t st = (hipStream_t)stream;
  fgls::launch_eig_reduce(e->p, e->n, e->ldn, e->Zt, dXR, ldxr, count, e->Wg, (int64_t)e->ldn,
                          e->ntr, e->part, (int64_t)e->max_block, st);
  HIP_TRY(hipGetLastError());
  e->last_rt = e->RT;
  return solve_common(e, count, dB, ldb, dV, ldv, snp_base, st);
}

int fgls_hip_eig_solve_block_wide(fgls_hip_eig* e, const double* dXR, int64_t ldxr, int count, int wXR,
                                  double* dB, int64_t ldb, double* dV, int64_t ldv, int64_t snp_base,
                                  void* stream) {
  int rc = pre_solve(e, count);
  if (rc || count == 0) return rc;
  if (e->mode != FGLS_MODE_FP64)
    return fail(FGLS_ERR_STATE, "the eigen variant with several columns per SNP runs in FP64");
  if (wXR < 1 || e->wXL + wXR > fgls::PMAX)
    return fail(FGLS_ERR_ARG, "wXR=%d: need wXR >= 1 and p = wXL + wXR <= %d", wXR, fgls::PMAX);
  if (!dXR || !dB || !dV) return fail(FGLS_ERR_ARG, "NULL pointer");
  const int64_t p = e->wXL + wXR;
  if (e->ntr > 1 && (ldb < (int64_t)count * p || ldv < (int64_t)count * p * p))
    return fail(FGLS_ERR_ARG, "per-trait output strides ldb=%lld ldv=%lld below count*p, count*p*p",
                (long long)ldb, (long long)ldv);
  if (ldxr < e->n || (ldxr & 3) || !aligned32(dXR))
    return fail(FGLS_ERR_ARG, "X_R needs ld >= n, ld %% 4 == 0 and 32-byte alignment");
  hipStream_t st = (hipStream_t)stream;
  const size_t ldn = (size_t)e->ldn;
  const size_t need = ldn * (size_t)count * wXR * sizeof(double);
  if (e->xp_cap < need) {
    HIP_TRY(hipStreamSynchronize(st));
    if (e->xp) HIP_TRY(hipFree(e->xp));
    e->xp = nullptr;
    e->xp_cap = 0;
    if (hipMalloc((void**)&e->xp, need) != hipSuccess)
      return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", need);
    e->xp_cap = need;
  }
  // x' = Z' x for every column of the block (fgls_eigen.c:494-512), then per
  // trait the per-SNP stage on the d_k-weighted sums (fgls_eigen.c:513-538)
  fgls::launch_dense_store(e->n, e->ldn, e->Zt, dXR, ldxr, count * wXR, e->xp, (int64_t)ldn, st);
  for (int k = 0; k < e->ntr; ++k) {
    const double* wg = e->Wg + (size_t)k * (e->p + 1) * ldn;   // [d_k, d_k Z'X_L, d_k Z'y_k]
    fgls::launch_snp_stage_weighted_any((int)p, e->n, count, e->wXL, wXR, e->xp, (int64_t)ldn, wg + ldn,
                                        (int64_t)ldn, wg, e->trait + (size_t)k * TRAIT_LD,
                                        dB + (size_t)k * ldb, dV + (size_t)k * ldv, e->err + k, snp_base, st);
  }
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_eig_solve_block_i8(fgls_hip_eig* e, const int8_t* dX8, int64_t ldx8, int count,
                                double* dB, int64_t ldb, double* dV, int64_t ldv, int64_t snp_base,
                                void* stream) {
  int rc = pre_solve(e, count);
  if (rc || count == 0) return rc;
  if ((rc = check_outputs(e, count, dB, ldb, dV, ldv))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = reduce_i8(e, dX8, ldx8, count, 1, 1.0, st))) return rc;
  return solve_common(e, count, dB, ldb, dV, ldv, snp_base, st);
}

int fgls_hip_eig_solve_block_dosage(fgls_hip_eig* e, const int8_t* dX8, int64_t ldx8, int count,
                                    double scale, double* dB, int64_t ldb, double* dV, int64_t ldv,
                                    int64_t snp_base, void* stream) {
  int rc = pre_solve(e, count);
  if (rc || count == 0) return rc;
  if ((rc = check_outputs(e, count, dB, ldb, dV, ldv))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = reduce_i8(e, dX8, ldx8, count, 2, scale, st))) return rc;
  return solve_common(e, count, dB, ldb, dV, ldv, snp_base, st);
}

int fgls_hip_eig_solve_block_i8t(fgls_hip_eig* e, const int8_t* dX8T, int count, int planes,
                                 double scale, double* dB, int64_t ldb, double* dV, int64_t ldv,
                                 int64_t snp_base, void* stream) {
  int rc = pre_solve(e, count);
  if (rc || count == 0) return rc;
  if ((rc = check_outputs(e, count, dB, ldb, dV, ldv))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = reduce_i8t(e, dX8T, count, planes, scale, st))) return rc;
  return solve_common(e, count, dB, ldb, dV, ldv, snp_base, st);
}

int fgls_hip_eig_solve_block_u2t(fgls_hip_eig* e, const uint8_t* dX2T, int count, double* dB, int64_t ldb,
                                 double* dV, int64_t ldv, int64_t snp_base, void* stream) {
  int rc = pre_solve(e, count);
  if (rc || count == 0) return rc;
  if ((rc = check_outputs(e, count, dB, ldb, dV, ldv))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = reduce_u2t(e, dX2T, count, st))) return rc;
  return solve_common(e, count, dB, ldb, dV, ldv, snp_base, st);
}

int fgls_hip_eig_check(fgls_hip_eig* e, void* stream, int* trait, int64_t* bad_snp, int* info) {
  int rc = check_eig(e);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  std::vector<unsigned long long> keys((size_t)e->max_traits, ~0ull);
  HIP_TRY(hipMemcpyAsync(keys.data(), e->err, keys.size() * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int k = 0; k < e->max_traits; ++k) {
    if (keys[k] == ~0ull) continue;
    const int64_t snp = (int64_t)(keys[k] >> 8);
    const int inf = (int)(keys[k] & 0xff);
    if (trait) *trait = k;
    if (bad_snp) *bad_snp = snp;
    if (info) *info = inf;
    HIP_TRY(hipMemsetAsync(e->err, 0xff, keys.size() * sizeof(unsigned long long), st));
    HIP_TRY(hipStreamSynchronize(st));
    return fail(FGLS_ERR_NOTSPD, "dpotrf failed (info: %d) - trait %d, i: %lld", inf, k, (long long)snp);
  }
  return FGLS_OK;
}

int fgls_hip_eig_ldn(const fgls_hip_eig* e) { return e ? e->ldn : 0; }

unsigned long long* fgls_eig_err_ptr(fgls_hip_eig* e) { return e ? e->err : nullptr; }

}  // extern "C"
