// fgls_potrf.hip -- the per-trait Cholesky factor on the device:
//   M = sigma2 (h2 Phi + (1 - h2) I) = L L'      (FP64, lower)
// replacing the host M construction and dpotrf of fgls_chol_gpu.c:406-430.
//
// Two-level right-looking factorisation.  Outer panels of OB = 256 columns;
// inside a panel, 64-column steps:
//   k_chol_diag   factors the 64 x 64 diagonal block in LDS and forms the
//                 inverse of its factor (one workgroup);
//   k_chol_panel  L21 = A21 L11^-T for the rows below (64-row workgroups);
//                 the rows are also written row-major to a panel buffer P
//                 (ld 256) so that updates read 4 consecutive k per lane;
//   k_chol_update A -= P P' on lower tiles: first the rest of the outer
//                 panel (K = 64), then, once the outer panel is done, the
//                 whole trailing matrix (K = 256: 4x fewer passes over it
//                 than 64-column steps would make).
// The first non-positive pivot is recorded as LAPACK's info (1-based order
// of the failing leading minor); later panels still run on harmless values.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fgls_ctx.h"
#include "fgls_hip.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int PB = 64;   // inner panel width
constexpr int OB = 256;  // outer panel width (ld of P)
constexpr int UT = 128;  // update tile (2 x 2 waves of 64 x 64)

using fgls::fail;

#define HIP_TRY(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(FGLS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// M (ldn x ldn, column-major) = alpha Phi + beta I on the n x n block, zero
// elsewhere (fgls_chol_gpu.c:406-414: alpha = sigma2 h2, beta = sigma2 (1-h2)).
__global__ __launch_bounds__(256) void k_build_M(int n, int ldn, const double* __restrict__ Phi,
                                                 int64_t ldphi, double alpha, double beta,
                                                 double* __restrict__ M) {
  const int c = blockIdx.y;
  for (int r = blockIdx.x * 256 + threadIdx.x; r < ldn; r += gridDim.x * 256) {
    double v = 0.0;
    if (r < n && c < n) v = alpha * Phi[(size_t)c * ldphi + r] + (r == c ? beta : 0.0);
    M[(size_t)c * ldn + r] = v;
  }
}

// Diagonal block at column j (nb = min(64, n - j) columns; the block is
// completed with the identity to 64): unblocked Cholesky in LDS, then the
// inverse of the factor by column-parallel forward substitution.  Writes the
// factor back (strict upper part of the block zeroed) and the inverse,
// column-major 64 x 64, to Dinv_j.
__global__ __launch_bounds__(256) void k_chol_diag(int n, int j, double* __restrict__ L, int64_t ldl,
                                                   double* __restrict__ Dinv_j, int* __restrict__ info) {
  __shared__ double A[PB][PB + 1];
  __shared__ double Di[PB][PB + 1];
  const int tid = threadIdx.x;
  const int nb = min(PB, n - j);
  for (int idx = tid; idx < PB * PB; idx += 256) {
    const int r = idx % PB, c = idx / PB;
    double v;
    if (r < nb && c < nb) v = (r >= c) ? L[(size_t)(j + c) * ldl + j + r] : 0.0;
    else v = (r == c) ? 1.0 : 0.0;
    A[r][c] = v;
  }
  __syncthreads();
  for (int k = 0; k < PB; ++k) {
    if (tid == 0) {
      double d = A[k][k];
      if (!(d > 0.0)) {
        if (*info == 0) *info = j + k + 1;
        d = 1.0;
      }
      A[k][k] = sqrt(d);
    }
    __syncthreads();
    const double dk = A[k][k];
    for (int i = k + 1 + tid; i < PB; i += 256) A[i][k] /= dk;
    __syncthreads();
    const int m = PB - 1 - k;
    for (int p = tid; p < m * m; p += 256) {
      const int i = k + 1 + p / m, l = k + 1 + p % m;
      if (l <= i) A[i][l] -= A[i][k] * A[l][k];
    }
    __syncthreads();
  }
  // thread t < 64 owns column t of the inverse: A * d = e_t
  if (tid < PB) {
    const int t = tid;
    for (int row = 0; row < PB; ++row) {
      double d = 0.0;
      if (row >= t) {
        double s = (row == t) ? 1.0 : 0.0;
        for (int q = t; q < row; ++q) s -= A[row][q] * Di[q][t];
        d = s / A[row][row];
      }
      Di[row][t] = d;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < PB * PB; idx += 256) {
    const int r = idx % PB, c = idx / PB;
    if (r < nb && c < nb) L[(size_t)(j + c) * ldl + j + r] = (r >= c) ? A[r][c] : 0.0;
    Dinv_j[(size_t)c * PB + r] = Di[r][c];
  }
}

// L21 = A21 L11^-T for rows [i0, i0 + 64) below the diagonal block at j
// (j + 64 < n, so the panel is full); also P[i * OB + poff + c] = L21(i, c).
__global__ __launch_bounds__(256) void k_chol_panel(int n, int j, double* __restrict__ L, int64_t ldl,
                                                    const double* __restrict__ Dinv_j,
                                                    double* __restrict__ P, int poff) {
  __shared__ double A[PB][PB + 1];
  __shared__ double Di[PB][PB + 1];
  const int tid = threadIdx.x;
  const int i0 = j + PB + blockIdx.x * PB;
  const int rows = min(PB, n - i0);
  for (int idx = tid; idx < PB * PB; idx += 256) {
    const int r = idx % PB, c = idx / PB;
    A[r][c] = (r < rows) ? L[(size_t)(j + c) * ldl + i0 + r] : 0.0;
    Di[r][c] = Dinv_j[(size_t)c * PB + r];
  }
  __syncthreads();
  // thread: row r = tid % 64, columns [16 (tid / 64), +16)
  const int r = tid % PB, c0 = 16 * (tid / PB);
  double x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = 0.0;
  for (int k = 0; k < PB; ++k) {
    const double a = A[r][k];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = fma(a, Di[c0 + q][k], x[q]);  // (L11^-T)(k, c) = L11^-1(c, k)
  }
  if (r < rows) {
#pragma unroll
    for (int q = 0; q < 16; ++q) L[(size_t)(j + c0 + q) * ldl + i0 + r] = x[q];
    v4d* prow = reinterpret_cast<v4d*>(P + (size_t)(i0 + r) * OB + poff + c0);
#pragma unroll
    for (int q = 0; q < 4; ++q) prow[q] = v4d{x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]};
  }
}

// A(row, col) -= sum_{k<K} P(row, k) P(col, k) for col in [c0, c1), row in
// [c0, n), row >= col; P row-major with ld OB (the columns of the current
// panel).  Grid: (row tiles, column tiles) of 128 from c0; tiles entirely
// above the diagonal exit.  MFMA operands as in the FP64 product: lane group
// g supplies k = 4g + s, one 32-byte load per lane and operand; swapped
// operands put 16 lanes on 16 consecutive rows of one column for the
// read-modify-write.
template <int K>
__global__ __launch_bounds__(256) void k_chol_update(int n, int c0g, int c1g, double* __restrict__ L,
                                                     int64_t ldl, const double* __restrict__ P) {
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (tj > ti) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w >> 1, wc = w & 1;
  const int li = lane & 15, lg = lane >> 4;
  const int r0 = c0g + ti * UT + wr * 64;
  const int c0 = c0g + tj * UT + wc * 64;
  if (r0 >= n || c0 >= c1g || c0 > r0 + 63) return;   // nothing of the region here
  const double* Ap[4];
  const double* Bp[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    Ap[q] = P + (size_t)(r0 + 16 * q + li) * OB + 4 * lg;
    Bp[q] = P + (size_t)(c0 + 16 * q + li) * OB + 4 * lg;
  }
  v4d acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4d av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      av[q] = *reinterpret_cast<const v4d*>(Ap[q] + k0);
      bv[q] = *reinterpret_cast<const v4d*>(Bp[q] + k0);
    }
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = mfma_f64(bv[c][s2], av[a][s2], acc[a][c]);
  }
  // element v of acc[a][c] is (row r0 + 16a + li, column c0 + 16c + lg + 4v)
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = c0 + 16 * c + lg + 4 * v;
      if (col >= c1g) continue;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int row = r0 + 16 * a + li;
        if (row < n && row >= col) L[(size_t)col * ldl + row] -= acc[a][c][v];
      }
    }
  }
}

template <int K>
void launch_update(int n, int c0, int c1, double* L, int64_t ldl, const double* P, hipStream_t st) {
  if (c0 >= n || c1 <= c0) return;
  const int tr = (n - c0 + UT - 1) / UT, tc = (c1 - c0 + UT - 1) / UT;
  hipLaunchKernelGGL(k_chol_update<K>, dim3(tr, tc), dim3(256), 0, st, n, c0, c1, L, ldl, P);
}

int factor_on_device(fgls_hip_ctx* c, const double* dPhi, int64_t ldphi, double h2, double sigma2,
                     hipStream_t st) {
  const int n = c->n, ldn = c->ldn;
  if (!c->panel) {
    const size_t bytes = (size_t)(ldn + UT) * OB * sizeof(double);
    if (hipMalloc((void**)&c->panel, bytes) != hipSuccess)
      return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
    HIP_TRY(hipMemset(c->panel, 0, bytes));
  }
  HIP_TRY(hipMemsetAsync(c->status + 1, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_build_M, dim3((ldn + 255) / 256, ldn), dim3(256), 0, st, n, ldn, dPhi, ldphi,
                     sigma2 * h2, sigma2 * (1.0 - h2), c->L);
  for (int j = 0; j < n; j += OB) {
    const int jend = min(j + OB, n);
    for (int jj = j; jj < jend; jj += PB) {
      double* Dj = c->Dinv + (size_t)(jj / PB) * PB * PB;
      hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(256), 0, st, n, jj, c->L, (int64_t)ldn, Dj,
                         c->status + 1);
      const int below = n - jj - PB;
      if (below <= 0) break;
      hipLaunchKernelGGL(k_chol_panel, dim3((below + PB - 1) / PB), dim3(256), 0, st, n, jj, c->L,
                         (int64_t)ldn, (const double*)Dj, c->panel, jj - j);
      // the rest of this outer panel, from this step's 64 columns
      launch_update<PB>(n, jj + PB, jend, c->L, (int64_t)ldn, c->panel + (jj - j), st);
    }
    // the trailing matrix, from the whole outer panel
    if (jend - j == OB) launch_update<OB>(n, jend, n, c->L, (int64_t)ldn, c->panel, st);
  }
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

}  // namespace

extern "C" {

int fgls_hip_set_trait_phi(fgls_hip_ctx* c, const double* dPhi, int64_t ldphi, double h2,
                           double sigma2, const double* dXL, int64_t ldxl, const double* dy,
                           void* stream) {
  if (!c) return fail(FGLS_ERR_ARG, "NULL context");
  if (!dPhi || !dy || (c->wXL > 0 && !dXL)) return fail(FGLS_ERR_ARG, "NULL input pointer");
  if (ldphi < c->n || (c->wXL > 0 && ldxl < c->n)) return fail(FGLS_ERR_ARG, "leading dimension below n");
  int rc = fgls::set_device(c);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  c->have_trait = false;
  if ((rc = factor_on_device(c, dPhi, ldphi, h2, sigma2, st))) return rc;
  static const int kPotrf = 1;
  return fgls::finish_trait(c, dXL, ldxl, dy, st, &kPotrf);
}

int fgls_hip_get_factor(fgls_hip_ctx* c, double* dL, int64_t ldl, void* stream) {
  if (!c) return fail(FGLS_ERR_ARG, "NULL context");
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (!dL || ldl < c->n) return fail(FGLS_ERR_ARG, "L needs a device pointer and ld >= n");
  int rc = fgls::set_device(c);
  if (rc) return rc;
  HIP_TRY(hipMemcpy2DAsync(dL, (size_t)ldl * sizeof(double), c->L, (size_t)c->ldn * sizeof(double),
                           (size_t)c->n * sizeof(double), (size_t)c->n, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  return FGLS_OK;
}

}  // extern "C"
