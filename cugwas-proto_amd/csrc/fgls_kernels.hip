// fgls_kernels.hip -- gfx950 (MI355X, CDNA4) kernels of the Cholesky FGLS
// hot path and the C-ABI declared in include/fgls_hip.h.
//
// Per SNP i the solver needs, with L L' = M (reference fgls_chol.c:251-368):
//   z_i  = L^-1 x_i                     (the block TRSM, fgls_chol_gpu.c:550)
//   S_i  = [[V_tl, .], [X~_L' z_i, z_i' z_i]],  r_i = [B_t ; z_i' y~]
//   b_i  = S_i^-1 r_i,  V_i = res_sigma * S_i^-1 (sqrt on the diagonal)
//                                        (per-SNP stage, fgls_chol.c:386-439)
//
// Device data layout (all FP64, padded to ldn = ceil(n/128)*128 rows, zero
// padding):
//   L     ldn x ldn  column-major copy of the factor (lower triangle used)
//   Linv  ldn x ldn  ROW-major L^-1, so that one MFMA A-fragment row of 16
//                    consecutive k is a contiguous 128-byte line
//   Dinv  nblk x 64 x 64  inverses of the 64x64 diagonal blocks of L
//   W     ldn x (wXL+1)   [L^-1 X_L | L^-1 y]
//   part  RT x (p+1) x max_block   per-128-row partial reductions
//
// MFMA: v_mfma_f64_16x16x4_f64.  Operand A (16x4): lane l holds A[l%16][l/16];
// B (4x16): lane l holds B[l/16][l%16]; D (16x16): lane l, element v holds
// D[l/16 + 4v][l%16] (checked on the device by the parity tests: any other
// assignment permutes rows inside each 16-row tile).  Every K=16 step is
// issued as four MFMAs in which
// lane group g = l/16 supplies k = 4g+s (s = 0..3): the contraction over k
// does not depend on the order, and it lets each lane fetch its four k
// values with one 32-byte load from both operands.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>
#include <float.h>
#include <mutex>
#include <math.h>

#include "fgls_device.h"
#include "fgls_hip.h"
#include "fgls_internal.h"
#include "fgls_ctx.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NB = 64;    // diagonal block of the TRSM
using fgls::TM;    // TRMM tile rows (2 waves of 64)
using fgls::TN;    // TRMM tile columns (2 waves of 64)
using fgls::PMAX;  // largest p supported by the per-SNP kernels
using fgls::trait_res_off;

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4d ld4(const double* p) {
  return *reinterpret_cast<const v4d*>(p);
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// ---------------------------------------------------------------------------
// Inverses of the 64x64 diagonal blocks of L: the diagonal blocks of
// T = L^-1 (row-major lower triangle of Linv, zero above), copied into Dinv
// (block I column-major at Dinv + I*4096); rows beyond n get the identity so
// that the padded solve is harmless.  One workgroup per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag_copy(int n, const double* __restrict__ Linv, int64_t ldli,
                                                   double* __restrict__ Dinv) {
  __shared__ double t[NB][NB + 1];
  const int base = blockIdx.x * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += 256) {
    const int r = idx / NB, c = idx % NB;
    t[r][c] = (base + r < n) ? Linv[(size_t)(base + r) * ldli + base + c] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  double* out = Dinv + (size_t)blockIdx.x * NB * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += 256) {
    const int c = idx / NB, r = idx % NB;
    out[(size_t)c * NB + r] = t[r][c];
  }
}

// ---------------------------------------------------------------------------
// Blocked left-looking TRSM, L Z = B in place (B column-major, ld ldb), one
// workgroup per panel of 64 columns.
// For every 64-row block I:   R = B_I - sum_{J<I} L_IJ Z_J   (FP64 MFMA)
//                             Z_I = Dinv_I R                 (FP64 MFMA, R in LDS)
// The sweep over row blocks is sequential inside a panel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trsm_panel(int n, int nrhs, const double* __restrict__ L,
                                                    int64_t ldl, const double* __restrict__ Dinv,
                                                    double* Z, int64_t ldz) {
  constexpr int NC = NB, TJ = NC / 16;
  __shared__ double Rs[NB][NC + 2];
  const int P = blockIdx.x;
  const int c0 = P * NC;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int nblk = (n + NB - 1) / NB;
  int colc[TJ];
#pragma unroll
  for (int tj = 0; tj < TJ; ++tj) colc[tj] = min(c0 + 16 * tj + li, nrhs - 1);

  for (int I = 0; I < nblk; ++I) {
    v4d acc[TJ];
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) acc[tj] = v4d{0.0, 0.0, 0.0, 0.0};
    const int rowA = I * NB + 16 * w + li;
    const bool rowok = rowA < n;
    for (int k0 = 0; k0 < I * NB; k0 += 16) {
      double a[4];
#pragma unroll
      for (int s = 0; s < 4; ++s)
        a[s] = rowok ? L[(size_t)(k0 + 4 * lg + s) * ldl + rowA] : 0.0;
      double b[TJ][4];
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj) {
#pragma unroll
        for (int s = 0; s < 4; ++s) b[tj][s] = Z[(size_t)colc[tj] * ldz + k0 + 4 * lg + s];
      }
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj) acc[tj] = mfma_f64(a[s], b[tj][s], acc[tj]);
    }
    // R = B_I - acc into LDS (natural 64 x NC layout)
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * w + lg + 4 * v, c = 16 * tj + li;
        const int grow = I * NB + r, gcol = c0 + c;
        const double bv = (grow < n && gcol < nrhs) ? Z[(size_t)gcol * ldz + grow] : 0.0;
        Rs[r][c] = bv - acc[tj][v];
      }
    }
    __syncthreads();
    // Z_I = Dinv_I * R
    v4d zc[TJ];
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) zc[tj] = v4d{0.0, 0.0, 0.0, 0.0};
    const double* D = Dinv + (size_t)I * NB * NB;
#pragma unroll
    for (int k0 = 0; k0 < NB; k0 += 16) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = k0 + 4 * lg + s;
        const double a = D[(size_t)k * NB + 16 * w + li];
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj) zc[tj] = mfma_f64(a, Rs[k][16 * tj + li], zc[tj]);
      }
    }
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int grow = I * NB + 16 * w + lg + 4 * v, gcol = c0 + 16 * tj + li;
        if (grow < n && gcol < nrhs) Z[(size_t)gcol * ldz + grow] = zc[tj][v];
      }
    }
    // Z_I is read back by other waves of this workgroup in the next blocks.
    __threadfence();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Triangular product Z = Linv * X over 128x128 tiles (2x2 waves of 64x64),
// FP64 MFMA, operands fetched straight into registers (one K=16 step ahead).
// Linv is row-major (ld ldli >= RT*128) and lower triangular: row tile rt
// only needs k < (rt+1)*128 (DENSE: a full row-major matrix, all k < n;
// UPPER: upper triangular, k from the wave's first row).
// X is column-major with ld ldx (% 4 == 0); rows k >= n are masked.
// Epilogue:
//   EPI_STORE : Z written column-major (ld ldz)
//   EPI_REDUCE: per column, the NQ sums over the tile's rows of
//               z*z and z*W[:,k] (k < NQ-1) go to part[(rt*NQ+q)*ldp + col]
//   EPI_DOT   : as EPI_REDUCE with z*X[row, col] (rows < n) in place of z*z
//   EPI_MULTI : ntr groups of NQ weight columns (group k at W + k*NQ*ldw):
//               column 0 weighs z*z, columns 1..NQ-1 weigh z; the sums of
//               group k go to part[((rt*ntr + k)*NQ + q)*ldp + col]
// Grid: (column tiles, row tiles), heaviest row tiles first.
// ---------------------------------------------------------------------------
enum { EPI_STORE = 0, EPI_REDUCE = 1, EPI_DOT = 2, EPI_MULTI = 3 };

template <int EPI, int NQ, bool DENSE = false, bool UPPER = false, bool LOWER_OUT = false>
__global__ __launch_bounds__(256) void k_trmm(int n, int count, int RT,
                                              const double* __restrict__ Linv, int64_t ldli,
                                              const double* __restrict__ X, int64_t ldx,
                                              double* __restrict__ Zout, int64_t ldz,
                                              const double* __restrict__ W, int64_t ldw,
                                              double* __restrict__ part, int64_t ldp,
                                              int ntr = 1) {
  const int ct = blockIdx.x;
  const int rt = UPPER ? (int)blockIdx.y : RT - 1 - (int)blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w >> 1, wc = w & 1;
  const int li = lane & 15, lg = lane >> 4;
  const int r0 = rt * TM + wr * 64;
  const int c0 = ct * TN + wc * 64;
  // LOWER_OUT (EPI_STORE, a symmetric result): only the 64 x 64 blocks on or
  // below the diagonal are formed; the caller mirrors the others
  static_assert(!LOWER_OUT || EPI == EPI_STORE, "the lower-half store has no reductions");
  if (LOWER_OUT && c0 > r0) return;
  const int n16 = (n + 15) & ~15;
  const int kend = (r0 < n16) ? ((DENSE || UPPER) ? n16 : min(r0 + 64, n16)) : 0;   // multiple of 16
  const int kfull = min(kend, n & ~15);                   // steps with no masking
  // UPPER: an upper triangular A (zeros below the diagonal): k from the
  // wave's first row
  const int kb = (UPPER && r0 < kfull) ? r0 : 0;

  const double* Ap[4];
  const double* Bp[4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti) Ap[ti] = Linv + (size_t)(r0 + 16 * ti + li) * ldli + 4 * lg;
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    const int col = min(c0 + 16 * tj + li, count - 1);
    Bp[tj] = X + (size_t)col * ldx + 4 * lg;
  }

  v4d acc[4][4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};

  // Main K loop: operands of step k0 + 16 are loaded while the 64 MFMAs of
  // step k0 run.  The prefetch is unconditional inside the loop and the last
  // full step is peeled, so the accumulators stay in AGPRs across iterations.
  v4d a[4], b[4];
  if (kfull > 0) {
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) a[ti] = ld4(Ap[ti] + kb);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) b[tj] = ld4(Bp[tj] + kb);
    for (int k0 = kb + 16; k0 < kfull; k0 += 16) {
      v4d an[4], bn[4];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) an[ti] = ld4(Ap[ti] + k0);
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) bn[tj] = ld4(Bp[tj] + k0);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
          for (int tj = 0; tj < 4; ++tj)
            acc[ti][tj] = (EPI == EPI_STORE) ? mfma_f64(b[tj][s], a[ti][s], acc[ti][tj])
                                             : mfma_f64(a[ti][s], b[tj][s], acc[ti][tj]);
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) a[ti] = an[ti];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) b[tj] = bn[tj];
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
          acc[ti][tj] = (EPI == EPI_STORE) ? mfma_f64(b[tj][s], a[ti][s], acc[ti][tj])
                                           : mfma_f64(a[ti][s], b[tj][s], acc[ti][tj]);
  }
  if (kend > kfull) {  // last partial K step: rows k >= n of X are masked
    const int k0 = kfull;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) a[ti] = ld4(Ap[ti] + k0);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = k0 + 4 * lg + s;
        b[tj][s] = (k < n) ? Bp[tj][k0 + s] : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
          acc[ti][tj] = (EPI == EPI_STORE) ? mfma_f64(b[tj][s], a[ti][s], acc[ti][tj])
                                           : mfma_f64(a[ti][s], b[tj][s], acc[ti][tj]);
  }

  if constexpr (EPI == EPI_STORE) {
    // operands swapped: element v of acc[ti][tj] is z(row r0+16ti+li, column
    // c0+16tj+lg+4v), so 16 lanes store 16 consecutive rows of one column.
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = c0 + 16 * tj + lg + 4 * v;
        if (col < count) {
#pragma unroll
          for (int ti = 0; ti < 4; ++ti) Zout[(size_t)col * ldz + r0 + 16 * ti + li] = acc[ti][tj][v];
        }
      }
    }
  } else if constexpr (EPI == EPI_MULTI) {
    // element v of acc[ti][tj] is z(row r0+16ti+lg+4v, column c0+16tj+li);
    // one pass over the accumulators per weight group
    __shared__ double red[2][64][NQ];
    for (int k = 0; k < ntr; ++k) {
      const double* Wk = W + (size_t)k * NQ * ldw;
      double q[4][NQ];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) q[tj][qi] = 0.0;
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + 16 * ti + lg + 4 * v;
          double wv[NQ];
#pragma unroll
          for (int j = 0; j < NQ; ++j) wv[j] = Wk[(size_t)j * ldw + row];
#pragma unroll
          for (int tj = 0; tj < 4; ++tj) {
            const double z = acc[ti][tj][v];
            q[tj][0] = fma(z * wv[0], z, q[tj][0]);
#pragma unroll
            for (int j = 1; j < NQ; ++j) q[tj][j] = fma(z, wv[j], q[tj][j]);
          }
        }
      }
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) {
          double x = q[tj][qi];
          x += __shfl_xor(x, 16, 64);
          x += __shfl_xor(x, 32, 64);
          q[tj][qi] = x;
        }
      if (wr == 1 && lg == 0) {
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
          for (int qi = 0; qi < NQ; ++qi) red[wc][16 * tj + li][qi] = q[tj][qi];
      }
      __syncthreads();
      if (wr == 0 && lg == 0) {
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
          const int col = c0 + 16 * tj + li;
          if (col < count) {
#pragma unroll
            for (int qi = 0; qi < NQ; ++qi)
              part[(((size_t)rt * ntr + k) * NQ + qi) * ldp + col] = q[tj][qi] + red[wc][16 * tj + li][qi];
          }
        }
      }
      __syncthreads();   // red is refilled by the next group
    }
  } else {
    // element v of acc[ti][tj] is z(row r0+16ti+lg+4v, column c0+16tj+li)
    __shared__ double red[2][64][NQ];
    double q[4][NQ];
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int qi = 0; qi < NQ; ++qi) q[tj][qi] = 0.0;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + 16 * ti + lg + 4 * v;
        double wv[NQ - 1];
#pragma unroll
        for (int k = 0; k < NQ - 1; ++k) wv[k] = W[(size_t)k * ldw + row];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
          const double z = acc[ti][tj][v];
          if (EPI == EPI_DOT) {
            const double x = (row < n) ? Bp[tj][(size_t)row - 4 * lg] : 0.0;
            q[tj][0] = fma(z, x, q[tj][0]);
          } else {
            q[tj][0] = fma(z, z, q[tj][0]);
          }
#pragma unroll
          for (int k = 0; k < NQ - 1; ++k) q[tj][1 + k] = fma(z, wv[k], q[tj][1 + k]);
        }
      }
    }
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int qi = 0; qi < NQ; ++qi) {
        double x = q[tj][qi];
        x += __shfl_xor(x, 16, 64);
        x += __shfl_xor(x, 32, 64);
        q[tj][qi] = x;
      }
    if (wr == 1 && lg == 0) {
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) red[wc][16 * tj + li][qi] = q[tj][qi];
    }
    __syncthreads();
    if (wr == 0 && lg == 0) {
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        const int col = c0 + 16 * tj + li;
        if (col < count) {
#pragma unroll
          for (int qi = 0; qi < NQ; ++qi)
            part[((size_t)rt * NQ + qi) * ldp + col] = q[tj][qi] + red[wc][16 * tj + li][qi];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Skinny triangular product Z = Linv X for q <= 16 columns (the trait's
// [X~_L, y~]): bound by reading the lower triangle of Linv (row-major, ld
// ldli, strict upper triangle zero) once from HBM.  One workgroup per 4
// consecutive rows, its 4 waves taking interleaved 64-column slices of the
// k range (so that the longest rows are split four ways); every row
// r < gridDim.x * 4 of Z is written with the sum over k < min(last row of the
// group + 1, n), reduced across each wave and then over the waves in a fixed
// order.  Grid: ldn / 4 blocks, 256 threads.
// ---------------------------------------------------------------------------
constexpr int TRMV_ROWS = 4;
constexpr int TRMV_QMAX = 16;

__global__ __launch_bounds__(256) void k_trmv_lower(int n, int q, const double* __restrict__ Linv,
                                                    int64_t ldli, const double* __restrict__ X, int64_t ldx,
                                                    double* __restrict__ Z, int64_t ldz) {
  __shared__ double red[4][TRMV_ROWS][TRMV_QMAX];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = blockIdx.x * TRMV_ROWS;
  const int kend = min(r0 + TRMV_ROWS, n);
  double acc[TRMV_ROWS][TRMV_QMAX];
#pragma unroll
  for (int i = 0; i < TRMV_ROWS; ++i)
#pragma unroll
    for (int j = 0; j < TRMV_QMAX; ++j) acc[i][j] = 0.0;
  for (int k = threadIdx.x; k < kend; k += 256) {
    double a[TRMV_ROWS];
#pragma unroll
    for (int i = 0; i < TRMV_ROWS; ++i) a[i] = Linv[(size_t)(r0 + i) * ldli + k];
#pragma unroll
    for (int j = 0; j < TRMV_QMAX; ++j) {
      if (j < q) {
        const double x = X[(size_t)j * ldx + k];
#pragma unroll
        for (int i = 0; i < TRMV_ROWS; ++i) acc[i][j] = fma(a[i], x, acc[i][j]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < TRMV_QMAX; ++j) {
    if (j < q) {
#pragma unroll
      for (int i = 0; i < TRMV_ROWS; ++i) {
        const double s = wave_sum(acc[i][j]);
        if (lane == 0) red[w][i][j] = s;
      }
    }
  }
  __syncthreads();
  const int t = threadIdx.x, i = t / TRMV_QMAX, j = t % TRMV_QMAX;
  if (t < TRMV_ROWS * TRMV_QMAX && j < q)
    Z[(size_t)j * ldz + r0 + i] = ((red[0][i][j] + red[1][i][j]) + red[2][i][j]) + red[3][i][j];
}

// ---------------------------------------------------------------------------
// Mixed precision: L^-1 as S int8 digit slices, genotypes as int8.
//
// Row i of L^-1 is scaled by sc_i = 127 / max_k |L^-1_ik| and written in
// balanced radix R = FGLS_SLICE_RADIX = 254:
//   L^-1_ik sc_i = sum_{s<S} a_s,ik R^(-s) + r_S R^(-S),  a_s,ik in [-127, 127],
// a_s = rint(r_s), r_{s+1} = R (r_s - a_s), so |r_s| <= 127 at every level
// and the digits use the whole int8 range (6 digits: 127 R^-6 / 127 = 3.7e-15
// of the row maximum; the FP64 rounding of the scaling and of R (r - a) is
// 2^-52 relative, far below).
// z_i = L^-1 x for integer genotypes x is then (sum_s R^(-s) (A_s x)_i) / sc_i,
// where every A_s x is an exact int32 product on the int8 MFMA
// (|a x| <= 127*127 per term, exact for n <= 133,000, checked at set_mode).
//
// Slice layout: fgls_a8_offset (1 KB blocks in MFMA A-operand lane order, so
// one 64-k step of a 16-row tile and one slice is a single contiguous load).
// ---------------------------------------------------------------------------
typedef fgls_v4i v4i;

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
  return x;
}

template <int S>
__global__ __launch_bounds__(256) void k_slice_linv(int ldn, const double* __restrict__ Linv,
                                                    int8_t* __restrict__ A8,
                                                    double* __restrict__ inv_scale) {
  __shared__ double red[4];
  const int row = blockIdx.x, tid = threadIdx.x;
  const double* Lr = Linv + (size_t)row * ldn;
  double mx = 0.0;
  for (int k = tid; k < ldn; k += 256) mx = fmax(mx, fabs(Lr[k]));
  mx = wave_max(mx);
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double sc = (mx > 0.0) ? 127.0 / mx : 1.0;   // |L^-1_ik| sc <= 127
  if (tid == 0) inv_scale[row] = (mx > 0.0) ? mx / 127.0 : 0.0;
  const int KC = ldn / 16;
  for (int kc = tid; kc < KC; kc += 256) {
    int8_t out[S][16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      double r = Lr[kc * 16 + b] * sc;
#pragma unroll
      for (int s2 = 0; s2 < S; ++s2) {
        const double a = fmin(fmax(rint(r), -127.0), 127.0);
        out[s2][b] = (int8_t)(int)a;
        r = (r - a) * (double)FGLS_SLICE_RADIX;
      }
    }
#pragma unroll
    for (int s2 = 0; s2 < S; ++s2)
      *reinterpret_cast<v4i*>(A8 + fgls_a8_offset(row, s2, kc, S, ldn / 64)) =
          *reinterpret_cast<const v4i*>(out[s2]);
  }
}

// Five-digit form of the triangular L^-1 (S = 5).  Row i is scaled by
//   sc_i = min(127 / off_i, 127 R / max_i),
// off_i the largest |L^-1_ik| outside the row's diagonal 64-k step and max_i
// the largest of the row: outside that step |L^-1_ik sc_i| <= 127 and five
// digits a_0..a_4 follow as in k_slice_linv; inside it |L^-1_ik sc_i| <=
// 127 R, and a top digit t = rint(r / R) (|t| <= 127) is split off first
// (r - R t is exact in FP64 and within [-127, 127]), then a_0..a_4.  The
// residual is 127 R^-5 / sc_i per entry, R off_i / max_i (<= the loss) times
// the six-digit form's 127 R^-6 / (127 / max_i); the largest such factor over
// the rows goes to *loss_bits (non-negative doubles order as their bits), and
// the host keeps this form only while it is at most FGLS_TOP_MAX_LOSS.  The
// top digits of a 16-row group's diagonal step go to A8top
// (fgls_a8top_offset), its five digits to A8 like every other step
// (k_slice_tri with TOP).
// The slices of the triangular L^-1 (row-major, ld ldn, zero above the
// diagonal) in either form: the digits of k_slice_linv (S = 6, TOP false)
// or the five-digit form above (S = 5, TOP), for the steps the triangular int8
// product reads: a 16-row group's 64-k steps up to its diagonal step (the
// later ones are neither read nor written).  One workgroup per 16-row group.
// Phase 1: wave w takes rows 4 w .. 4 w + 3, its lanes along k < the end of
// the diagonal step, for the row maximum (TOP: and the maximum before the
// diagonal step).  Phase 2: wave w forms the steps w, w + 4, ...; lane
// 16 q + r takes row r and the 16 k's of chunk 4 step + q, so that the wave
// writes each digit's 1 KB block of the step contiguously.
template <int S, bool TOP>
__global__ __launch_bounds__(256) void k_slice_tri(int ldn, const double* __restrict__ Linv,
                                                   int8_t* __restrict__ A8, int8_t* __restrict__ A8top,
                                                   double* __restrict__ inv_scale,
                                                   unsigned long long* __restrict__ loss_bits) {
  __shared__ double shsc[16];
  constexpr double R = (double)FGLS_SLICE_RADIX;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int row0 = 16 * (int)blockIdx.x;
  const int d0 = row0 & ~63;     // the group's diagonal step
  const int nst = d0 / 64 + 1;   // steps 0 .. diagonal
  const int KS = ldn / 64;
  for (int i = 0; i < 4; ++i) {
    const int r = 4 * w + i;
    const double* Lr = Linv + (size_t)(row0 + r) * ldn;
    double mx = 0.0, off = 0.0;
    // 64-column chunks j = 0 .. nst - 1 (the last is the diagonal step),
    // eight loads in flight per lane
    int j = 0;
    for (; j + 8 <= nst; j += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = Lr[64 * (j + u) + lane];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double a = fabs(v[u]);
        mx = fmax(mx, a);
        if (TOP && j + u < nst - 1) off = fmax(off, a);
      }
    }
    for (; j < nst; ++j) {
      const double a = fabs(Lr[64 * j + lane]);
      mx = fmax(mx, a);
      if (TOP && j < nst - 1) off = fmax(off, a);
    }
    mx = wave_max(mx);
    if (TOP) off = wave_max(off);
    if (lane == 0) {
      double sc, isc;
      if constexpr (TOP) {
        sc = 1.0;
        if (mx > 0.0) {
          sc = 127.0 * R / mx;
          if (off > 0.0) sc = fmin(sc, 127.0 / off);
        }
        isc = (mx > 0.0) ? 1.0 / sc : 0.0;
        const double loss = (mx > 0.0) ? R * off / mx : 0.0;
        atomicMax(loss_bits, (unsigned long long)__double_as_longlong(loss));
      } else {
        sc = (mx > 0.0) ? 127.0 / mx : 1.0;   // |L^-1_ik| sc <= 127
        isc = (mx > 0.0) ? mx / 127.0 : 0.0;
      }
      shsc[r] = sc;
      inv_scale[row0 + r] = isc;
    }
  }
  __syncthreads();
  const int r = lane & 15, q = lane >> 4;
  const int row = row0 + r;
  const double sc = shsc[r];
  const double* Lr = Linv + (size_t)row * ldn;
  for (int st = w; st < nst; st += 4) {
    const int kc = 4 * st + q;
    const bool diag = TOP && st == nst - 1;
    double v[16];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const v4d x = *reinterpret_cast<const v4d*>(Lr + 16 * kc + 4 * j);
#pragma unroll
      for (int e = 0; e < 4; ++e) v[4 * j + e] = x[e];
    }
    int8_t out[S][16], top[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      double x = v[b] * sc;
      double t = 0.0;
      if (diag) {
        t = fmin(fmax(rint(x / R), -127.0), 127.0);
        x -= R * t;
      }
      top[b] = (int8_t)(int)t;
#pragma unroll
      for (int s2 = 0; s2 < S; ++s2) {
        const double a = fmin(fmax(rint(x), -127.0), 127.0);
        out[s2][b] = (int8_t)(int)a;
        x = (x - a) * R;
      }
    }
#pragma unroll
    for (int s2 = 0; s2 < S; ++s2)
      *reinterpret_cast<v4i*>(A8 + fgls_a8_offset(row, s2, kc, S, KS)) = *reinterpret_cast<const v4i*>(out[s2]);
    if (diag) *reinterpret_cast<v4i*>(A8top + fgls_a8top_offset(row, kc)) = *reinterpret_cast<const v4i*>(top);
  }
}

// FP64 genotypes -> int8 (column-major, ld ldx8), exact only: a value that is
// not an integer in [-127, 127] is counted in *bad.
__global__ __launch_bounds__(256) void k_to_i8(int n, const double* __restrict__ X,
                                               int64_t ldx, int8_t* __restrict__ X8, int64_t ldx8,
                                               unsigned int* __restrict__ bad) {
  const int col = blockIdx.x;   // one workgroup per column
  unsigned int nb = 0;
  for (int r = threadIdx.x; r < ldx8; r += 256) {
    int8_t o = 0;
    if (r < n) {
      const double v = X[(size_t)col * ldx + r];
      const double rv = rint(v);
      if (!(rv == v && fabs(v) <= 127.0)) ++nb;
      else o = (int8_t)(int)rv;
    }
    X8[(size_t)col * ldx8 + r] = o;
  }
  if (nb) atomicAdd(bad, nb);
}

// FP64 dosages -> two int8 digit planes: v = rint(x * scale) = 128 hi + lo,
// lo in [0, 127] at X8, hi in [-127, 127] at X8 + plane_off (rows n..ldx8-1
// zeroed).  A value with |x * scale - v| > 8 eps max(1, |x * scale|) (eps =
// 2^-52: x holds the nearest double to v / scale, and x * scale adds one
// more rounding, so an exact multiple of 1/scale is off by at most
// 2 eps |x * scale|) or v outside [-16256, 16383] is counted in *bad and
// written as 0.
__global__ __launch_bounds__(256) void k_to_dosage_i8(int n, const double* __restrict__ X, int64_t ldx,
                                                      double scale, int8_t* __restrict__ X8,
                                                      int64_t ldx8, int64_t plane_off,
                                                      unsigned int* __restrict__ bad) {
  const int col = blockIdx.x;   // one workgroup per column
  unsigned int nb = 0;
  for (int r = threadIdx.x; r < ldx8; r += 256) {
    int lo = 0, hi = 0;
    if (r < n) {
      const double y = X[(size_t)col * ldx + r] * scale;
      const double v = rint(y);
      if (!(fabs(y - v) <= 8.0 * DBL_EPSILON * fmax(1.0, fabs(y)) && v >= -16256.0 && v <= 16383.0)) {
        ++nb;
      } else {
        const int iv = (int)v;
        hi = iv >> 7;     // floor(v / 128)
        lo = iv & 127;
      }
    }
    X8[(size_t)col * ldx8 + r] = (int8_t)lo;
    X8[plane_off + (size_t)col * ldx8 + r] = (int8_t)hi;
  }
  if (nb) atomicAdd(bad, nb);
}

// 2-bit genotypes -> int8 (column-major, ld ldx8, rows n..ldx8-1 zeroed):
// column col is ceil(n/4) bytes at P + col*ldp, row r in bits 2(r%4)+1..2(r%4)
// of byte r/4; codes 0, 1, 2 are the genotype, code 3 is counted in *bad and
// written as 0.  One workgroup per column, one packed byte per thread.
__global__ __launch_bounds__(256) void k_u2_to_i8(int n, const uint8_t* __restrict__ P, int64_t ldp,
                                                  int8_t* __restrict__ X8, int64_t ldx8,
                                                  unsigned int* __restrict__ bad) {
  const int col = blockIdx.x;
  const int nb = (n + 3) / 4;
  unsigned int nbad = 0;
  for (int b = threadIdx.x; b < (int)(ldx8 / 4); b += 256) {
    const unsigned int byte = (b < nb) ? P[(size_t)col * ldp + b] : 0u;
    unsigned int out = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const unsigned int c = (byte >> (2 * j)) & 3u;
      if (4 * b + j < n) {
        if (c == 3u) ++nbad;
        else out |= c << (8 * j);
      }
    }
    *reinterpret_cast<unsigned int*>(X8 + (size_t)col * ldx8 + 4 * b) = out;
  }
  if (nbad) atomicAdd(bad, nbad);
}

// 16-bit dosage file values v = x * scale (column-major, ld ldv) -> the two
// int8 digit planes of the mixed path (v = 128 hi + lo, lo at X8, hi at
// X8 + plane_off; rows n..ldx8-1 zeroed); v outside [-16256, 16383] is
// counted in *bad and written as 0.
__global__ __launch_bounds__(256) void k_i16_to_dosage_i8(int n, const int16_t* __restrict__ V,
                                                          int64_t ldv, int8_t* __restrict__ X8,
                                                          int64_t ldx8, int64_t plane_off,
                                                          unsigned int* __restrict__ bad) {
  const int col = blockIdx.x;
  unsigned int nb = 0;
  for (int r = threadIdx.x; r < ldx8; r += 256) {
    int lo = 0, hi = 0;
    if (r < n) {
      const int v = V[(size_t)col * ldv + r];
      if (v < -16256 || v > 16383) {
        ++nb;
      } else {
        hi = v >> 7;
        lo = v & 127;
      }
    }
    X8[(size_t)col * ldx8 + r] = (int8_t)lo;
    X8[plane_off + (size_t)col * ldx8 + r] = (int8_t)hi;
  }
  if (nb) atomicAdd(bad, nb);
}

// 16-bit dosage file values -> FP64 x = v / scale (rows n..ldx-1 zeroed).
__global__ __launch_bounds__(256) void k_i16_to_f64(int n, const int16_t* __restrict__ V, int64_t ldv,
                                                    double scale, double* __restrict__ X, int64_t ldx) {
  const int col = blockIdx.x;
  for (int r = threadIdx.x; r < ldx; r += 256)
    X[(size_t)col * ldx + r] = (r < n) ? (double)V[(size_t)col * ldv + r] / scale : 0.0;
}

// int8 genotypes as given (column-major, ld ldx8): values of -128 in rows
// < n are counted in *bad (the int8 product's exactness bound assumes
// |x| <= 127).  One workgroup per column, one byte per thread and step.
__global__ __launch_bounds__(256) void k_check_i8(int n, const int8_t* __restrict__ X8, int64_t ldx8,
                                                  unsigned int* __restrict__ bad) {
  const int col = blockIdx.x;
  unsigned int nb = 0;
  for (int r = threadIdx.x; r < n; r += 256) nb += (X8[(size_t)col * ldx8 + r] == -128) ? 1u : 0u;
  if (nb) atomicAdd(bad, nb);
}

// int8 genotypes -> FP64 (column-major, ld ldx; rows n..ldx-1 written as 0).
__global__ __launch_bounds__(256) void k_i8_to_f64(int n, const int8_t* __restrict__ X8, int64_t ldx8,
                                                   double* __restrict__ X, int64_t ldx) {
  const int col = blockIdx.x;   // one workgroup per column
  for (int r = threadIdx.x; r < ldx; r += 256)
    X[(size_t)col * ldx + r] = (r < n) ? (double)X8[(size_t)col * ldx8 + r] : 0.0;
}

// ---------------------------------------------------------------------------
// p x p tail of the per-SNP stage (fgls_chol.c:401-438): Cholesky of the
// lower triangle of S, two triangular solves for b, inverse through the
// factor (trtri + L'L product, as dpotri does), scaling by res_sigma and
// sqrt of the diagonal.  S is column-major-lower in registers; all indices
// are compile-time.  Returns the LAPACK-style info (0 = success).
// ---------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ int finalize_snp(double (&S)[P][P], double (&b)[P], double res_sigma,
                                            double* __restrict__ Bo, double* __restrict__ Vo) {
  int info = 0;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    double d = S[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= S[j][k] * S[j][k];
    if (!(d > 0.0) && info == 0) info = j + 1;
    d = sqrt(d);
    S[j][j] = d;
#pragma unroll
    for (int i = j + 1; i < P; ++i) {
      double s = S[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= S[i][k] * S[j][k];
      S[i][j] = s / d;
    }
  }
  if (info) {
    const double nan = __builtin_nan("");
#pragma unroll
    for (int i = 0; i < P; ++i) Bo[i] = nan;
#pragma unroll
    for (int i = 0; i < P * P; ++i) Vo[i] = nan;
    return info;
  }
  // C y = b, then C' x = y
#pragma unroll
  for (int i = 0; i < P; ++i) {
    double s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= S[i][k] * b[k];
    b[i] = s / S[i][i];
  }
#pragma unroll
  for (int i = P - 1; i >= 0; --i) {
    double s = b[i];
#pragma unroll
    for (int k = i + 1; k < P; ++k) s -= S[k][i] * b[k];
    b[i] = s / S[i][i];
  }
#pragma unroll
  for (int i = 0; i < P; ++i) Bo[i] = b[i];
  // C^-1 in place (lower, non-unit), columns right to left
#pragma unroll
  for (int j = P - 1; j >= 0; --j) {
    S[j][j] = 1.0 / S[j][j];
    const double ajj = -S[j][j];
    // x = S[j+1:, j]; x <- T x with T = S[j+1:, j+1:] (already inverted), bottom-up
#pragma unroll
    for (int i = P - 1; i > j; --i) {
      double s = 0.0;
#pragma unroll
      for (int k = j + 1; k <= i; ++k) s += S[i][k] * S[k][j];
      S[i][j] = s;
    }
#pragma unroll
    for (int i = j + 1; i < P; ++i) S[i][j] *= ajj;
  }
  // lower(C^-T C^-1) in place, rows top to bottom (dlauu2, lower)
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const double aii = S[i][i];
    double d = 0.0;
#pragma unroll
    for (int k = i; k < P; ++k) d += S[k][i] * S[k][i];
    S[i][i] = d;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      double s = aii * S[i][j];
#pragma unroll
      for (int k = i + 1; k < P; ++k) s += S[k][i] * S[k][j];
      S[i][j] = s;
    }
  }
#pragma unroll
  for (int c = 0; c < P; ++c) {
#pragma unroll
    for (int r = 0; r < P; ++r) {
      double v;
      if (r > c) v = res_sigma * S[r][c];
      else if (r == c) v = sqrt(res_sigma * S[r][r]);
      else v = 0.0;
      Vo[c * P + r] = v;
    }
  }
  return 0;
}

__device__ __forceinline__ void record_failure(unsigned long long* err, int64_t snp, int info) {
  const unsigned long long key = ((unsigned long long)snp << 8) | (unsigned long long)(info & 0xff);
  atomicMin(err, key);
}

// Per-trait scalars live in a small device array:
//   [0, wXL)               B_t
//   [wXL, wXL + wXL^2)     V_tl (column-major, upper triangle 0)
//   [wXL + wXL^2]          res_sigma
//   [wXL + wXL^2 + 1, +wXL) b_L

// Finalize from the TRMM partials (wXR = 1, p = P): lane per SNP, one
// grid row per trait k of ntr (partials of trait k at
// part[((rt*ntr + k)*NQ + q)*ldp + i]; its scalars at trait + k*trait_ld,
// its outputs at Bout + k*ldb, Vout + k*ldv, its failure record at err[k]).
//
// Workgroup: FIN_SPLIT waves over the same 64 SNPs (lane = SNP), wave w
// summing the row tiles rt = w, w + FIN_SPLIT, ... (FIN_SPLIT times the
// loads in flight of one wave per SNP); wave 0 then adds the other waves'
// sums in wave order (a fixed order: the result does not depend on timing)
// and runs the p x p stage.
constexpr int FIN_SPLIT = 4;
template <int P>
__global__ __launch_bounds__(64 * FIN_SPLIT) void k_finalize(int count, int RT, const double* __restrict__ part,
                                                  int64_t ldp, int ntr, const double* __restrict__ trait,
                                                  int64_t trait_ld, double* __restrict__ Bout, int64_t ldb,
                                                  double* __restrict__ Vout, int64_t ldv,
                                                  unsigned long long* __restrict__ err, int64_t snp_base) {
  constexpr int NQ = P + 1;
  constexpr int WXL = P - 1;
  __shared__ double red[FIN_SPLIT - 1][NQ][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int k = blockIdx.y;
  const int ic = min(i, count - 1);   // lanes past count read a valid column and store nothing
  double sums[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) sums[q] = 0.0;
  for (int rt = w; rt < RT; rt += FIN_SPLIT) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) sums[q] += part[(((size_t)rt * ntr + k) * NQ + q) * ldp + ic];
  }
  if (w > 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) red[w - 1][q][lane] = sums[q];
  }
  __syncthreads();
  if (w > 0 || i >= count) return;
#pragma unroll
  for (int j = 0; j < FIN_SPLIT - 1; ++j)
#pragma unroll
    for (int q = 0; q < NQ; ++q) sums[q] += red[j][q][lane];
  trait += (size_t)k * trait_ld;
  Bout += (size_t)k * ldb;
  Vout += (size_t)k * ldv;
  double S[P][P], b[P];
#pragma unroll
  for (int r = 0; r < P; ++r)
#pragma unroll
    for (int c = 0; c < P; ++c) S[r][c] = 0.0;
#pragma unroll
  for (int r = 0; r < WXL; ++r) {
#pragma unroll
    for (int c = 0; c <= r; ++c) S[r][c] = trait[WXL + c * WXL + r];
    b[r] = trait[r];
  }
#pragma unroll
  for (int c = 0; c < WXL; ++c) S[WXL][c] = sums[1 + c];
  S[WXL][WXL] = sums[0];
  b[WXL] = sums[NQ - 1];
  const double res = trait[trait_res_off(WXL)];
  const int info = finalize_snp<P>(S, b, res, Bout + (size_t)i * P, Vout + (size_t)i * P * P);
  if (info) record_failure(err + k, snp_base + i, info);
}

// Per-SNP stage from transformed columns Z (any wXR): one wave per SNP,
// wavefront reductions over the n rows, then the p x p tail on lane 0.
// dwt (the eigen variant; nullptr for the Cholesky path): row weights of
// the Z'Z sums, sum_r dwt_r z_a(r) z_b(r); W then holds the weighted
// columns dwt [Z'X_L, Z'y], so that the z'W sums need no weight.
template <int P>
__global__ __launch_bounds__(256) void k_snp_stage(int n, int count, int wXL, int wXR,
                                                   const double* __restrict__ Z, int64_t ldz,
                                                   const double* __restrict__ W, int64_t ldw,
                                                   const double* __restrict__ trait,
                                                   double* __restrict__ Bout, double* __restrict__ Vout,
                                                   unsigned long long* __restrict__ err, int64_t snp_base,
                                                   const double* __restrict__ dwt) {
  __shared__ double Ssh[4][P * P + P];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int snp = blockIdx.x * 4 + w;
  if (snp >= count) return;  // no workgroup barrier below
  double* Sw = Ssh[w];
  for (int a = 0; a < wXR; ++a) {
    const double* za = Z + (size_t)(snp * wXR + a) * ldz;
    double t[P], u[P];
#pragma unroll
    for (int i = 0; i < P; ++i) t[i] = u[i] = 0.0;
    for (int r = lane; r < n; r += 64) {
      const double x = za[r];
      const double xt = dwt ? x * dwt[r] : x;
#pragma unroll
      for (int bb = 0; bb < P; ++bb)
        if (bb <= a) t[bb] = fma(xt, Z[(size_t)(snp * wXR + bb) * ldz + r], t[bb]);
#pragma unroll
      for (int k = 0; k < P; ++k)
        if (k <= wXL) u[k] = fma(x, W[(size_t)k * ldw + r], u[k]);
    }
#pragma unroll
    for (int i = 0; i < P; ++i) {
      t[i] = wave_sum(t[i]);
      u[i] = wave_sum(u[i]);
    }
    if (lane == 0) {
#pragma unroll
      for (int bb = 0; bb < P; ++bb)
        if (bb <= a) Sw[(wXL + a) + (wXL + bb) * P] = t[bb];
#pragma unroll
      for (int k = 0; k < P; ++k)
        if (k < wXL) Sw[(wXL + a) + k * P] = u[k];
      double uy = 0.0;
#pragma unroll
      for (int k = 0; k < P; ++k)
        if (k == wXL) uy = u[k];
      Sw[P * P + wXL + a] = uy;
    }
  }
  if (lane != 0) return;
  for (int r = 0; r < wXL; ++r) {
    for (int c = 0; c <= r; ++c) Sw[r + c * P] = trait[wXL + c * wXL + r];
    Sw[P * P + r] = trait[r];
  }
  double S[P][P], b[P];
#pragma unroll
  for (int r = 0; r < P; ++r) {
#pragma unroll
    for (int c = 0; c < P; ++c) S[r][c] = (c <= r) ? Sw[r + c * P] : 0.0;
    b[r] = Sw[P * P + r];
  }
  const double res = trait[trait_res_off(wXL)];
  const int info = finalize_snp<P>(S, b, res, Bout + (size_t)snp * P, Vout + (size_t)snp * P * P);
  if (info) record_failure(err, snp_base + snp, info);
}

// Per-trait scalars from W = [X~_L | y~]: B_t, V_tl, b_L, res_sigma.
// One workgroup; fixed summation order.
__device__ double block_sum(double x, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  x = wave_sum(x);
  __syncthreads();
  if (lane == 0) red[w] = x;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;
}

__global__ __launch_bounds__(256) void k_trait_stats(int n, int wXL, const double* __restrict__ W,
                                                     int64_t ldw, double* __restrict__ trait,
                                                     int* __restrict__ status) {
  __shared__ double red[4];
  __shared__ double S[PMAX * PMAX];
  __shared__ double bl[PMAX];
  const int tid = threadIdx.x;
  // V_tl (lower) and B_t
  for (int a = 0; a < wXL; ++a) {
    for (int c = 0; c <= a; ++c) {
      double s = 0.0;
      for (int r = tid; r < n; r += blockDim.x) s = fma(W[(size_t)a * ldw + r], W[(size_t)c * ldw + r], s);
      s = block_sum(s, red);
      if (tid == 0) S[a + c * PMAX] = s;
    }
    double s = 0.0;
    for (int r = tid; r < n; r += blockDim.x) s = fma(W[(size_t)a * ldw + r], W[(size_t)wXL * ldw + r], s);
    s = block_sum(s, red);
    if (tid == 0) bl[a] = s;
  }
  __syncthreads();
  if (tid == 0) {
    for (int a = 0; a < wXL; ++a) {
      trait[a] = bl[a];
      for (int c = 0; c < wXL; ++c) trait[wXL + c * wXL + a] = (c <= a) ? S[a + c * PMAX] : 0.0;
    }
    // b_L = V_tl^-1 B_t by Cholesky (fgls_chol.c:306-316)
    int info = 0;
    for (int j = 0; j < wXL && !info; ++j) {
      double d = S[j + j * PMAX];
      for (int k = 0; k < j; ++k) d -= S[j + k * PMAX] * S[j + k * PMAX];
      if (!(d > 0.0)) { info = j + 1; break; }
      d = sqrt(d);
      S[j + j * PMAX] = d;
      for (int i = j + 1; i < wXL; ++i) {
        double s = S[i + j * PMAX];
        for (int k = 0; k < j; ++k) s -= S[i + k * PMAX] * S[j + k * PMAX];
        S[i + j * PMAX] = s / d;
      }
    }
    status[0] = info;
    if (!info) {
      for (int i = 0; i < wXL; ++i) {
        double s = bl[i];
        for (int k = 0; k < i; ++k) s -= S[i + k * PMAX] * bl[k];
        bl[i] = s / S[i + i * PMAX];
      }
      for (int i = wXL - 1; i >= 0; --i) {
        double s = bl[i];
        for (int k = i + 1; k < wXL; ++k) s -= S[k + i * PMAX] * bl[k];
        bl[i] = s / S[i + i * PMAX];
      }
    }
    for (int i = 0; i < wXL; ++i) trait[trait_res_off(wXL) + 1 + i] = bl[i];
  }
  __syncthreads();
  // res_sigma = |y~ - X~_L b_L|^2 / (n - wXL)   (L^-1 is linear: L^-1(y - X_L b_L))
  double s = 0.0;
  for (int r = tid; r < n; r += blockDim.x) {
    double e = W[(size_t)wXL * ldw + r];
    for (int k = 0; k < wXL; ++k) e -= W[(size_t)k * ldw + r] * bl[k];
    s = fma(e, e, s);
  }
  s = block_sum(s, red);
  if (tid == 0) trait[trait_res_off(wXL)] = s / (double)(n - wXL);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(FGLS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace


namespace {

int check_ctx(const fgls_hip_ctx* c) {
  if (!c) return fail(FGLS_ERR_ARG, "NULL context");
  return FGLS_OK;
}

int set_device(const fgls_hip_ctx* c) {
  HIP_TRY(hipSetDevice(c->device));
  return FGLS_OK;
}

// The transformed columns L^-1 X_R of a block for wXR > 1 (ldn rows,
// max_block * wXR columns), allocated at the first such reduce.
int ensure_z(fgls_hip_ctx* c) {
  if (c->z) return FGLS_OK;
  const size_t bytes = (size_t)c->ldn * (size_t)c->max_block * c->wXR * sizeof(double);
  if (hipMalloc((void**)&c->z, bytes) != hipSuccess) {
    c->z = nullptr;
    return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
  }
  return FGLS_OK;
}

// b and V from RT row tiles of partial sums in `part` (leading dimension
// ldp): the context's own buffer, or the no-alloc entry point's scratch.
template <int P>
void launch_finalize_from(const fgls_hip_ctx* c, int RT, const double* part, int64_t ldp, int count,
                          double* dB, double* dV, int64_t base, hipStream_t st) {
  if (count <= 0) return;
  dim3 grid((count + 63) / 64);
  hipLaunchKernelGGL(k_finalize<P>, grid, dim3(64 * FIN_SPLIT), 0, st, count, RT, part, ldp, 1, c->trait,
                     (int64_t)0, dB, (int64_t)0, dV, (int64_t)0, c->err, base);
}

template <int P>
void launch_finalize(const fgls_hip_ctx* c, int count, double* dB, double* dV, int64_t base,
                     hipStream_t st) {
  launch_finalize_from<P>(c, c->last_rt, c->part, (int64_t)c->max_block, count, dB, dV, base, st);
}

// The FP64 triangular product with the fused per-row-tile reductions into
// `part` (leading dimension ldp).
template <int P>
void launch_trmm_reduce_into(const fgls_hip_ctx* c, const double* X, int64_t ldx, int count, double* part,
                             int64_t ldp, hipStream_t st) {
  dim3 grid((count + TN - 1) / TN, c->RT);
  hipLaunchKernelGGL((k_trmm<EPI_REDUCE, P + 1>), grid, dim3(256), 0, st, c->n, count, c->RT,
                     c->Linv, (int64_t)c->ldn, X, ldx, (double*)nullptr, (int64_t)0, c->W,
                     (int64_t)c->ldn, part, ldp);
}

template <int P>
void launch_trmm_reduce(const fgls_hip_ctx* c, const double* X, int64_t ldx, int count,
                        hipStream_t st) {
  launch_trmm_reduce_into<P>(c, X, ldx, count, c->part, (int64_t)c->max_block, st);
}

constexpr int NSLICE = FGLS_NSLICE;   // int8 digit slices of L^-1 in the mixed mode

// The slices of L^-1 in the form the context uses: the five-digit form
// (k_slice_tri with TOP) when the slice form is automatic, n admits it and its
// error bound over the six-digit one (the loss) is at most
// FGLS_TOP_MAX_LOSS; otherwise six digits.  Synchronizes `st`
// (the loss is read back) when it tries the five-digit form.
int build_slices(fgls_hip_ctx* c, hipStream_t st) {
  c->nslice = NSLICE;
  c->slice_loss = -1.0;
  if (c->slice_form == FGLS_SLICES_AUTO && c->n <= FGLS_I8_TOP_MAX_N) {
    unsigned long long* lb = reinterpret_cast<unsigned long long*>(c->invsc + c->ldn);
    HIP_TRY(hipMemsetAsync(lb, 0, sizeof(*lb), st));
    hipLaunchKernelGGL((k_slice_tri<FGLS_TOP_NSLICE, true>), dim3(c->ldn / 16), dim3(256), 0, st, c->ldn,
                       c->Linv, c->A8, c->A8 + fgls_a8top_base(c->ldn / 64), c->invsc, lb);
    HIP_TRY(hipGetLastError());
    double loss = 0.0;
    HIP_TRY(hipMemcpyAsync(&loss, lb, sizeof(loss), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    c->slice_loss = loss;
    if (loss <= FGLS_TOP_MAX_LOSS) {
      c->nslice = FGLS_TOP_NSLICE;
      return FGLS_OK;
    }
  }
  hipLaunchKernelGGL((k_slice_tri<NSLICE, false>), dim3(c->ldn / 16), dim3(256), 0, st, c->ldn, c->Linv, c->A8,
                     (int8_t*)nullptr, c->invsc, (unsigned long long*)nullptr);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

template <int P>
void launch_snp_stage(const fgls_hip_ctx* c, const double* Z, int64_t ldz, int count, double* dB,
                      double* dV, int64_t base, hipStream_t st) {
  dim3 grid((count + 3) / 4);
  hipLaunchKernelGGL(k_snp_stage<P>, grid, dim3(256), 0, st, c->n, count, c->wXL, c->wXR, Z, ldz,
                     c->W, (int64_t)c->ldn, c->trait, dB, dV, c->err, base, (const double*)nullptr);
}

template <int P>
void launch_snp_stage_weighted(int n, int count, int wXL, int wXR, const double* Z, int64_t ldz,
                               const double* W, int64_t ldw, const double* dwt, const double* trait,
                               double* dB, double* dV, unsigned long long* err, int64_t base,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_snp_stage<P>, dim3((count + 3) / 4), dim3(256), 0, st, n, count, wXL, wXR, Z, ldz, W,
                     ldw, trait, dB, dV, err, base, dwt);
}

bool aligned32(const void* p) { return ((uintptr_t)p & 31u) == 0; }

template <int P>
void launch_trmm_dot(const fgls_hip_ctx* c, const double* A, const double* X, int64_t ldx,
                     int count, hipStream_t st) {
  dim3 grid((count + TN - 1) / TN, c->RT);
  hipLaunchKernelGGL((k_trmm<EPI_DOT, P + 1, true>), grid, dim3(256), 0, st, c->n, count, c->RT, A,
                     (int64_t)c->ldn, X, ldx, (double*)nullptr, (int64_t)0, c->W, (int64_t)c->ldn,
                     c->part, (int64_t)c->max_block);
}

}  // namespace

namespace fgls {

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int set_device(const fgls_hip_ctx* c) { return ::set_device(c); }

void launch_product_store(const fgls_hip_ctx* c, const double* A, bool dense, const double* X,
                          int64_t ldx, int ncols, double* Z, int64_t ldz, hipStream_t st) {
  dim3 grid((ncols + TN - 1) / TN, c->RT);
  if (dense)
    hipLaunchKernelGGL((k_trmm<EPI_STORE, 2, true>), grid, dim3(256), 0, st, c->n, ncols, c->RT, A,
                       (int64_t)c->ldn, X, ldx, Z, ldz, c->W, (int64_t)c->ldn, (double*)nullptr,
                       (int64_t)0);
  else
    hipLaunchKernelGGL((k_trmm<EPI_STORE, 2>), grid, dim3(256), 0, st, c->n, ncols, c->RT, A,
                       (int64_t)c->ldn, X, ldx, Z, ldz, c->W, (int64_t)c->ldn, (double*)nullptr,
                       (int64_t)0);
}

void launch_product_store_lower_half(const fgls_hip_ctx* c, const double* X, int64_t ldx,
                                     double* Z, int64_t ldz, hipStream_t st) {
  dim3 grid((c->n + TN - 1) / TN, c->RT);
  hipLaunchKernelGGL((k_trmm<EPI_STORE, 2, false, false, true>), grid, dim3(256), 0, st, c->n, c->n,
                     c->RT, c->Linv, (int64_t)c->ldn, X, ldx, Z, ldz, c->W, (int64_t)c->ldn,
                     (double*)nullptr, (int64_t)0);
}

void launch_product_store_upper(const fgls_hip_ctx* c, const double* A, const double* X,
                                int64_t ldx, int ncols, double* Z, int64_t ldz, hipStream_t st) {
  dim3 grid((ncols + TN - 1) / TN, c->RT);
  hipLaunchKernelGGL((k_trmm<EPI_STORE, 2, false, true>), grid, dim3(256), 0, st, c->n, ncols, c->RT,
                     A, (int64_t)c->ldn, X, ldx, Z, ldz, c->W, (int64_t)c->ldn, (double*)nullptr,
                     (int64_t)0);
}

void launch_product_dot(const fgls_hip_ctx* c, const double* A, const double* X, int64_t ldx,
                        int count, hipStream_t st) {
  FGLS_P_SWITCH(c->p, launch_trmm_dot, c, A, X, ldx, count, st);
}

void launch_dense_store(int n, int ldn, const double* A, const double* X, int64_t ldx, int ncols,
                        double* Z, int64_t ldz, hipStream_t st) {
  dim3 grid((ncols + TN - 1) / TN, ldn / TM);
  hipLaunchKernelGGL((k_trmm<EPI_STORE, 2, true>), grid, dim3(256), 0, st, n, ncols, ldn / TM, A,
                     (int64_t)ldn, X, ldx, Z, ldz, (const double*)nullptr, (int64_t)0,
                     (double*)nullptr, (int64_t)0, 1);
}

template <int P>
void launch_eig_reduce_p(int n, int ldn, const double* A, const double* X, int64_t ldx, int count,
                         const double* W, int64_t ldw, int ntr, double* part, int64_t ldp,
                         hipStream_t st) {
  dim3 grid((count + TN - 1) / TN, ldn / TM);
  hipLaunchKernelGGL((k_trmm<EPI_MULTI, P + 1, true>), grid, dim3(256), 0, st, n, count, ldn / TM,
                     A, (int64_t)ldn, X, ldx, (double*)nullptr, (int64_t)0, W, ldw, part, ldp, ntr);
}

void launch_eig_reduce(int p, int n, int ldn, const double* A, const double* X, int64_t ldx,
                       int count, const double* W, int64_t ldw, int ntr, double* part, int64_t ldp,
                       hipStream_t st) {
  FGLS_P_SWITCH(p, launch_eig_reduce_p, n, ldn, A, X, ldx, count, W, ldw, ntr, part, ldp, st);
}

template <int P>
void launch_finalize_multi_p(int count, int RT, const double* part, int64_t ldp, int ntr,
                             const double* trait, int64_t trait_ld, double* dB, int64_t ldb,
                             double* dV, int64_t ldv, unsigned long long* err, int64_t base,
                             hipStream_t st) {
  if (count <= 0) return;
  dim3 grid((count + 63) / 64, ntr);
  hipLaunchKernelGGL(k_finalize<P>, grid, dim3(64 * FIN_SPLIT), 0, st, count, RT, part, ldp, ntr, trait,
                     trait_ld, dB, ldb, dV, ldv, err, base);
}

void launch_finalize_multi(int p, int count, int RT, const double* part, int64_t ldp, int ntr,
                           const double* trait, int64_t trait_ld, double* dB, int64_t ldb,
                           double* dV, int64_t ldv, unsigned long long* err, int64_t base,
                           hipStream_t st) {
  FGLS_P_SWITCH(p, launch_finalize_multi_p, count, RT, part, ldp, ntr, trait, trait_ld, dB, ldb,
                dV, ldv, err, base, st);
}

void launch_trait_stats(int n, int wXL, const double* W, int64_t ldw, double* trait, int* status,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_trait_stats, dim3(1), dim3(256), 0, st, n, wXL, W, ldw, trait, status);
}

int ensure_buffer(int8_t** buf, size_t* cap, size_t need) {
  if (*cap >= need && *buf) return FGLS_OK;
  HIP_TRY(hipDeviceSynchronize());
  if (*buf) HIP_TRY(hipFree(*buf));
  *buf = nullptr;
  *cap = 0;
  if (hipMalloc((void**)buf, need) != hipSuccess) {
    *buf = nullptr;
    return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", need);
  }
  *cap = need;
  return FGLS_OK;
}

void launch_slices(int ldn, const double* A, int8_t* A8, double* invsc, hipStream_t st) {
  hipLaunchKernelGGL((k_slice_linv<NSLICE>), dim3(ldn), dim3(256), 0, st, ldn, A, A8, invsc);
}

void launch_snp_stage_any(const fgls_hip_ctx* c, const double* Z, int64_t ldz, int count,
                          double* dB, double* dV, int64_t base, hipStream_t st) {
  FGLS_P_SWITCH(c->p, launch_snp_stage, c, Z, ldz, count, dB, dV, base, st);
}

void launch_snp_stage_weighted_any(int p, int n, int count, int wXL, int wXR, const double* Z, int64_t ldz,
                                   const double* W, int64_t ldw, const double* dwt, const double* trait,
                                   double* dB, double* dV, unsigned long long* err, int64_t base,
                                   hipStream_t st) {
  FGLS_P_SWITCH(p, launch_snp_stage_weighted, n, count, wXL, wXR, Z, ldz, W, ldw, dwt, trait, dB, dV, err,
                base, st);
}

int finish_trait(fgls_hip_ctx* c, const double* dXL, int64_t ldxl, const double* dy, hipStream_t st,
                 const int* potrf_info) {
  int rc;
  const int n = c->n, wXL = c->wXL;
  const size_t ldn = (size_t)c->ldn;
  fgls_adjoint_invalidate(c->adj);
  c->have_trait = false;
  if (wXL > 0)
    HIP_TRY(hipMemcpy2DAsync(c->Wtmp, ldn * sizeof(double), dXL, (size_t)ldxl * sizeof(double),
                             (size_t)n * sizeof(double), (size_t)wXL, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->Wtmp + (size_t)wXL * ldn, dy, (size_t)n * sizeof(double),
                         hipMemcpyDeviceToDevice, st));
  // L^-1 (already formed with the device factor)
  if (!potrf_info && (rc = chol_inv(c, false, st))) return rc;
  // inverses of L's 64 x 64 diagonal blocks, for fgls_hip_trsm: the diagonal
  // blocks of L^-1
  hipLaunchKernelGGL(k_diag_copy, dim3(c->nblk), dim3(256), 0, st, n, c->Linv, (int64_t)ldn, c->Dinv);
  // [X~_L | y~] = L^-1 [X_L | y] through the explicit inverse: wXL + 1 columns,
  // one pass over the triangle of L^-1 with every row independent (a TRSM
  // would run the n rows one after another)
  static_assert(PMAX <= TRMV_QMAX, "the skinny product holds p columns");
  hipLaunchKernelGGL(k_trmv_lower, dim3(ldn / TRMV_ROWS), dim3(256), 0, st, n, wXL + 1, c->Linv,
                     (int64_t)ldn, c->Wtmp, (int64_t)ldn, c->W, (int64_t)ldn);
  hipLaunchKernelGGL(k_trait_stats, dim3(1), dim3(256), 0, st, n, wXL, c->W, (int64_t)ldn, c->trait,
                     c->status);
  HIP_TRY(hipGetLastError());
  if (c->mode == FGLS_MODE_MIXED_I8 && (rc = build_slices(c, st))) return rc;
  int status[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(status, c->status, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (potrf_info && status[1] != 0)
    return fail(FGLS_ERR_NOTSPD, "dpotrf(M) failed (info: %d)", status[1]);
  if (status[0] != 0) return fail(FGLS_ERR_NOTSPD, "dpotrf(scoreV) failed (info: %d)", status[0]);
  c->have_trait = true;
  return FGLS_OK;
}

}  // namespace fgls

extern "C" {

const char* fgls_hip_last_error(void) { return g_err; }

int fgls_set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int fgls_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int fgls_hip_ctx_create(fgls_hip_ctx** out, int n, int wXL, int wXR, int max_block, int device) {
  if (!out) return fail(FGLS_ERR_ARG, "NULL output pointer");
  *out = nullptr;
  if (n < 1 || wXL < 0 || wXR < 1 || wXL + wXR > PMAX || max_block < 1)
    return fail(FGLS_ERR_ARG, "invalid sizes n=%d wXL=%d wXR=%d max_block=%d (need n>=1, wXR>=1, p<=%d)",
                n, wXL, wXR, max_block, PMAX);
  if (wXL >= n) return fail(FGLS_ERR_ARG, "wXL=%d must be smaller than n=%d", wXL, n);
  int ndev = fgls_hip_device_count();
  if (device < 0 || device >= ndev)
    return fail(FGLS_ERR_ARG, "device %d not available (%d visible)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  fgls_hip_ctx* c = new fgls_hip_ctx();
  memset(c, 0, sizeof(*c));
  c->device = device;
  c->n = n;
  c->wXL = wXL;
  c->wXR = wXR;
  c->p = wXL + wXR;
  c->ldn = (int)round_up(n, TM);
  c->nblk = (n + NB - 1) / NB;
  c->RT = c->ldn / TM;
  c->max_block = (int)round_up(max_block, 4);
  c->RT8 = c->ldn / I8_TM;
  c->part_rt = c->RT;
  c->mode = FGLS_MODE_FP64;
  c->nslice = FGLS_NSLICE;
  c->slice_form = FGLS_SLICES_AUTO;
  c->slice_loss = -1.0;
  const size_t ldn = (size_t)c->ldn;
  const size_t nq = (size_t)c->p + 1;
  struct { void** ptr; size_t bytes; } allocs[] = {
      {(void**)&c->L, ldn * ldn * sizeof(double)},
      {(void**)&c->Linv, ldn * ldn * sizeof(double)},
      {(void**)&c->Dinv, (size_t)c->nblk * NB * NB * sizeof(double)},
      {(void**)&c->W, ldn * (size_t)(wXL + 1) * sizeof(double)},
      {(void**)&c->Wtmp, ldn * (size_t)(wXL + 1) * sizeof(double)},
      {(void**)&c->trait, (size_t)(2 * PMAX + PMAX * PMAX + 8) * sizeof(double)},
      {(void**)&c->part, (size_t)c->RT * nq * (size_t)c->max_block * sizeof(double)},
      {(void**)&c->err, sizeof(unsigned long long)},
      {(void**)&c->status, 4 * sizeof(int)},
      {(void**)&c->bad, 4 * sizeof(unsigned int)},
  };
  for (auto& a : allocs) {
    if (hipMalloc(a.ptr, a.bytes) != hipSuccess) {
      fgls_hip_ctx_destroy(c);
      return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", a.bytes);
    }
    if (hipMemset(*a.ptr, 0, a.bytes) != hipSuccess) {
      fgls_hip_ctx_destroy(c);
      return fail(FGLS_ERR_HIP, "hipMemset failed");
    }
  }
  HIP_TRY(hipMemset(c->err, 0xff, sizeof(unsigned long long)));
  HIP_TRY(hipDeviceSynchronize());
  *out = c;
  return FGLS_OK;
}

int fgls_hip_ctx_destroy(fgls_hip_ctx* c) {
  if (!c) return FGLS_OK;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  void* ptrs[] = {c->L, c->Linv, c->Dinv, c->W, c->Wtmp, c->trait, c->part, c->z, c->err, c->status,
                  c->A8, c->invsc, c->bad, c->X8t};
  for (void* p : ptrs)
    if (p) hipFree(p);
  fgls_adjoint_free(c->adj);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->side) hipStreamDestroy(c->side);
  delete c;
  return FGLS_OK;
}

int fgls_hip_ctx_ldn(const fgls_hip_ctx* c) { return c ? c->ldn : 0; }

unsigned long long* fgls_ctx_err_ptr(fgls_hip_ctx* c) { return c ? c->err : nullptr; }

int fgls_hip_set_trait(fgls_hip_ctx* c, const double* dL, int64_t ldl, const double* dXL,
                       int64_t ldxl, const double* dy, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!dL || !dy || (c->wXL > 0 && !dXL)) return fail(FGLS_ERR_ARG, "NULL input pointer");
  if (ldl < c->n || (c->wXL > 0 && ldxl < c->n)) return fail(FGLS_ERR_ARG, "leading dimension below n");
  if ((rc = set_device(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const size_t ldn = (size_t)c->ldn;
  HIP_TRY(hipMemcpy2DAsync(c->L, ldn * sizeof(double), dL, (size_t)ldl * sizeof(double),
                           (size_t)c->n * sizeof(double), (size_t)c->n, hipMemcpyDeviceToDevice, st));
  return fgls::finish_trait(c, dXL, ldxl, dy, st, nullptr);
}

int fgls_hip_get_trait(fgls_hip_ctx* c, double* hBt, double* hVtl, double* hres) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if ((rc = set_device(c))) return rc;
  const int wXL = c->wXL;
  double buf[2 * PMAX + PMAX * PMAX + 8];
  HIP_TRY(hipMemcpy(buf, c->trait, sizeof(buf), hipMemcpyDeviceToHost));
  if (hBt) memcpy(hBt, buf, (size_t)wXL * sizeof(double));
  if (hVtl) memcpy(hVtl, buf + wXL, (size_t)wXL * wXL * sizeof(double));
  if (hres) *hres = buf[trait_res_off(wXL)];
  return FGLS_OK;
}

int fgls_hip_trmm(fgls_hip_ctx* c, const double* dX, int64_t ldx, int ncols, double* dZ,
                  int64_t ldz, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (ncols < 0) return fail(FGLS_ERR_ARG, "ncols < 0");
  if (ncols == 0) return FGLS_OK;
  if (!dX || !dZ) return fail(FGLS_ERR_ARG, "NULL pointer");
  if (ldx < c->n || (ldx & 3) || !aligned32(dX))
    return fail(FGLS_ERR_ARG, "X needs ld >= n, ld %% 4 == 0 and 32-byte alignment");
  if (ldz < c->ldn || (ldz & 3) || !aligned32(dZ))
    return fail(FGLS_ERR_ARG, "Z needs ld >= %d, ld %% 4 == 0 and 32-byte alignment", c->ldn);
  if ((rc = set_device(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  dim3 grid((ncols + TN - 1) / TN, c->RT);
  hipLaunchKernelGGL((k_trmm<EPI_STORE, 2>), grid, dim3(256), 0, st, c->n, ncols, c->RT, c->Linv,
                     (int64_t)c->ldn, dX, ldx, dZ, ldz, c->W, (int64_t)c->ldn, (double*)nullptr,
                     (int64_t)0);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_trsm(fgls_hip_ctx* c, double* dX, int64_t ldx, int ncols, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (ncols < 0) return fail(FGLS_ERR_ARG, "ncols < 0");
  if (ncols == 0) return FGLS_OK;
  if (!dX || ldx < c->n) return fail(FGLS_ERR_ARG, "X needs ld >= n");
  if ((rc = set_device(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_trsm_panel, dim3((ncols + NB - 1) / NB), dim3(256), 0, st, c->n, ncols,
                     c->L, (int64_t)c->ldn, c->Dinv, dX, ldx);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_snp_stage(fgls_hip_ctx* c, const double* dZ, int64_t ldz, int count, double* dB,
                       double* dV, int64_t snp_base, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (count < 0) return fail(FGLS_ERR_ARG, "count < 0");
  if (count == 0) return FGLS_OK;
  if (!dZ || !dB || !dV || ldz < c->n) return fail(FGLS_ERR_ARG, "bad Z/B/V arguments");
  if ((rc = set_device(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  FGLS_P_SWITCH(c->p, launch_snp_stage, c, dZ, ldz, count, dB, dV, snp_base, st);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_block_reduce(fgls_hip_ctx* c, const double* dXR, int64_t ldxr, int count,
                          void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (count < 0 || count > c->max_block)
    return fail(FGLS_ERR_ARG, "count=%d outside [0, %d]", count, c->max_block);
  if (count == 0) return FGLS_OK;
  if (!dXR) return fail(FGLS_ERR_ARG, "NULL pointer");
  if (ldxr < c->n || (ldxr & 3) || !aligned32(dXR))
    return fail(FGLS_ERR_ARG, "X_R needs ld >= n, ld %% 4 == 0 and 32-byte alignment");
  if ((rc = set_device(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (c->wXR == 1) {
    FGLS_P_SWITCH(c->p, launch_trmm_reduce, c, dXR, ldxr, count, st);
  } else {
    if ((rc = ensure_z(c))) return rc;
    rc = fgls_hip_trmm(c, dXR, ldxr, count * c->wXR, c->z, c->ldn, stream);
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  c->reduced = count;
  c->last_rt = c->RT;
  return FGLS_OK;
}

int fgls_hip_ctx_set_slice_form(fgls_hip_ctx* c, int form) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (form != FGLS_SLICES_AUTO && form != FGLS_SLICES_SIX) return fail(FGLS_ERR_ARG, "unknown slice form %d", form);
  if (form == c->slice_form) return FGLS_OK;
  c->slice_form = form;
  if (c->mode == FGLS_MODE_MIXED_I8 && c->have_trait) {
    if ((rc = set_device(c))) return rc;
    if ((rc = build_slices(c, nullptr))) return rc;
    HIP_TRY(hipDeviceSynchronize());
  }
  return FGLS_OK;
}

int fgls_hip_ctx_slice_info(const fgls_hip_ctx* c, int* nslice, double* loss) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const bool built = c->A8 && c->have_trait && c->mode == FGLS_MODE_MIXED_I8;
  if (nslice) *nslice = built ? c->nslice : 0;
  if (loss) *loss = built ? c->slice_loss : -1.0;
  return FGLS_OK;
}

int fgls_hip_ctx_set_mode(fgls_hip_ctx* c, int mode) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (mode != FGLS_MODE_FP64 && mode != FGLS_MODE_MIXED_I8) return fail(FGLS_ERR_ARG, "unknown mode %d", mode);
  if (mode == FGLS_MODE_MIXED_I8 && c->n > FGLS_I8_MAX_N)
    return fail(FGLS_ERR_ARG, "the mixed mode is exact for n <= %d (got %d)", FGLS_I8_MAX_N, c->n);
  if ((rc = set_device(c))) return rc;
  if (mode == FGLS_MODE_MIXED_I8) {
    if (!c->A8) {
      const size_t bytes = (size_t)c->ldn * c->ldn * NSLICE;
      if (hipMalloc((void**)&c->A8, bytes) != hipSuccess)
        return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
      // the row scales and, after them, the loss cell of build_slices
      if (hipMalloc((void**)&c->invsc, ((size_t)c->ldn + 1) * sizeof(double)) != hipSuccess)
        return fail(FGLS_ERR_NOMEM, "hipMalloc of the row scales failed");
    }
    if ((rc = fgls::ensure_buffer(&c->X8t, &c->X8t_cap, fgls_x8t_bytes(c->n, c->max_block * c->wXR, 1))))
      return rc;
    if (c->wXR > 1 && (rc = ensure_z(c))) return rc;
    if (c->part_rt < c->RT8) {
      double* np = nullptr;
      const size_t bytes = (size_t)c->RT8 * (c->p + 1) * (size_t)c->max_block * sizeof(double);
      HIP_TRY(hipDeviceSynchronize());
      if (hipMalloc((void**)&np, bytes) != hipSuccess)
        return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
      HIP_TRY(hipFree(c->part));
      c->part = np;
      c->part_rt = c->RT8;
    }
    if (c->have_trait) {
      if ((rc = build_slices(c, nullptr))) return rc;
      HIP_TRY(hipDeviceSynchronize());
    }
  }
  c->mode = mode;
  return FGLS_OK;
}

int fgls_hip_to_i8(const double* dX, int64_t ldx, int n, int ncols, int8_t* dX8, int64_t ldx8,
                   unsigned int* dbad, void* stream) {
  if (n < 1 || ncols < 0 || ldx < n || ldx8 < n || !dX || !dX8 || !dbad)
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_to_i8");
  if (ncols == 0) return FGLS_OK;
  hipLaunchKernelGGL(k_to_i8, dim3(ncols), dim3(256), 0, (hipStream_t)stream, n, dX, ldx, dX8, ldx8,
                     dbad);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_u2_to_i8(const uint8_t* dP, int64_t ldp, int n, int ncols, int8_t* dX8, int64_t ldx8,
                      unsigned int* dbad, void* stream) {
  if (n < 1 || ncols < 0 || !dP || !dX8 || !dbad || ldp < (n + 3) / 4 || ldx8 < n || (ldx8 & 3) ||
      ((uintptr_t)dX8 & 3))
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_u2_to_i8");
  if (ncols == 0) return FGLS_OK;
  hipLaunchKernelGGL(k_u2_to_i8, dim3(ncols), dim3(256), 0, (hipStream_t)stream, n, dP, ldp, dX8, ldx8,
                     dbad);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_i16_to_dosage_i8(const int16_t* dV, int64_t ldv, int n, int ncols, int8_t* dX8,
                              int64_t ldx8, unsigned int* dbad, void* stream) {
  if (n < 1 || ncols < 0 || !dV || !dX8 || !dbad || ldv < n || ldx8 < n)
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_i16_to_dosage_i8");
  if (ncols == 0) return FGLS_OK;
  hipLaunchKernelGGL(k_i16_to_dosage_i8, dim3(ncols), dim3(256), 0, (hipStream_t)stream, n, dV, ldv, dX8,
                     ldx8, (int64_t)ncols * ldx8, dbad);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_i16_to_f64(const int16_t* dV, int64_t ldv, int n, int ncols, double scale, double* dX,
                        int64_t ldx, void* stream) {
  if (n < 1 || ncols < 0 || !dV || !dX || ldv < n || ldx < n || !(scale > 0.0))
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_i16_to_f64");
  if (ncols == 0) return FGLS_OK;
  hipLaunchKernelGGL(k_i16_to_f64, dim3(ncols), dim3(256), 0, (hipStream_t)stream, n, dV, ldv, scale, dX,
                     ldx);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_check_i8(const int8_t* dX8, int64_t ldx8, int n, int ncols, unsigned int* dbad,
                      void* stream) {
  if (n < 1 || ncols < 0 || ldx8 < n || !dX8 || !dbad)
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_check_i8");
  if (ncols == 0) return FGLS_OK;
  hipLaunchKernelGGL(k_check_i8, dim3(ncols), dim3(256), 0, (hipStream_t)stream, n, dX8, ldx8, dbad);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_i8_to_f64(const int8_t* dX8, int64_t ldx8, int n, int ncols, double* dX, int64_t ldx,
                       void* stream) {
  if (n < 1 || ncols < 0 || ldx < n || ldx8 < n || !dX || !dX8)
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_i8_to_f64");
  if (ncols == 0) return FGLS_OK;
  hipLaunchKernelGGL(k_i8_to_f64, dim3(ncols), dim3(256), 0, (hipStream_t)stream, n, dX8, ldx8, dX, ldx);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

}  // extern "C"

namespace {

// The int8 product of a block of tiled genotypes (planes = 1) or dosage
// digits (planes = 2, x = (128 hi + lo) / scale).  wXR = 1: the fused
// reductions into `part`; wXR > 1: the count * wXR tiled columns (SNP i's at
// i*wXR + a) transformed into z for the per-SNP stage.
int reduce_i8t(fgls_hip_ctx* c, const int8_t* dX8T, int count, int planes, double scale,
               void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (c->mode != FGLS_MODE_MIXED_I8) return fail(FGLS_ERR_STATE, "context is not in the mixed mode");
  if (count < 0 || count > c->max_block)
    return fail(FGLS_ERR_ARG, "count=%d outside [0, %d]", count, c->max_block);
  if (planes != 1 && planes != 2) return fail(FGLS_ERR_ARG, "planes must be 1 or 2");
  if (planes == 2 && c->wXR != 1)
    return fail(FGLS_ERR_ARG, "dosage digits need wXR = 1 (got %d)", c->wXR);
  if (count == 0) return FGLS_OK;
  if (!dX8T || ((uintptr_t)dX8T & 15)) return fail(FGLS_ERR_ARG, "tiled genotypes need 16-byte alignment");
  // 32-bit offsets to the hi plane of the tiled dosage digits
  if (planes == 2 && fgls_x8t_plane(c->n, count) > ((int64_t)1 << 31) - 1)
    return fail(FGLS_ERR_ARG, "int8 X_R block of %d columns too large", count);
  if (planes == 2 && !(scale > 0.0)) return fail(FGLS_ERR_ARG, "dosage scale must be positive");
  if ((rc = set_device(c))) return rc;
  const hipStream_t st = (hipStream_t)stream;
  if (c->wXR > 1) {
    if ((rc = ensure_z(c))) return rc;
    fgls_launch_i8_store(c->n, count * c->wXR, c->RT8, c->ldn / 64, c->A8, c->invsc, dX8T, round_up(c->n, 64),
                         c->z, (int64_t)c->ldn, c->nslice, st);
  } else {
    fgls_launch_trmm_i8(c->p, c->n, count, c->RT8, c->ldn / 64, c->A8, c->invsc, dX8T,
                        round_up(c->n, 64), c->W, (int64_t)c->ldn, c->part, (int64_t)c->max_block,
                        planes == 2 ? fgls_x8t_plane(c->n, count) : 0, planes == 2 ? 1.0 / scale : 1.0,
                        0, c->nslice, st);
  }
  HIP_TRY(hipGetLastError());
  c->reduced = count;
  c->last_rt = c->RT8;
  return FGLS_OK;
}

int check_x8(int n, const int8_t* dX8, int64_t ldx8) {
  if (!dX8) return fail(FGLS_ERR_ARG, "NULL pointer");
  if (ldx8 < (int64_t)round_up(n, 64) || (ldx8 & 15) || ((uintptr_t)dX8 & 15))
    return fail(FGLS_ERR_ARG, "int8 X_R needs ld >= ceil(n/64)*64, ld %% 16 == 0 and 16-byte alignment");
  return FGLS_OK;
}

// The int8 product of a block of 2-bit tiled genotypes (fgls_x2t_bytes of
// count * wXR columns), as reduce_i8t.
int reduce_u2t(fgls_hip_ctx* c, const uint8_t* dX2T, int count, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (c->mode != FGLS_MODE_MIXED_I8) return fail(FGLS_ERR_STATE, "context is not in the mixed mode");
  if (count < 0 || count > c->max_block)
    return fail(FGLS_ERR_ARG, "count=%d outside [0, %d]", count, c->max_block);
  if (count == 0) return FGLS_OK;
  if (!dX2T || ((uintptr_t)dX2T & 3)) return fail(FGLS_ERR_ARG, "2-bit tiles need 4-byte alignment");
  if ((rc = set_device(c))) return rc;
  const hipStream_t st = (hipStream_t)stream;
  if (c->wXR > 1) {
    if ((rc = ensure_z(c))) return rc;
    fgls_launch_u2_store(c->n, count * c->wXR, c->RT8, c->ldn / 64, c->A8, c->invsc, dX2T, c->z,
                         (int64_t)c->ldn, c->nslice, st);
  } else {
    fgls_launch_trmm_u2(c->p, c->n, count, c->RT8, c->ldn / 64, c->A8, c->invsc, dX2T, c->W, (int64_t)c->ldn,
                        c->part, (int64_t)c->max_block, 0, c->nslice, st);
  }
  HIP_TRY(hipGetLastError());
  c->reduced = count;
  c->last_rt = c->RT8;
  return FGLS_OK;
}

// Column-major genotypes: tiled into the context's buffer, then the product.
int reduce_i8(fgls_hip_ctx* c, const int8_t* dX8, int64_t ldx8, int count, int planes, double scale,
              void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count < 0 || count > c->max_block)
    return fail(FGLS_ERR_ARG, "count=%d outside [0, %d]", count, c->max_block);
  if (count == 0 || !c->have_trait || c->mode != FGLS_MODE_MIXED_I8 || (planes == 2 && c->wXR != 1))
    return reduce_i8t(c, nullptr, count, planes, scale, stream);   // the state errors
  if ((rc = check_x8(c->n, dX8, ldx8))) return rc;
  if ((rc = set_device(c))) return rc;
  const int ncols = count * c->wXR;
  if ((rc = fgls::ensure_buffer(&c->X8t, &c->X8t_cap, fgls_x8t_bytes(c->n, ncols, planes)))) return rc;
  fgls_launch_tile_i8(c->n, ncols, planes, dX8, ldx8, c->X8t, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return reduce_i8t(c, c->X8t, count, planes, scale, stream);
}

}  // namespace

extern "C" {

int fgls_hip_block_reduce_i8(fgls_hip_ctx* c, const int8_t* dX8, int64_t ldx8, int count,
                             void* stream) {
  return reduce_i8(c, dX8, ldx8, count, 1, 1.0, stream);
}

size_t fgls_hip_x8t_bytes(int n, int count, int planes) {
  if (n < 1 || count < 0 || planes < 1 || planes > 2) return 0;
  return fgls_x8t_bytes(n, count, planes);
}

int fgls_hip_tile_i8(const int8_t* dX8, int64_t ldx8, int n, int count, int planes, int8_t* dX8T,
                     void* stream) {
  if (n < 1 || count < 0 || (planes != 1 && planes != 2) || !dX8T || ((uintptr_t)dX8T & 15))
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_tile_i8");
  int rc = check_x8(n, dX8, ldx8);
  if (rc) return rc;
  if (count == 0) return FGLS_OK;
  fgls_launch_tile_i8(n, count, planes, dX8, ldx8, dX8T, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

size_t fgls_hip_x2t_bytes(int n, int count) {
  if (n < 1 || count < 0) return 0;
  return fgls_x2t_bytes(n, count);
}

int fgls_hip_tile_u2(const int8_t* dX8, int64_t ldx8, int n, int count, uint8_t* dX2T, unsigned int* dbad,
                     void* stream) {
  if (n < 1 || count < 0 || !dX2T || !dbad || ((uintptr_t)dX2T & 3))
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_tile_u2");
  int rc = check_x8(n, dX8, ldx8);
  if (rc) return rc;
  if (count == 0) return FGLS_OK;
  fgls_launch_tile_u2(n, count, dX8, ldx8, dX2T, dbad, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_tile_u2_packed(const uint8_t* dP, int64_t ldp, int n, int count, uint8_t* dX2T,
                            unsigned int* dbad, void* stream) {
  if (n < 1 || count < 0 || !dP || !dX2T || !dbad || ((uintptr_t)dX2T & 3) || ldp < (n + 3) / 4)
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_tile_u2_packed");
  if (count == 0) return FGLS_OK;
  fgls_launch_tile_u2_packed(n, count, dP, ldp, dX2T, dbad, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_block_reduce_u2t(fgls_hip_ctx* c, const uint8_t* dX2T, int count, void* stream) {
  return reduce_u2t(c, dX2T, count, stream);
}

int fgls_hip_solve_block_u2t(fgls_hip_ctx* c, const uint8_t* dX2T, int count, double* dB, double* dV,
                             int64_t snp_base, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count > 0 && (!dB || !dV)) return fail(FGLS_ERR_ARG, "NULL pointer");
  rc = reduce_u2t(c, dX2T, count, stream);
  if (rc) return rc;
  return fgls_hip_block_finalize(c, count, dB, dV, snp_base, stream);
}

int fgls_hip_block_reduce_i8t(fgls_hip_ctx* c, const int8_t* dX8T, int count, int planes,
                              double scale, void* stream) {
  return reduce_i8t(c, dX8T, count, planes, scale, stream);
}

int fgls_hip_solve_block_i8t(fgls_hip_ctx* c, const int8_t* dX8T, int count, int planes, double scale,
                             double* dB, double* dV, int64_t snp_base, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count > 0 && (!dB || !dV)) return fail(FGLS_ERR_ARG, "NULL pointer");
  rc = reduce_i8t(c, dX8T, count, planes, scale, stream);
  if (rc) return rc;
  return fgls_hip_block_finalize(c, count, dB, dV, snp_base, stream);
}

int fgls_hip_solve_block_i8(fgls_hip_ctx* c, const int8_t* dX8, int64_t ldx8, int count, double* dB,
                            double* dV, int64_t snp_base, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count > 0 && (!dB || !dV)) return fail(FGLS_ERR_ARG, "NULL pointer");
  rc = fgls_hip_block_reduce_i8(c, dX8, ldx8, count, stream);
  if (rc) return rc;
  return fgls_hip_block_finalize(c, count, dB, dV, snp_base, stream);
}

int fgls_hip_block_reduce_dosage(fgls_hip_ctx* c, const int8_t* dX8, int64_t ldx8, int count,
                                 double scale, void* stream) {
  return reduce_i8(c, dX8, ldx8, count, 2, scale, stream);
}

int fgls_hip_solve_block_dosage(fgls_hip_ctx* c, const int8_t* dX8, int64_t ldx8, int count,
                                double scale, double* dB, double* dV, int64_t snp_base, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count > 0 && (!dB || !dV)) return fail(FGLS_ERR_ARG, "NULL pointer");
  rc = fgls_hip_block_reduce_dosage(c, dX8, ldx8, count, scale, stream);
  if (rc) return rc;
  return fgls_hip_block_finalize(c, count, dB, dV, snp_base, stream);
}

int fgls_hip_to_dosage_i8(const double* dX, int64_t ldx, int n, int ncols, double scale,
                          int8_t* dX8, int64_t ldx8, unsigned int* dbad, void* stream) {
  if (n < 1 || ncols < 0 || ldx < n || ldx8 < n || !dX || !dX8 || !dbad || !(scale > 0.0))
    return fail(FGLS_ERR_ARG, "bad arguments to fgls_hip_to_dosage_i8");
  if (ncols == 0) return FGLS_OK;
  hipLaunchKernelGGL(k_to_dosage_i8, dim3(ncols), dim3(256), 0, (hipStream_t)stream, n, dX, ldx,
                     scale, dX8, ldx8, (int64_t)ncols * ldx8, dbad);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_block_finalize(fgls_hip_ctx* c, int count, double* dB, double* dV, int64_t snp_base,
                            void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count < 0 || count != c->reduced)
    return fail(FGLS_ERR_STATE, "finalize of %d SNPs after a reduce of %d", count, c->reduced);
  if (count == 0) return FGLS_OK;
  if (!dB || !dV) return fail(FGLS_ERR_ARG, "NULL pointer");
  if ((rc = set_device(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (c->wXR == 1) {
    FGLS_P_SWITCH(c->p, launch_finalize, c, count, dB, dV, snp_base, st);
  } else {
    FGLS_P_SWITCH(c->p, launch_snp_stage, c, c->z, (int64_t)c->ldn, count, dB, dV, snp_base, st);
  }
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_solve_block(fgls_hip_ctx* c, const double* dXR, int64_t ldxr, int count, double* dB,
                         double* dV, int64_t snp_base, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count > 0 && (!dB || !dV)) return fail(FGLS_ERR_ARG, "NULL pointer");
  rc = fgls_hip_block_reduce(c, dXR, ldxr, count, stream);
  if (rc) return rc;
  return fgls_hip_block_finalize(c, count, dB, dV, snp_base, stream);
}

int fgls_hip_check(fgls_hip_ctx* c, void* stream, int64_t* bad_snp, int* info) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = set_device(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  unsigned long long key = 0;
  HIP_TRY(hipMemcpyAsync(&key, c->err, sizeof(key), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (key == ~0ull) return FGLS_OK;
  const int64_t snp = (int64_t)(key >> 8);
  const int inf = (int)(key & 0xff);
  if (bad_snp) *bad_snp = snp;
  if (info) *info = inf;
  HIP_TRY(hipMemsetAsync(c->err, 0xff, sizeof(unsigned long long), st));
  HIP_TRY(hipStreamSynchronize(st));
  return fail(FGLS_ERR_NOTSPD, "dpotrf(V) failed (info: %d) - i: %lld", inf, (long long)snp);
}

// ---------------------------------------------------------------------------
// fgls_hip_solve_block_noalloc: fgls_hip_solve_block with its scratch in a
// file-scope device array instead of context buffers, for callers that cannot
// allocate.  The scratch holds the per-row-tile partial sums of the fused
// product (wXR = 1) or the transformed columns L^-1 X_R (wXR > 1); every
// entry a call reads was written earlier in the same call, so nothing is
// cleared.  Each device holds its own copy for the life of the process;
// calls on one device share it, so calls that overlap in time on one device
// (other streams, other threads) race.

}  // extern "C"

namespace {

constexpr size_t kScratchBytes = FGLS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices];   // device address of `scratch` per device

// The device address of this device's copy of `scratch` (the symbol's name
// is a host-side stand-in); looked up once per device with that device
// current.
int scratch_ptr(int device, double** out) {
  if (device < 0 || device >= kMaxScratchDevices)
    return fail(FGLS_ERR_ARG, "device %d beyond the %d devices the scratch table holds", device,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FGLS_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = (double*)p;
  }
  *out = g_scratch_ptr[device];
  return FGLS_OK;
}

// SNPs per call that fit the scratch: (p + 1) partial sums per 128-row tile
// and SNP for wXR = 1, the wXR transformed columns (ldn rows) otherwise.
int noalloc_max_count(const fgls_hip_ctx* c) {
  const size_t per = c->wXR == 1 ? (size_t)c->RT * (c->p + 1) * sizeof(double)
                                 : (size_t)c->ldn * c->wXR * sizeof(double);
  const size_t cap = kScratchBytes / per;
  return (int)(cap < (size_t)INT32_MAX ? cap : (size_t)INT32_MAX);
}

}  // namespace

extern "C" {

int fgls_hip_solve_block_noalloc(fgls_hip_ctx* c, const double* dXR, int64_t ldxr, int count, double* dB,
                                 double* dV, int64_t snp_base, void* stream) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  const int cap = noalloc_max_count(c);
  if (count < 0 || count > cap)
    return fail(FGLS_ERR_ARG,
                "count=%d outside [0, %d]: the no-alloc solve holds its scratch in %zu bytes "
                "(at n=%d, p=%d, wXR=%d at most %d SNPs per call)",
                count, cap, kScratchBytes, c->n, c->p, c->wXR, cap);
  if (count == 0) return FGLS_OK;
  if (!dXR || !dB || !dV) return fail(FGLS_ERR_ARG, "NULL pointer");
  if (ldxr < c->n || (ldxr & 3) || !aligned32(dXR))
    return fail(FGLS_ERR_ARG, "X_R needs ld >= n, ld %% 4 == 0 and 32-byte alignment");
  if ((rc = set_device(c))) return rc;
  double* ws = nullptr;
  if ((rc = scratch_ptr(c->device, &ws))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (c->wXR == 1) {
    const int64_t ldp = count;
    FGLS_P_SWITCH(c->p, launch_trmm_reduce_into, c, dXR, ldxr, count, ws, ldp, st);
    FGLS_P_SWITCH(c->p, launch_finalize_from, c, c->RT, ws, ldp, count, dB, dV, snp_base, st);
  } else {
    dim3 grid((count * c->wXR + TN - 1) / TN, c->RT);
    hipLaunchKernelGGL((k_trmm<EPI_STORE, 2>), grid, dim3(256), 0, st, c->n, count * c->wXR, c->RT, c->Linv,
                       (int64_t)c->ldn, dXR, ldxr, ws, (int64_t)c->ldn, c->W, (int64_t)c->ldn,
                       (double*)nullptr, (int64_t)0);
    FGLS_P_SWITCH(c->p, launch_snp_stage, c, ws, (int64_t)c->ldn, count, dB, dV, snp_base, st);
  }
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

}  // extern "C"
