// fgls_potrf.hip -- the per-trait Cholesky factor and its inverse on the
// device:
//   M = sigma2 (h2 Phi + (1 - h2) I) = L L',   T = L^-1      (FP64, lower)
// replacing the host M construction and dpotrf of fgls_chol_gpu.c:406-430 and
// the triangular solves with L that the reference repeats for every block
// (fgls_chol_gpu.c:493-499, 536-542), which this library replaces by
// products with T.
//
// Recursive factorisation and inversion over the padded order N = ldn (a
// multiple of 128), split at multiples of 64 (top-down, Gustavson's
// recursion with the inverse carried along):
//
//   [S11  .  ]   [L11   0 ] [L11'  L21']        T11 = L11^-1
//   [S21 S22 ] = [L21  L22] [ 0    L22'],       T22 = L22^-1
//
//   node(S):  node(S11)                  -> L11, T11
//             L21  = S21 T11'                     (product, B upper)
//             S22 -= L21 L21'                     (product, lower half)
//             node(S22)                  -> L22, T22
//             X    = L21 T11                      (product, B lower)
//             T21  = -T22 X     (as T21' = -X' T22', product, B upper)
//
// 64 x 64 leaves factor and invert in one workgroup (k_leaf), 128-row nodes
// in one workgroup (k_node128).  Every product runs on k_tgemm, an FP64 MFMA
// kernel that reads each operand along k or, transposed, across k, with a
// column-major result, or, when it has at most 4,096 64 x 64 blocks of
// output, on k_tgemm_sk, which splits K over the four waves of a block.  So
// each factor needs one layout ("R-view" = the buffer read row-major, element
// (i, j) at i ld + j):
//   L buffer: L column-major in the R-view upper triangle (the final factor);
//             the R-view lower triangle holds M and then the Schur
//             complements S.
//   T buffer: R-view lower = T row-major (the final L^-1); the R-view upper
//             holds temporaries and is zeroed at the end.
// The temporary X of a node lives in the T buffer's R-view upper-right block
// of that node.
// Rows and columns n..N-1 are identity in the leaves' arithmetic and zero in
// the stored factor and inverse.  The first non-positive pivot is recorded as
// LAPACK's info (1-based order of the failing leading minor); later leaves
// still run on harmless values.
//
// Flops: n^3/3 for the factor and n^3/3 for the inverse; 3/4 of them are in
// the top level's four products, which fill the device.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fgls_ctx.h"
#include "fgls_hip.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

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

__device__ __forceinline__ v4d ld4(const double* p) {
  return *reinterpret_cast<const v4d*>(p);
}

// M = alpha Phi + beta I on the n x n block, zero elsewhere
// (fgls_chol_gpu.c:406-414: alpha = sigma2 h2, beta = sigma2 (1-h2)), from the
// lower triangle of Phi only, written to the lower triangle of the ldn x ldn
// buffer read row-major, which is what the recursion reads (it overwrites
// the other triangle with the column-major factor).  One workgroup per
// 64 x 64 block (I, J) with I >= J: the block is read once (coalesced down
// the columns of Phi) and written transposed through LDS.
__global__ __launch_bounds__(256) void k_build_M(int n, int ldn, const double* __restrict__ Phi,
                                                 int64_t ldphi, double alpha, double beta,
                                                 double* __restrict__ M) {
  __shared__ double t[64][65];
  const int I = blockIdx.x, J = blockIdx.y;
  if (J > I) return;
  const int r0 = I * 64, c0 = J * 64;
  const bool diag = I == J;
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    const int cc = idx >> 6, rr = idx & 63;
    const int r = r0 + rr, c = c0 + cc;
    double v = 0.0;
    if (r < n && c < n && (!diag || rr >= cc))
      v = alpha * Phi[(size_t)c * ldphi + r] + (r == c ? beta : 0.0);
    t[rr][cc] = v;
  }
  __syncthreads();
  // M(r, c) at M[r ldn + c], r >= c
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    const int rr = idx >> 6, cc = idx & 63;
    if (!diag || rr >= cc) M[(size_t)(r0 + rr) * ldn + c0 + cc] = t[rr][cc];
  }
}

// Value of x in lane l (l wave-uniform), as a scalar broadcast.
__device__ __forceinline__ double lane_bcast(double x, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}

// acc + C for one 16 x 16 block by one wave on FP64 MFMA, C(r, c) =
// sum_{q < 16} U(r, q) V(c, q) with U(r, q) = U[ur + r][uc + q] and V(c, q) =
// V[vr + c][vc + q] (VT: V[vr + q][vc + c]), both in LDS.  Element v of the
// result is (row li, column lg + 4 v), li = lane % 16, lg = lane / 16.
template <bool VT, typename TU, typename TV>
__device__ __forceinline__ v4d blk16(const TU& U, int ur, int uc, const TV& V, int vr, int vc, int li, int lg,
                                     v4d acc) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int q = 4 * lg + s;
    const double a = U[ur + li][uc + q];
    const double b = VT ? V[vr + q][vc + li] : V[vr + li][vc + q];
    acc = mfma_f64(b, a, acc);
  }
  return acc;
}

// ---------------------------------------------------------------------------
// 64 x 64 diagonal block at o = o0 + 64 blockIdx.x, one workgroup of 4 waves,
// blocked by 16 with the block in LDS (A: S, then L; Tm: T = L^-1, lower
// triangles):
// FACTOR: S from the L buffer's R-view lower triangle; for each 16-column
//         panel jb: wave 0 factors the 16 x 16 diagonal block in registers
//         (lane r holds row r; the column and the pivot are scalar
//         broadcasts, v_readlane; 1 / sqrt of the pivot from v_rsq_f64 and
//         two Newton steps), the rows below it in the other lanes of the
//         wave, so that the same 16 pivot steps form the whole column panel;
//         then the waves update the trailing lower triangle,
//         S_ik -= L_ij L_kj', a 16 x 16 block per wave on FP64 MFMA.
// else:   L from the caller's factor (column-major lower triangle).
// Each wave then inverts one 16 x 16 diagonal block (lane c: column c of
// T_jj), off the chain of the factorisation.
// Then the off-diagonal blocks of T by block rows, T_IJ = -T_II sum_{J<=K<I}
// L_IK T_KJ (wave J, FP64 MFMA).  L (FACTOR) is written column-major and T
// row-major, every store instruction covering a contiguous row segment.
// Rows >= n are identity rows; their stored values are zero.  The first
// non-positive pivot is recorded in info (1-based), and the pivot is replaced
// by 1 so that the later arithmetic stays finite.
// ---------------------------------------------------------------------------
// The LDS of one 64 x 64 leaf: A holds S, then L; Tm holds T = L^-1 (lower
// triangles; T's strict upper triangle is zero); X the block sums of T.
struct LeafLds {
  double A[64][65];
  double Tm[64][65];
  double X[3][16][17];
  double dg[64];   // 1 / L(k, k)
};

// Leaf body for the block at o, run by all 256 threads of the workgroup.
// FROM_LDS: S is already in A (rows >= n as identity rows), written by the
// caller before a barrier; otherwise it is read from the buffer.
template <bool FACTOR, bool FROM_LDS>
__device__ __forceinline__ void leaf64(int n, int o, double* __restrict__ Lb, double* __restrict__ Tb, int64_t ld,
                       int* __restrict__ info, LeafLds& sh) {
  auto& A = sh.A;
  auto& Tm = sh.Tm;
  auto& X = sh.X;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int nl = min(64, n - o);   // live rows of the block (may be <= 0)
  // all 16 loads of a thread in flight at once (every address is inside the
  // buffer; the values outside the live lower triangle are replaced)
  double sv[16];
  if (!FROM_LDS) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      // FACTOR: row-major source, c fastest; else column-major, r fastest
      const int idx = tid + 256 * e;
      const int r = FACTOR ? idx >> 6 : idx & 63, c = FACTOR ? idx & 63 : idx >> 6;
      sv[e] = FACTOR ? Lb[(size_t)(o + r) * ld + o + c] : Lb[(size_t)(o + c) * ld + o + r];
    }
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int idx = tid + 256 * e;
    const int r = FACTOR ? idx >> 6 : idx & 63, c = FACTOR ? idx & 63 : idx >> 6;
    if (!FROM_LDS) A[r][c] = (r < nl && c <= r) ? sv[e] : ((r == c) ? 1.0 : 0.0);
    Tm[r][c] = 0.0;   // the strict upper triangle of T stays zero: the block
                      // products below run over whole 16-blocks
  }
  __syncthreads();

  // T_jj = L_jj^-1 of the diagonal block at c0 by one wave: lane c (< 16)
  // forms column c by forward substitution; a[] is row `lane` of L_jj and
  // irr[r] = 1 / L(r, r) (the same in every lane)
  auto diag_inverse = [&](int c0, const double (&a)[16], const double (&irr)[16]) {
    double t[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < r; ++q) s4[q & 3] = fma(lane_bcast(a[q], r), t[q], s4[q & 3]);
      const double sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
      t[r] = (r == lane) ? irr[r] : ((r > lane) ? -sum * irr[r] : 0.0);
    }
    if (lane < 16) {   // whole columns: t[r] = 0 for r < lane
#pragma unroll
      for (int r = 0; r < 16; ++r) Tm[c0 + r][c0 + lane] = t[r];
    }
  };

  int bad = 0;
  if (FACTOR) {
    for (int jb = 0; jb < 4; ++jb) {
      const int c0 = 16 * jb;
      if (w == 0) {
        // the 16-column panel from row c0 down, lane i holding row c0 + i:
        // the diagonal block's rows and the 48 - c0 rows below it, all
        // factored by the same pivot steps (the lanes past the panel repeat
        // its last row)
        double a[16];
        const int rr = min(lane, 63 - c0);
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = A[c0 + rr][c0 + j];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          double d = lane_bcast(a[k], k);   // S(k, k) after k updates
          if (!(d > 0.0)) {
            if (bad == 0) bad = o + c0 + k + 1;
            d = 1.0;
          }
          // 1 / sqrt(d) from the hardware estimate (v_rsq_f64) and two
          // Newton steps y += y (1 - d y^2) / 2, each of which squares the
          // relative error: 7 dependent operations on the pivot chain in
          // place of a square root and a division
          double y = __builtin_amdgcn_rsq(d);
#pragma unroll
          for (int it = 0; it < 2; ++it) y = fma(0.5 * y, fma(-d * y, y, 1.0), y);
          if (lane == 0) sh.dg[c0 + k] = y;
          const double lk = (lane == k) ? d * y : a[k] * y;   // L(c0 + lane, c0 + k)
          a[k] = lk;
#pragma unroll
          for (int j = k + 1; j < 16; ++j) a[j] = fma(-lk, lane_bcast(lk, j), a[j]);
        }
        // whole rows (unmasked stores): the strict upper triangle of a
        // diagonal block of A is never read
        if (lane < 64 - c0) {
#pragma unroll
          for (int j = 0; j < 16; ++j) A[c0 + lane][c0 + j] = a[j];
        }
      }
      __syncthreads();
      // trailing lower triangle: S_IJ -= L_Ij L_Jj' for jb < J <= I < 4, the
      // blocks of the lower triangle dealt round the waves (only the lower
      // half of a diagonal block is written)
      {
        const int m = 3 - jb;
        for (int t = w; t < m * (m + 1) / 2; t += 4) {
          int I = 0;
          while ((I + 1) * (I + 2) / 2 <= t) ++I;
          const int J = t - I * (I + 1) / 2 + jb + 1;
          I += jb + 1;
          const v4d u = blk16<false>(A, 16 * I, c0, A, 16 * J, c0, li, lg, v4d{0.0, 0.0, 0.0, 0.0});
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = 16 * I + li, c = 16 * J + lg + 4 * v;
            if (I != J || c <= r) A[r][c] -= u[v];
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0 && bad != 0 && *info == 0) *info = bad;
  }
  // the diagonal blocks of T, off the factorisation's chain: wave w inverts
  // block w (FACTOR: 1 / L(r, r) kept from the pivots)
  {
    double a[16], irr[16];
    const int c0 = 16 * w, rr = min(lane, 15);
#pragma unroll
    for (int j = 0; j < 16; ++j) a[j] = A[c0 + rr][c0 + j];
#pragma unroll
    for (int r = 0; r < 16; ++r) irr[r] = FACTOR ? sh.dg[c0 + r] : 1.0 / lane_bcast(a[r], r);
    diag_inverse(c0, a, irr);
    __syncthreads();
  }

  // off-diagonal blocks of T, block row I at a time, wave J < I forming
  // T_IJ on FP64 MFMA: X_J = sum_{J <= K < I} L_IK T_KJ, then T_IJ = -T_II X_J
  for (int I = 1; I < 4; ++I) {
    if (w < I) {
      v4d x = {0.0, 0.0, 0.0, 0.0};
      for (int K = w; K < I; ++K) x = blk16<true>(A, 16 * I, 16 * K, Tm, 16 * K, 16 * w, li, lg, x);
#pragma unroll
      for (int v = 0; v < 4; ++v) X[w][li][lg + 4 * v] = x[v];
    }
    __syncthreads();
    if (w < I) {
      const v4d t = blk16<true>(Tm, 16 * I, 16 * I, X[w], 0, 0, li, lg, v4d{0.0, 0.0, 0.0, 0.0});
#pragma unroll
      for (int v = 0; v < 4; ++v) Tm[16 * I + li][16 * w + lg + 4 * v] = -t[v];
    }
    __syncthreads();
  }

  // buffer row o + a, column o + b: L column-major (b >= a holds L(b, a)),
  // live rows only; T row-major (b <= a holds T(a, b)), zero for rows >= n
  for (int idx = tid; idx < 4096; idx += 256) {
    const int ra = idx >> 6, cb = idx & 63;
    if (FACTOR && cb >= ra && cb < nl) Lb[(size_t)(o + ra) * ld + o + cb] = A[cb][ra];
    if (cb <= ra) Tb[(size_t)(o + ra) * ld + o + cb] = (ra < nl) ? Tm[ra][cb] : 0.0;
  }
}

template <bool FACTOR>
__global__ __launch_bounds__(256) void k_leaf(int n, int o0, double* __restrict__ Lb,
                                              double* __restrict__ Tb, int64_t ld,
                                              int* __restrict__ info) {
  __shared__ LeafLds sh;
  leaf64<FACTOR, false>(n, o0 + 64 * (int)blockIdx.x, Lb, Tb, ld, info, sh);
}

// ---------------------------------------------------------------------------
// A whole 128-row node of the factorising recursion (rows [o, o + 128)) in
// one workgroup, in place of 2 leaves, 4 products and 2 transposed copies:
//   leaf(S11) -> L11, T11;  L21 = S21 T11';  S22 - L21 L21' and X = L21 T11
//   (into LDS);  leaf(S22) -> L22, T22;  T21 = -T22 X.
// Everything the node produces for itself stays in LDS (L21, X, the updated
// S22, T11 and T22); only S11 and S21, written by earlier launches, are read
// from the buffers.  The node writes L column-major (its two diagonal blocks
// and L21) and T row-major (its two diagonal blocks and T21).  The
// 64 x 64 x 64 products run on FP64 MFMA with both operands in LDS (a
// quadrant per wave), T's rows >= n zeroed first so that they match the
// stored factor.
// ---------------------------------------------------------------------------
struct NodeLds {
  LeafLds leaf;
  double P[64][65];   // S21, then L21
  double Q[64][65];   // X = L21 T11
};

// C(r, c) = sum_k U(r, k) V(c, k) (V_T: V(k, c) instead), k = 0..63, on FP64
// MFMA: wave w takes the 32 x 32 quadrant at rows 32 (w >> 1), columns
// 32 (w & 1); element v of acc[ti][tj] is (row r0 + 16 ti + li, column
// c0 + 16 tj + lg + 4 v), the operand order of k_tgemm.
template <bool V_T>
__device__ __forceinline__ void lds_product(const double (&U)[64][65], const double (&V)[64][65], int w,
                                            int lane, v4d (&acc)[2][2]) {
  const int li = lane & 15, lg = lane >> 4;
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
  for (int k0 = 0; k0 < 64; k0 += 16) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = k0 + 4 * lg + s;
      double a[2], b[2];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) a[ti] = U[r0 + 16 * ti + li][k];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) b[tj] = V_T ? V[k][c0 + 16 * tj + li] : V[c0 + 16 * tj + li][k];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = mfma_f64(b[tj], a[ti], acc[ti][tj]);
    }
  }
}

// f(r, c, value) for every element of the wave's quadrant
template <typename F>
__device__ __forceinline__ void for_quadrant(const v4d (&acc)[2][2], int w, int lane, F f) {
  const int li = lane & 15, lg = lane >> 4;
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) f(r0 + 16 * ti + li, c0 + 16 * tj + lg + 4 * v, acc[ti][tj][v]);
}

__global__ __launch_bounds__(256) void k_node128(int n, int o, double* __restrict__ Lb,
                                                 double* __restrict__ Tb, int64_t ld,
                                                 int* __restrict__ info) {
  __shared__ NodeLds sh;
  auto& Tm = sh.leaf.Tm;
  auto& A = sh.leaf.A;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int o2 = o + 64;
  const int nl1 = min(64, n - o), nl2 = min(64, n - o2);

  // S21 (R-view lower-left, row-major) and this wave's quadrant of S22 are
  // loaded before the first leaf, so that their latency overlaps the leaf's
  // own load (nothing in this node writes them before they are used)
  double s21[16], s22[2][2][4];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int idx = tid + 256 * e, r = idx >> 6, c = idx & 63;
    s21[e] = Lb[(size_t)(o2 + r) * ld + o + c];
  }
  {
    const int li = lane & 15, lg = lane >> 4;
    const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + 16 * ti + li, c = c0 + 16 * tj + lg + 4 * v;
          s22[ti][tj][v] = Lb[(size_t)(o2 + r) * ld + o2 + c];   // in the buffer; used where c <= r < nl2
        }
  }

  leaf64<true, false>(n, o, Lb, Tb, ld, info, sh.leaf);
  __syncthreads();   // the leaf's last reads of Tm are done
  // T11 as stored: rows >= n zero; S21 into P
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int idx = tid + 256 * e, r = idx >> 6, c = idx & 63;
    if (r >= nl1) Tm[r][c] = 0.0;
    sh.P[r][c] = s21[e];
  }
  __syncthreads();
  // L21(r, c) = sum_{k <= c} S21(r, k) T11(c, k)
  v4d acc[2][2];
  lds_product<false>(sh.P, Tm, w, lane, acc);
  __syncthreads();   // every wave is done reading S21
  for_quadrant(acc, w, lane, [&](int r, int c, double v) {
    sh.P[r][c] = v;
    Lb[(size_t)(o + c) * ld + o2 + r] = v;   // column-major (R-view upper-right)
  });
  __syncthreads();
  // X(r, c) = sum_{k >= c} L21(r, k) T11(k, c) into Q
  lds_product<true>(sh.P, Tm, w, lane, acc);
  for_quadrant(acc, w, lane, [&](int r, int c, double v) { sh.Q[r][c] = v; });
  // S22(r, c) - sum_k L21(r, k) L21(c, k): the second leaf's input in A (its
  // lower triangle; rows >= n identity rows, as the leaf's own load)
  lds_product<false>(sh.P, sh.P, w, lane, acc);
  {
    const int li = lane & 15, lg = lane >> 4;
    const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + 16 * ti + li, c = c0 + 16 * tj + lg + 4 * v;
          double a = (r == c) ? 1.0 : 0.0;
          if (r < nl2 && c <= r) a = s22[ti][tj][v] - acc[ti][tj][v];
          A[r][c] = a;
        }
  }
  __syncthreads();   // A complete; the first leaf's T11 in Tm is no longer needed
  leaf64<true, true>(n, o2, Lb, Tb, ld, info, sh.leaf);
  __syncthreads();
  for (int idx = tid; idx < 4096; idx += 256) {
    const int r = idx >> 6, c = idx & 63;
    if (r >= nl2) Tm[r][c] = 0.0;   // T22 as stored
  }
  __syncthreads();
  // T21(c, r) = -sum_{k <= c} T22(c, k) X(k, r)
  lds_product<true>(Tm, sh.Q, w, lane, acc);
  for_quadrant(acc, w, lane, [&](int c, int r, double v) {
    Tb[(size_t)(o2 + c) * ld + o + r] = -v;   // row-major (R-view lower-left)
  });
}

// ---------------------------------------------------------------------------
// C = alpha A B (+ C with ACC) on an M x N block, K terms, FP64 MFMA, over
// 128 x 128 tiles (2 x 2 waves of 64 x 64).  C(r, c) = C[c ld + r]; M, N, K
// multiples of 64.  The operands are read along k, A(r, k) = A[r ld + k] and
// B(k, c) = B[c ld + k], or across k (transposed storage: ATR, BTR):
// A(r, k) = A[k ld + r], B(k, c) = B[k ld + c] (16 lanes read 16 consecutive
// rows or columns of one k).
// TRI selects the terms and tiles that can be non-zero:
//   TG_BLOWER  B lower triangular (k >= c): k from the wave's first column;
//   TG_BUPPER  B upper triangular (k <= c): k up to the wave's last column;
//   TG_CUPPER  only the blocks with r <= c (one triangle of a symmetric C).
// The triangular B is a diagonal block of a buffer whose other triangle is
// not zero: inside the wave's 64 x 64 diagonal block the terms of that
// triangle are masked.
// Grid: (row tiles, column tiles), heaviest column tiles first.  Operand
// fetch and MFMA order as the triangular product of the block loop: per K = 16
// step one 32-byte load per lane and operand (transposed: four 8-byte
// loads), one step ahead.
// ---------------------------------------------------------------------------
enum { TG_BLOWER = 1, TG_BUPPER = 2, TG_CUPPER = 3 };
#ifndef FGLS_TG_SK_BLOCKS
#define FGLS_TG_SK_BLOCKS 4096
#endif
#ifndef FGLS_TG_COPY_BLOCKS
#define FGLS_TG_COPY_BLOCKS 1024
#endif

// The operands of one wave: rows r0 + 16 t + li of A and columns c0 + 16 t +
// li of B, lane (li, lg) holding k = k0 + 4 lg + s in element s.
template <bool ATR, bool BTR>
struct TgOperands {
  const double* Ap[4];
  const double* Bp[4];
  int64_t ld;
  __device__ __forceinline__ TgOperands(const double* A, const double* B, int64_t ld_, int r0, int c0, int li,
                                        int lg)
      : ld(ld_) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      Ap[t] = ATR ? A + (size_t)(4 * lg) * ld + r0 + 16 * t + li : A + (size_t)(r0 + 16 * t + li) * ld + 4 * lg;
      Bp[t] = BTR ? B + (size_t)(4 * lg) * ld + c0 + 16 * t + li : B + (size_t)(c0 + 16 * t + li) * ld + 4 * lg;
    }
  }
  template <bool TR>
  __device__ __forceinline__ v4d get(const double* p, int k0) const {
    if constexpr (TR) {
      const double* q = p + (size_t)k0 * ld;
      return v4d{q[0], q[ld], q[2 * ld], q[3 * ld]};
    } else {
      return ld4(p + k0);
    }
  }
  __device__ __forceinline__ void load(int k0, v4d (&a)[4], v4d (&b)[4]) const {
#pragma unroll
    for (int t = 0; t < 4; ++t) a[t] = get<ATR>(Ap[t], k0);
#pragma unroll
    for (int t = 0; t < 4; ++t) b[t] = get<BTR>(Bp[t], k0);
  }
};

// A wave's 64 x 64 block of alpha acc (+ C): element v of acc[ti][tj] is (row
// r0 + 16 ti + li, column c0 + 16 tj + lg + 4 v), so 16 lanes write 16
// consecutive rows of one column.
template <bool ACC>
__device__ __forceinline__ void tg_store(const v4d (&acc)[4][4], double alpha, double* __restrict__ C, int64_t ld,
                                         int r0, int c0, int li, int lg) {
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double* Cc = C + (size_t)(c0 + 16 * tj + lg + 4 * v) * ld + r0 + li;
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) {
        const double z = alpha * acc[ti][tj][v];
        Cc[16 * ti] = ACC ? Cc[16 * ti] + z : z;
      }
    }
  }
}

template <int TRI, bool ACC, bool ATR, bool BTR>
__global__ __launch_bounds__(256) void k_tgemm(int M, int N, int K, double alpha,
                                               const double* __restrict__ A,
                                               const double* __restrict__ B,
                                               double* __restrict__ C, int64_t ld) {
  const int CT = (N + 127) / 128;
  const int rt = blockIdx.x;
  const int ct = (TRI == TG_BUPPER) ? CT - 1 - (int)blockIdx.y : (int)blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w >> 1, wc = w & 1;
  const int li = lane & 15, lg = lane >> 4;
  const int r0 = rt * 128 + wr * 64;
  const int c0 = ct * 128 + wc * 64;
  if (r0 >= M || c0 >= N) return;
  if (TRI == TG_CUPPER && r0 > c0) return;
  // unmasked range [kb, ke); a triangular B adds the diagonal block [c0, c0 + 64)
  const int kb = (TRI == TG_BLOWER) ? c0 + 64 : 0;
  const int ke = (TRI == TG_BUPPER) ? c0 : K;
  const TgOperands<ATR, BTR> op(A, B, ld, r0, c0, li, lg);

  v4d acc[4][4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};

  v4d a[4], b[4];
  auto mfma_step = [&]() {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = mfma_f64(b[tj][s], a[ti][s], acc[ti][tj]);
  };
  // the 4 K-steps of the diagonal block, B's other triangle masked
  auto diag_steps = [&]() {
#pragma unroll
    for (int k0 = c0; k0 < c0 + 64; k0 += 16) {
      op.load(k0, a, b);
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        const int c = c0 + 16 * tj + li;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int k = k0 + 4 * lg + s;
          const bool keep = (TRI == TG_BUPPER) ? (k <= c) : (k >= c);
          b[tj][s] = keep ? b[tj][s] : 0.0;
        }
      }
      mfma_step();
    }
  };

  // The diagonal block first; a B-upper product then runs k downwards from
  // it (7 % faster at n = 30,000 than upwards from k = 0), a B-lower one
  // upwards; the dense update runs upwards (downwards or from a
  // tile-dependent start measured no gain).
  if (TRI == TG_BLOWER || TRI == TG_BUPPER) diag_steps();
  if (kb < ke) {
    constexpr int DK = (TRI == TG_BUPPER) ? -16 : 16;
    const int kfirst = (TRI == TG_BUPPER) ? ke - 16 : kb;
    const int nst = (ke - kb) / 16;
    op.load(kfirst, a, b);
    for (int st = 1, k0 = kfirst + DK; st < nst; ++st, k0 += DK) {
      v4d an[4], bn[4];
      op.load(k0, an, bn);
      mfma_step();
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) a[ti] = an[ti];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) b[tj] = bn[tj];
    }
    mfma_step();
  }
  tg_store<ACC>(acc, alpha, C, ld, r0, c0, li, lg);
}

// The same product for small blocks, split over K: one workgroup per 64 x 64
// block of C, its 4 waves take consecutive quarters of the block's K range
// (in 16-steps) and their partial sums are added through LDS in a fixed
// order.  Four times the waves of k_tgemm on the levels of the recursion
// whose products have few 64 x 64 blocks.  Grid: (row blocks, column blocks).
template <int TRI, bool ACC, bool ATR, bool BTR>
__global__ __launch_bounds__(256) void k_tgemm_sk(int M, int N, int K, double alpha,
                                                  const double* __restrict__ A,
                                                  const double* __restrict__ B,
                                                  double* __restrict__ C, int64_t ld) {
  __shared__ double red[64][64];   // one wave's partial block, lane-major
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  if (r0 >= M || c0 >= N) return;
  if (TRI == TG_CUPPER && r0 > c0) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int kb = (TRI == TG_BLOWER) ? c0 : 0;
  const int ke = (TRI == TG_BUPPER) ? min(c0 + 64, K) : K;
  const int q = (((ke - kb) / 16 + 3) / 4) * 16;      // this wave's share, whole 16-steps
  const int kbw = min(kb + w * q, ke), kew = min(kbw + q, ke);
  const TgOperands<ATR, BTR> op(A, B, ld, r0, c0, li, lg);

  v4d acc[4][4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
  v4d a[4], b[4];
  // operands of step k0; inside the diagonal block B's other triangle is masked
  auto load = [&](int k0, v4d (&av)[4], v4d (&bv)[4]) {
    op.load(k0, av, bv);
    if ((TRI == TG_BLOWER || TRI == TG_BUPPER) && k0 >= c0 && k0 < c0 + 64) {
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        const int c = c0 + 16 * tj + li;
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) {
          const int k = k0 + 4 * lg + s2;
          const bool keep = (TRI == TG_BUPPER) ? (k <= c) : (k >= c);
          bv[tj][s2] = keep ? bv[tj][s2] : 0.0;
        }
      }
    }
  };
  auto mfma_step = [&]() {
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = mfma_f64(b[tj][s2], a[ti][s2], acc[ti][tj]);
  };
  if (kbw < kew) {
    load(kbw, a, b);
    for (int k0 = kbw + 16; k0 < kew; k0 += 16) {
      v4d an[4], bn[4];
      load(k0, an, bn);
      mfma_step();
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) a[ti] = an[ti];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) b[tj] = bn[tj];
    }
    mfma_step();
  }
  // wave 0 adds the partial sums of waves 1, 2, 3 in turn
  for (int src = 1; src < 4; ++src) {
    if (w == src) {
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
          for (int v = 0; v < 4; ++v) red[16 * ti + 4 * tj + v][lane] = acc[ti][tj][v];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[ti][tj][v] += red[16 * ti + 4 * tj + v][lane];
    }
    __syncthreads();
  }
  if (w != 0) return;
  tg_store<ACC>(acc, alpha, C, ld, r0, c0, li, lg);
}

// Zero the strict upper triangle of the row-major N x N matrix T.
__global__ __launch_bounds__(256) void k_zero_upper(int N, double* __restrict__ T, int64_t ld) {
  const int r = blockIdx.x;
  for (int c = r + 1 + threadIdx.x; c < N; c += 256) T[(size_t)r * ld + c] = 0.0;
}

// dst(c, r) = src(r, c) for an R x C block (64-multiples): src(r, c) at
// src[r ld + c], written to dst[c ld + r].  Grid (R / 64, C / 64).
__global__ __launch_bounds__(256) void k_tcopy(const double* __restrict__ src,
                                               double* __restrict__ dst, int64_t ld) {
  __shared__ double t[64][65];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    const int rr = idx >> 6, cc = idx & 63;
    t[rr][cc] = src[(size_t)(r0 + rr) * ld + c0 + cc];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    const int cc = idx >> 6, rr = idx & 63;
    dst[(size_t)(c0 + cc) * ld + r0 + rr] = t[rr][cc];
  }
}

template <int TRI, bool ACC, bool ATR, bool BTR>
void tgemm_launch(int M, int N, int K, double alpha, const double* A, const double* B, double* C, int64_t ld,
                  hipStream_t st) {
  // 64 x 64 blocks that hold work (the symmetric update forms one triangle)
  const long blocks = (TRI == TG_CUPPER) ? (long)(M / 64) * (M / 64 + 1) / 2 : (long)(M / 64) * (N / 64);
  if (blocks <= FGLS_TG_SK_BLOCKS)   // few 64 x 64 blocks: split K over the waves
    hipLaunchKernelGGL((k_tgemm_sk<TRI, ACC, ATR, BTR>), dim3(M / 64, N / 64), dim3(256), 0, st, M, N, K, alpha,
                       A, B, C, ld);
  else
    hipLaunchKernelGGL((k_tgemm<TRI, ACC, ATR, BTR>), dim3((M + 127) / 128, (N + 127) / 128), dim3(256), 0, st,
                       M, N, K, alpha, A, B, C, ld);
}

// The second stream of the recursion: on the nodes of at most
// FGLS_SIDE_MAX_SIZE rows, X = L21 T11 needs only L21 and T11, so it runs on
// `s` (forked after S22 -= L21 L21') while the main stream recurses into
// S22, and the main stream joins before T21' = -X' T22'.  One fork and one
// join event serve every node: each wait is enqueued right after the record
// it needs, and a later record of `join` follows every earlier X on the
// in-order side stream.
struct Side {
  hipStream_t s;
  hipEvent_t fork, join;
};
#ifndef FGLS_SIDE_MAX_SIZE
#define FGLS_SIDE_MAX_SIZE 5120
#endif

// One node of the recursion on rows/columns [o, o + size) (see the top).
// The products: L21 = S21 T11' and T21' = -X' T22' (B upper, operands along
// k), X = L21 T11 (B lower, T11 read across k from its row-major lower
// triangle) and S22 -= L21 L21' (the symmetric update, accumulated).  L21 is
// read from its column-major copy across k on the small nodes; on the nodes
// with more than FGLS_TG_COPY_BLOCKS 64 x 64 blocks of L21 it is first copied
// row-major over S21 and read along k, which is faster there than the
// transposed reads (n = 10,000: the top two levels).
int chol_inv_node(int n, int o, int size, bool factor, double* L, double* T, int64_t ld, int* info,
                  hipStream_t st, const Side* sd) {
  if (size == 64) {
    if (factor) hipLaunchKernelGGL(k_leaf<true>, dim3(1), dim3(256), 0, st, n, o, L, T, ld, info);
    return FGLS_OK;   // without FACTOR every leaf ran up front
  }
  if (size == 128 && factor) {
    hipLaunchKernelGGL(k_node128, dim3(1), dim3(256), 0, st, n, o, L, T, ld, info);
    return FGLS_OK;
  }
  const int s = (size <= 128) ? 64 : ((size / 2 + 127) / 128) * 128;
  const int s2 = size - s;
  int rc = chol_inv_node(n, o, s, factor, L, T, ld, info, st, sd);
  if (rc) return rc;
  double* L21r = L + (size_t)(o + s) * ld + o;   // R-view lower-left: S21 (then L21 row-major)
  double* L21c = L + (size_t)o * ld + o + s;     // R-view upper-right: L21 column-major
  const double* T11 = T + (size_t)o * ld + o;
  const double* T22 = T + (size_t)(o + s) * ld + o + s;
  double* Xc = T + (size_t)o * ld + o + s;       // R-view upper-right: X column-major
  double* T21r = T + (size_t)(o + s) * ld + o;   // R-view lower-left: T21 row-major
  const bool big = (long)(s2 / 64) * (s / 64) > FGLS_TG_COPY_BLOCKS;
  const bool aside = sd && size <= FGLS_SIDE_MAX_SIZE;
  if (factor)
    tgemm_launch<TG_BUPPER, false, false, false>(s2, s, s, 1.0, L21r, T11, L21c, ld, st);   // L21 = S21 T11'
  if (big) hipLaunchKernelGGL(k_tcopy, dim3(s / 64, s2 / 64), dim3(256), 0, st, L21c, L21r, ld);
  if (factor) {
    double* S22 = L + (size_t)(o + s) * ld + o + s;   // S22 -= L21 L21'
    if (big) tgemm_launch<TG_CUPPER, true, false, false>(s2, s2, s, -1.0, L21r, L21r, S22, ld, st);
    else tgemm_launch<TG_CUPPER, true, true, true>(s2, s2, s, -1.0, L21c, L21c, S22, ld, st);
  }
  // X = L21 T11, forked after the symmetric update (beside which it would
  // only share the device) so that it runs beside the right child
  const hipStream_t xs = aside ? sd->s : st;
  if (aside) {
    HIP_TRY(hipEventRecord(sd->fork, st));
    HIP_TRY(hipStreamWaitEvent(sd->s, sd->fork, 0));
  }
  if (big) tgemm_launch<TG_BLOWER, false, false, true>(s2, s, s, 1.0, L21r, T11, Xc, ld, xs);
  else tgemm_launch<TG_BLOWER, false, true, true>(s2, s, s, 1.0, L21c, T11, Xc, ld, xs);
  if (aside) HIP_TRY(hipEventRecord(sd->join, sd->s));
  if ((rc = chol_inv_node(n, o + s, s2, factor, L, T, ld, info, st, sd))) return rc;
  if (aside) HIP_TRY(hipStreamWaitEvent(st, sd->join, 0));
  tgemm_launch<TG_BUPPER, false, false, false>(s, s2, s2, -1.0, Xc, T22, T21r, ld, st);   // T21' = -X' T22'
  return FGLS_OK;
}

int factor_on_device(fgls_hip_ctx* c, const double* dPhi, int64_t ldphi, double h2, double sigma2,
                     hipStream_t st) {
  const int n = c->n, ldn = c->ldn;
  HIP_TRY(hipMemsetAsync(c->status + 1, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_build_M, dim3(ldn / 64, ldn / 64), dim3(256), 0, st, n, ldn, dPhi, ldphi,
                     sigma2 * h2, sigma2 * (1.0 - h2), c->L);
  HIP_TRY(hipGetLastError());
  return fgls::chol_inv(c, true, st);
}

}  // namespace

namespace fgls {

int chol_inv(fgls_hip_ctx* c, bool factor, hipStream_t st) {
  const int n = c->n, N = c->ldn;
  const int64_t ld = N;
  if (!c->side) {
    HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  const Side sd{c->side, c->ev_fork, c->ev_join};
  if (!factor)
    hipLaunchKernelGGL(k_leaf<false>, dim3(N / 64), dim3(256), 0, st, n, 0, c->L, c->Linv, ld,
                       (int*)nullptr);
  int rc = chol_inv_node(n, 0, N, factor, c->L, c->Linv, ld, c->status + 1, st, &sd);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_upper, dim3(N), dim3(256), 0, st, N, c->Linv, ld);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

}  // namespace fgls

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
