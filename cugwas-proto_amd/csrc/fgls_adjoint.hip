// fgls_adjoint.hip -- adjoint of the FGLS solver for the loss
//   l = sum_i w_i beta_i^2,   beta_i = b_i[p-1] (the SNP coefficient),
// with respect to y and h^2 (M = sigma2 (h^2 Phi + (1 - h^2) I) = L L').
// The reference has no adjoint; this is an addition on top of its forward
// path (b_i as in fgls_chol.c:386-439).
//
// Derivation (X_i = [X_L, x_i], S_i = X_i' M^-1 X_i, r_i = y - X_i b_i):
//   d beta_i = e_p' S_i^-1 X_i' dM^-1 (y - X_i b_i)
//            = -(M^-1 X_i u_i)' dM (M^-1 r_i),        u_i = S_i^-1 e_p
//   with g_i = 2 w_i beta_i u_i (p values):
//   dl/dy   = M^-1 sum_i X_i g_i = L^-T (X~_L sum_i g_L,i + sum_i gamma_i z_i)
//   dl/dh^2 = -sigma2 sum_i p~_i' G q~_i,   G = L^-1 (Phi - I) L^-T,
//   p~_i = X~_i g_i = W g^_i + gamma_i z_i,  q~_i = L^-1 r_i = W b^_i - beta_i z_i,
// where z_i = L^-1 x_i, W = [X~_L | y~], gamma_i = g_i[p-1],
// g^_i = [g_L,i ; 0] and b^_i = [-b_L,i ; 1].  Expanding with s_i = W' G z_i
// and t_i = z_i' G z_i:
//   p~_i' G q~_i = g^' (W'GW) b^ - beta (g^ . s) + gamma (s . b^) - beta gamma t.
// Per block of SNPs this needs Z = L^-1 X_R (recomputed, not stored from
// the forward pass), b_i and S_i^-1 e_p (the forward per-SNP stage on Z),
// one dense product G Z whose epilogue reduces t_i and s_i, and the sum
// Z gamma.  G, G W and W' G W are built once per trait.
//
// Mixed mode (int8 genotypes): G Z = G L^-1 X_R = H X_R with H = G L^-1
// built once per trait and sliced like L^-1, so both per-block products are
// exact int8 products against the genotypes: Z = L^-1 X_R written out, and
// Y = H X_R with t_i = y_i' z_i and s_i = W' y_i fused into its epilogue.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "fgls_ctx.h"
#include "fgls_device.h"
#include "fgls_hip.h"

struct fgls_adjoint_state {
  double sigma2;
  bool ready;      // adjoint_begin done for the current trait
  double* G;       // ldn x ldn, L^-1 (Phi - I) L^-T, zero padded
  double* T;       // ldn x ldn scratch
  double* GW;      // ldn x p: G W
  double* WGW;     // p x p (column-major)
  double* Bs;      // max_block x p: b of the block
  double* Vs;      // max_block x p^2: V of the block
  double* gam;     // max_block: gamma_i
  double* rec;     // max_block x (p + 1): w beta^2, p~'Gq~, g_L (p - 1)
  double* gzp;     // nchunk x ldn: partial sums of gamma_i z_i
  double* gz;      // ldn: sum_i gamma_i z_i
  double* acc;     // p + 1: loss, sum p~'Gq~, sum g_L
  double* gyt;     // ldn: X~_L sum g_L + gz
  double* H;       // ldn x ldn row-major G L^-1 (mixed mode)
  int8_t* H8;      // its int8 digit slices
  double* hinvsc;  // their row scales
  bool h8_ready;   // H8 built for the current trait
};

namespace {

constexpr int CHUNK = 256;  // SNPs per partial of sum_i gamma_i z_i

using fgls::fail;

#define HIP_TRY(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(FGLS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Sum over a 256-thread workgroup in a fixed order.
__device__ double group_sum(double x, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  x = wave_sum(x);
  __syncthreads();
  if (lane == 0) red[w] = x;
  __syncthreads();
  const double s = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ x, size_t len, double v) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < len; i += (size_t)gridDim.x * 256)
    x[i] = v;
}

// out (ldn x ldn, column-major) = Phi - I on the n x n block, 0 elsewhere.
__global__ __launch_bounds__(256) void k_phi_minus_identity(int n, int ldn,
                                                            const double* __restrict__ Phi,
                                                            int64_t ldphi, double* __restrict__ out) {
  const int c = blockIdx.y;
  for (int r = blockIdx.x * 256 + threadIdx.x; r < ldn; r += gridDim.x * 256) {
    double v = 0.0;
    if (r < n && c < n) v = Phi[(size_t)c * ldphi + r] - (r == c ? 1.0 : 0.0);
    out[(size_t)c * ldn + r] = v;
  }
}

// out = in' for ldn x ldn column-major matrices, through 32 x 33 LDS tiles.
__global__ __launch_bounds__(256) void k_transpose(int ldn, const double* __restrict__ in,
                                                   double* __restrict__ out) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty 0..7
#pragma unroll
  for (int j = 0; j < 32; j += 8) tile[ty + j][tx] = in[(size_t)(by + ty + j) * ldn + bx + tx];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 32; j += 8) out[(size_t)(bx + ty + j) * ldn + by + tx] = tile[tx][ty + j];
}

// The 64 x 64 blocks above the diagonal of a symmetric ldn x ldn column-major
// matrix from those below it, through 32 x 33 LDS tiles (32-tiles whose
// 64-blocks lie above the diagonal; the others exit).
__global__ __launch_bounds__(256) void k_mirror_upper(int ldn, double* __restrict__ Z) {
  __shared__ double tile[32][33];
  const int ti = blockIdx.x, tj = blockIdx.y;   // 32-tile (rows, columns) to fill
  if ((ti >> 1) >= (tj >> 1)) return;
  const int r0 = ti * 32, c0 = tj * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty 0..7
  // source: Z(c0 + tx, r0 + ty + j), below the diagonal
#pragma unroll
  for (int j = 0; j < 32; j += 8) tile[ty + j][tx] = Z[(size_t)(r0 + ty + j) * ldn + c0 + tx];
  __syncthreads();
  // Z(r0 + tx, c0 + ty + j) = Z(c0 + ty + j, r0 + tx)
#pragma unroll
  for (int j = 0; j < 32; j += 8) Z[(size_t)(c0 + ty + j) * ldn + r0 + tx] = tile[tx][ty + j];
}

// WGW = W' (G W), p x p column-major; one workgroup.
__global__ __launch_bounds__(256) void k_wgw(int n, int p, const double* __restrict__ W,
                                             const double* __restrict__ GW, int64_t ldn,
                                             double* __restrict__ WGW) {
  __shared__ double red[4];
  for (int a = 0; a < p; ++a)
    for (int b = 0; b < p; ++b) {
      double s = 0.0;
      for (int r = threadIdx.x; r < n; r += 256) s = fma(W[(size_t)a * ldn + r], GW[(size_t)b * ldn + r], s);
      s = group_sum(s, red);
      if (threadIdx.x == 0) WGW[b * p + a] = s;
    }
}

// Per SNP (thread): t_i, s_i from the dense-product partials; b_i and V_i of
// the forward stage; then gamma_i and the record (w beta^2, p~'Gq~, g_L).
template <int P>
__global__ __launch_bounds__(128) void k_adj_snp(int count, int RT, const double* __restrict__ part,
                                                 int64_t ldp, const double* __restrict__ Bs,
                                                 const double* __restrict__ Vs,
                                                 const double* __restrict__ trait,
                                                 const double* __restrict__ wt,
                                                 const double* __restrict__ WGW,
                                                 double* __restrict__ gam, double* __restrict__ rec) {
  constexpr int NQ = P + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double q[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = 0.0;
  for (int rt = 0; rt < RT; ++rt) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] += part[((size_t)rt * NQ + k) * ldp + i];
  }
  const double t = q[0];  // z' G z; s_k = q[1 + k] = W_k' G z
  const double res = trait[fgls::trait_res_off(P - 1)];
  const double* b = Bs + (size_t)i * P;
  const double* V = Vs + (size_t)i * P * P;
  const double beta = b[P - 1];
  // u = S^-1 e_p: row p-1 of V (lower triangle) / res; its diagonal holds
  // sqrt(res * S^-1_{p-1,p-1})
  double g[P], bh[P];
  const double vd = V[(P - 1) * P + (P - 1)];
  const double scale = 2.0 * wt[i] * beta / res;
#pragma unroll
  for (int c = 0; c < P - 1; ++c) g[c] = scale * V[c * P + (P - 1)];
  g[P - 1] = scale * vd * vd;
  const double gamma = g[P - 1];
#pragma unroll
  for (int c = 0; c < P - 1; ++c) bh[c] = -b[c];
  bh[P - 1] = 1.0;
  // g^' WGW b^ with g^ = [g_L ; 0]
  double quad = 0.0;
#pragma unroll
  for (int a = 0; a < P - 1; ++a) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < P; ++c) s = fma(WGW[c * P + a], bh[c], s);
    quad = fma(g[a], s, quad);
  }
  double gs = 0.0, sb = 0.0;
#pragma unroll
  for (int a = 0; a < P - 1; ++a) gs = fma(g[a], q[1 + a], gs);
#pragma unroll
  for (int a = 0; a < P; ++a) sb = fma(q[1 + a], bh[a], sb);
  const double h = quad - beta * gs + gamma * sb - beta * gamma * t;
  gam[i] = gamma;
  double* r = rec + (size_t)i * NQ;
  r[0] = wt[i] * beta * beta;
  r[1] = h;
#pragma unroll
  for (int a = 0; a < P - 1; ++a) r[2 + a] = g[a];
}

// gzp[chunk * ldn + r] = sum over the chunk's SNPs of gamma_i z_i[r].
__global__ __launch_bounds__(256) void k_adj_gz_partial(int n, int count, const double* __restrict__ Z,
                                                        int64_t ldz, const double* __restrict__ gam,
                                                        double* __restrict__ gzp, int64_t ldn) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int i0 = blockIdx.y * CHUNK, i1 = min(i0 + CHUNK, count);
  if (r >= n) return;
  double s = 0.0;
  for (int i = i0; i < i1; ++i) s = fma(Z[(size_t)i * ldz + r], gam[i], s);
  gzp[(size_t)blockIdx.y * ldn + r] = s;
}

// gz[r] += sum over chunks (in order); acc[j] += sum_i rec[i][j] (one
// workgroup per j, fixed order).
__global__ __launch_bounds__(256) void k_adj_gz_reduce(int n, int nchunk, const double* __restrict__ gzp,
                                                       int64_t ldn, double* __restrict__ gz) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += gzp[(size_t)k * ldn + r];
  gz[r] += s;
}

__global__ __launch_bounds__(256) void k_adj_accum(int count, int nq, const double* __restrict__ rec,
                                                   double* __restrict__ acc) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += 256) s += rec[(size_t)i * nq + j];
  s = group_sum(s, red);
  if (threadIdx.x == 0) acc[j] += s;
}

// gyt = X~_L sum g_L + gz;  grad_y = L^-T gyt (Linv row-major: column c of
// L^-1 is read down the rows r >= c).
__global__ __launch_bounds__(256) void k_adj_gyt(int n, int wXL, const double* __restrict__ W,
                                                 int64_t ldn, const double* __restrict__ acc,
                                                 const double* __restrict__ gz,
                                                 double* __restrict__ gyt) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double s = gz[r];
  for (int k = 0; k < wXL; ++k) s = fma(W[(size_t)k * ldn + r], acc[2 + k], s);
  gyt[r] = s;
}

__global__ __launch_bounds__(256) void k_adj_grad_y(int n, const double* __restrict__ Linv,
                                                    int64_t ldn, const double* __restrict__ gyt,
                                                    double* __restrict__ out) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  double s = 0.0;
  for (int r = c; r < n; ++r) s = fma(Linv[(size_t)r * ldn + c], gyt[r], s);
  out[c] = s;
}

template <int P>
void launch_adj_snp(const fgls_hip_ctx* c, int count, hipStream_t st, const double* dw, int RT) {
  const fgls_adjoint_state* a = c->adj;
  hipLaunchKernelGGL(k_adj_snp<P>, dim3((count + 127) / 128), dim3(128), 0, st, count, RT,
                     c->part, (int64_t)c->max_block, a->Bs, a->Vs, c->trait, dw, a->WGW, a->gam,
                     a->rec);
}

int fill(double* x, size_t len, double v, hipStream_t st) {
  const size_t blocks = std::min<size_t>((len + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, x, len, v);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int ensure_state(fgls_hip_ctx* c) {
  if (!c->z) {
    const size_t bytes = (size_t)c->ldn * c->max_block * sizeof(double);
    if (hipMalloc((void**)&c->z, bytes) != hipSuccess)
      return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
  }
  if (c->adj) return FGLS_OK;
  fgls_adjoint_state* a = new fgls_adjoint_state();
  memset(a, 0, sizeof(*a));
  c->adj = a;
  const size_t ldn = (size_t)c->ldn, p = (size_t)c->p, mb = (size_t)c->max_block;
  const size_t nchunk = (mb + CHUNK - 1) / CHUNK;
  struct { double** ptr; size_t len; } allocs[] = {
      {&a->G, ldn * ldn},     {&a->T, ldn * ldn},       {&a->GW, ldn * p},
      {&a->WGW, p * p},       {&a->Bs, mb * p},         {&a->Vs, mb * p * p},
      {&a->gam, mb},          {&a->rec, mb * (p + 1)},  {&a->gzp, nchunk * ldn},
      {&a->gz, ldn},          {&a->acc, p + 1},         {&a->gyt, ldn},
  };
  for (auto& al : allocs) {
    if (hipMalloc((void**)al.ptr, al.len * sizeof(double)) != hipSuccess)
      return fail(FGLS_ERR_NOMEM, "hipMalloc of %zu bytes failed", al.len * sizeof(double));
    if (hipMemset(*al.ptr, 0, al.len * sizeof(double)) != hipSuccess)
      return fail(FGLS_ERR_HIP, "hipMemset failed");
  }
  return FGLS_OK;
}

}  // namespace

void fgls_adjoint_free(fgls_adjoint_state* a) {
  if (!a) return;
  double* ptrs[] = {a->G, a->T, a->GW, a->WGW, a->Bs, a->Vs, a->gam, a->rec, a->gzp, a->gz,
                    a->acc, a->gyt, a->H, a->hinvsc};
  for (double* p : ptrs)
    if (p) hipFree(p);
  if (a->H8) hipFree(a->H8);
  delete a;
}

void fgls_adjoint_invalidate(fgls_adjoint_state* a) {
  if (a) a->ready = a->h8_ready = false;
}

extern "C" {

int fgls_hip_adjoint_begin(fgls_hip_ctx* c, const double* dPhi, int64_t ldphi, double sigma2,
                           void* stream) {
  if (!c) return fail(FGLS_ERR_ARG, "NULL context");
  if (!c->have_trait) return fail(FGLS_ERR_STATE, "no trait set");
  if (c->wXR != 1) return fail(FGLS_ERR_ARG, "the adjoint needs wXR = 1 (got %d)", c->wXR);
  if (!dPhi || ldphi < c->n) return fail(FGLS_ERR_ARG, "Phi needs a device pointer and ld >= n");
  if (!(sigma2 > 0.0)) return fail(FGLS_ERR_ARG, "sigma2 must be positive");
  int rc = fgls::set_device(c);
  if (rc) return rc;
  if ((rc = ensure_state(c))) return rc;
  fgls_adjoint_state* a = c->adj;
  a->ready = a->h8_ready = false;
  a->sigma2 = sigma2;
  hipStream_t st = (hipStream_t)stream;
  const int n = c->n, ldn = c->ldn, p = c->p;
  // T = L^-1 (Phi - I);  G = L^-1 T' = L^-1 (Phi - I) L^-T  (Phi symmetric)
  hipLaunchKernelGGL(k_phi_minus_identity, dim3((ldn + 255) / 256, ldn), dim3(256), 0, st, n, ldn,
                     dPhi, ldphi, a->G);
  fgls::launch_product_store(c, c->Linv, false, a->G, ldn, n, a->T, ldn, st);
  hipLaunchKernelGGL(k_transpose, dim3(ldn / 32, ldn / 32), dim3(256), 0, st, ldn, a->T, a->G);
  // G is symmetric: its blocks on and below the diagonal, then the mirror
  fgls::launch_product_store_lower_half(c, a->G, ldn, a->T, ldn, st);
  hipLaunchKernelGGL(k_mirror_upper, dim3(ldn / 32, ldn / 32), dim3(256), 0, st, ldn, a->T);
  // columns n..ldn-1 of the product were not written: T holds G with the
  // padding of the buffer it was cleared with; copy it to G
  HIP_TRY(hipMemcpyAsync(a->G, a->T, (size_t)ldn * ldn * sizeof(double), hipMemcpyDeviceToDevice, st));
  // GW = G W (G is symmetric, so its column-major buffer is also row-major)
  fgls::launch_product_store(c, a->G, true, c->W, ldn, p, a->GW, ldn, st);
  hipLaunchKernelGGL(k_wgw, dim3(1), dim3(256), 0, st, n, p, c->W, a->GW, (int64_t)ldn, a->WGW);
  HIP_TRY(hipGetLastError());
  if (c->mode == FGLS_MODE_MIXED_I8) {
    // H = G L^-1 row-major = (L^-T G)' : T = (L^-1)' as a row-major L^-T,
    // then the upper triangular product T G (G symmetric) stored column-major
    const size_t nn = (size_t)ldn * ldn;
    if (!a->H) {
      if (hipMalloc((void**)&a->H, nn * sizeof(double)) != hipSuccess ||
          hipMalloc((void**)&a->H8, nn * FGLS_NSLICE) != hipSuccess ||
          hipMalloc((void**)&a->hinvsc, (size_t)ldn * sizeof(double)) != hipSuccess)
        return fail(FGLS_ERR_NOMEM, "hipMalloc of the mixed adjoint's %zu bytes failed",
                    nn * (sizeof(double) + FGLS_NSLICE));
      if ((rc = fill(a->H, nn, 0.0, st))) return rc;
    }
    hipLaunchKernelGGL(k_transpose, dim3(ldn / 32, ldn / 32), dim3(256), 0, st, ldn, c->Linv, a->T);
    fgls::launch_product_store_upper(c, a->T, a->G, ldn, n, a->H, ldn, st);
    fgls::launch_slices(ldn, a->H, a->H8, a->hinvsc, st);
    HIP_TRY(hipGetLastError());
    a->h8_ready = true;
  }
  if ((rc = fill(a->gz, (size_t)ldn, 0.0, st))) return rc;
  if ((rc = fill(a->acc, (size_t)p + 1, 0.0, st))) return rc;
  a->ready = true;
  return FGLS_OK;
}

int fgls_hip_adjoint_block(fgls_hip_ctx* c, const double* dXR, int64_t ldxr, int count,
                           const double* dw, int64_t snp_base, void* stream) {
  if (!c) return fail(FGLS_ERR_ARG, "NULL context");
  if (!c->adj || !c->adj->ready) return fail(FGLS_ERR_STATE, "fgls_hip_adjoint_begin not called");
  if (count < 0 || count > c->max_block)
    return fail(FGLS_ERR_ARG, "count=%d outside [0, %d]", count, c->max_block);
  if (count == 0) return FGLS_OK;
  if (!dXR || !dw) return fail(FGLS_ERR_ARG, "NULL pointer");
  if (ldxr < c->n || (ldxr & 3) || ((uintptr_t)dXR & 31))
    return fail(FGLS_ERR_ARG, "X needs ld >= n, ld %% 4 == 0 and 32-byte alignment");
  int rc = fgls::set_device(c);
  if (rc) return rc;
  fgls_adjoint_state* a = c->adj;
  hipStream_t st = (hipStream_t)stream;
  const int n = c->n, ldn = c->ldn, p = c->p;
  // Z = L^-1 X_R, then b and V of every SNP of the block
  fgls::launch_product_store(c, c->Linv, false, dXR, ldxr, count, c->z, ldn, st);
  fgls::launch_snp_stage_any(c, c->z, ldn, count, a->Bs, a->Vs, snp_base, st);
  // t_i = z' G z and s_i = W' G z from one dense product
  fgls::launch_product_dot(c, a->G, c->z, ldn, count, st);
  FGLS_P_SWITCH(p, launch_adj_snp, c, count, st, dw, c->RT);
  const int nchunk = (count + CHUNK - 1) / CHUNK;
  hipLaunchKernelGGL(k_adj_gz_partial, dim3((n + 255) / 256, nchunk), dim3(256), 0, st, n, count,
                     c->z, (int64_t)ldn, a->gam, a->gzp, (int64_t)ldn);
  hipLaunchKernelGGL(k_adj_gz_reduce, dim3((n + 255) / 256), dim3(256), 0, st, n, nchunk, a->gzp,
                     (int64_t)ldn, a->gz);
  hipLaunchKernelGGL(k_adj_accum, dim3(p + 1), dim3(256), 0, st, count, p + 1, a->rec, a->acc);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_adjoint_block_i8(fgls_hip_ctx* c, const int8_t* dX8, int64_t ldx8, int count,
                              const double* dw, int64_t snp_base, void* stream) {
  if (!c) return fail(FGLS_ERR_ARG, "NULL context");
  if (!c->adj || !c->adj->ready) return fail(FGLS_ERR_STATE, "fgls_hip_adjoint_begin not called");
  if (c->mode != FGLS_MODE_MIXED_I8 || !c->adj->h8_ready)
    return fail(FGLS_ERR_STATE, "fgls_hip_adjoint_begin was not called in the mixed mode");
  if (count < 0 || count > c->max_block)
    return fail(FGLS_ERR_ARG, "count=%d outside [0, %d]", count, c->max_block);
  if (count == 0) return FGLS_OK;
  if (!dX8 || !dw) return fail(FGLS_ERR_ARG, "NULL pointer");
  const int64_t ld8 = (c->n + 63) / 64 * 64;
  if (ldx8 < ld8 || (ldx8 & 15) || ((uintptr_t)dX8 & 15))
    return fail(FGLS_ERR_ARG, "int8 X_R needs ld >= ceil(n/64)*64, ld %% 16 == 0 and 16-byte alignment");
  int rc = fgls::set_device(c);
  if (rc) return rc;
  if ((rc = fgls::ensure_buffer(&c->X8t, &c->X8t_cap, fgls_x8t_bytes(c->n, count, 1)))) return rc;
  fgls_adjoint_state* a = c->adj;
  hipStream_t st = (hipStream_t)stream;
  const int n = c->n, ldn = c->ldn, p = c->p;
  fgls_launch_tile_i8(n, count, 1, dX8, ldx8, c->X8t, st);
  // Z = L^-1 X_R on the int8 slices, then b and V of every SNP of the block
  fgls_launch_i8_store(n, count, c->RT8, ldn / 64, c->A8, c->invsc, c->X8t, ld8, c->z, ldn, c->nslice, st);
  fgls::launch_snp_stage_any(c, c->z, ldn, count, a->Bs, a->Vs, snp_base, st);
  // Y = H X_R = G Z on the slices of H: t_i = y_i' z_i, s_i = W' y_i
  fgls_launch_i8_dot(p, n, count, c->RT8, ldn / 64, a->H8, a->hinvsc, c->X8t, ld8, c->W, ldn, c->z,
                     ldn, c->part, (int64_t)c->max_block, st);
  FGLS_P_SWITCH(p, launch_adj_snp, c, count, st, dw, c->RT8);
  const int nchunk = (count + CHUNK - 1) / CHUNK;
  hipLaunchKernelGGL(k_adj_gz_partial, dim3((n + 255) / 256, nchunk), dim3(256), 0, st, n, count,
                     c->z, (int64_t)ldn, a->gam, a->gzp, (int64_t)ldn);
  hipLaunchKernelGGL(k_adj_gz_reduce, dim3((n + 255) / 256), dim3(256), 0, st, n, nchunk, a->gzp,
                     (int64_t)ldn, a->gz);
  hipLaunchKernelGGL(k_adj_accum, dim3(p + 1), dim3(256), 0, st, count, p + 1, a->rec, a->acc);
  HIP_TRY(hipGetLastError());
  return FGLS_OK;
}

int fgls_hip_adjoint_end(fgls_hip_ctx* c, double* dgrad_y, double* hloss, double* hgrad_h2,
                         void* stream) {
  if (!c) return fail(FGLS_ERR_ARG, "NULL context");
  if (!c->adj || !c->adj->ready) return fail(FGLS_ERR_STATE, "fgls_hip_adjoint_begin not called");
  int rc = fgls::set_device(c);
  if (rc) return rc;
  fgls_adjoint_state* a = c->adj;
  hipStream_t st = (hipStream_t)stream;
  const int n = c->n, ldn = c->ldn, p = c->p;
  if (dgrad_y) {
    hipLaunchKernelGGL(k_adj_gyt, dim3((n + 255) / 256), dim3(256), 0, st, n, c->wXL, c->W,
                       (int64_t)ldn, a->acc, a->gz, a->gyt);
    hipLaunchKernelGGL(k_adj_grad_y, dim3((n + 255) / 256), dim3(256), 0, st, n, c->Linv,
                       (int64_t)ldn, a->gyt, dgrad_y);
    HIP_TRY(hipGetLastError());
  }
  double acc[2] = {0.0, 0.0};
  HIP_TRY(hipMemcpyAsync(acc, a->acc, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  (void)p;
  if (hloss) *hloss = acc[0];
  if (hgrad_h2) *hgrad_h2 = -a->sigma2 * acc[1];
  return FGLS_OK;
}

}  // extern "C"
