// fgls_device.h -- device helpers shared by the translation units of the
// FGLS device layer (internal; not part of the C-ABI).
#ifndef FGLS_DEVICE_H
#define FGLS_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double fgls_v4d __attribute__((ext_vector_type(4)));
typedef int fgls_v4i __attribute__((ext_vector_type(4)));

// Mixed mode: L^-1 as NSLICE int8 digit slices; the int8 product runs on
// workgroup tiles of I8_TM rows x I8_TN SNP columns, grouped I8_GR x I8_GC
// per XCD (see fgls_tile_of_block).
constexpr int FGLS_NSLICE = 6;
// Radix of the slice digits (balanced, digits in [-127, 127]) and the
// largest n for which the int32 sums stay exact: 127 * 127 * n < 2^31.
constexpr int FGLS_SLICE_RADIX = 254;
constexpr int FGLS_I8_MAX_N = 133000;
// Five-digit form of a triangular L^-1 (k_slice_tri, five digits): row i scaled by
// its largest entry outside its diagonal 64-k step, five digits everywhere
// and one more (top) digit on the diagonal step.  Used when its error bound,
// at most FGLS_TOP_MAX_LOSS times the six-digit one, allows, and for
// n <= FGLS_I8_TOP_MAX_N, where the fold of the top digit into digit 0
// (127 * 127 * (n + 254 * 64) < 2^31) stays exact.
constexpr int FGLS_TOP_NSLICE = 5;
constexpr double FGLS_TOP_MAX_LOSS = 16.0;
constexpr int FGLS_I8_TOP_MAX_N = 116000;
// The default grouping is a build-time tunable (-DFGLS_I8_GR=.. -DFGLS_I8_GC=..);
// the variants of the int8 product that measured faster with another shape
// set their own (I8Group in fgls_i8.hip).
#ifndef FGLS_I8_GR
#define FGLS_I8_GR 16
#endif
#ifndef FGLS_I8_GC
#define FGLS_I8_GC 4
#endif
constexpr int I8_TM = 32, I8_TN = 256, I8_GR = FGLS_I8_GR, I8_GC = FGLS_I8_GC;

// Tile order of the hot kernels: workgroup b of a 1-D grid runs on XCD b % 8
// (round-robin dispatch), so each XCD takes whole groups of GR x GC tiles (one
// group = the workgroups its 32 CUs hold at once, sharing row panels of L^-1
// and column panels of X_R in that XCD's L2).  Groups are dealt to the XCDs
// in turn, heaviest row tiles (longest K) first.  Returns false for padding
// workgroups.
template <int GR, int GC>
__host__ __device__ __forceinline__ bool fgls_tile_of_block(int b, int RT, int CT, int& rt, int& ct) {
  constexpr int GS = GR * GC;
  const int xcd = b & 7, k = b >> 3;
  const int group = (k / GS) * 8 + xcd, j = k % GS;
  const int NGC = (CT + GC - 1) / GC, NGR = (RT + GR - 1) / GR;
  if (group >= NGR * NGC) return false;
  const int gr = NGR - 1 - group / NGC, gc = group % NGC;
  rt = gr * GR + j / GC;
  ct = gc * GC + j % GC;
  return rt < RT && ct < CT;
}

__host__ __device__ constexpr int fgls_tile_grid(int RT, int CT, int GR, int GC) {
  return (((RT + GR - 1) / GR) * ((CT + GC - 1) / GC) + 7) / 8 * 8 * GR * GC;
}

// Slice layout of the int8 product: 1 KB blocks, one per (16-row tile rt16,
// 64-k step ks, slice s), each in the A-operand lane order of
// v_mfma_i32_16x16x64_i8: lane 16*g + i holds k = 64 ks + 16 g .. +15 of row
// 16 rt16 + i.  KS = ldn / 64 steps per row tile; one step of a row tile
// is S contiguous blocks.
__host__ __device__ constexpr size_t fgls_a8_offset(int row, int s, int kc, int S, int KS) {
  return (((((size_t)(row >> 4) * KS + (kc >> 2)) * S + s) * 64) + 16 * (kc & 3) + (row & 15)) * 16;
}

// Top-digit blocks of the five-digit form: one 1 KB block per 16-row tile
// (its diagonal 64-k step), in the same lane order; they follow the five
// slices (fgls_a8_offset with S = 5) in the slice buffer, at byte
// fgls_a8top_base(KS).
__host__ __device__ constexpr size_t fgls_a8top_offset(int row, int kc) {
  return ((size_t)(row >> 4) * 64 + 16 * (kc & 3) + (row & 15)) * 16;
}
__host__ __device__ constexpr size_t fgls_a8top_base(int KS) {
  return (size_t)KS * 64 * KS * 64 * FGLS_TOP_NSLICE;
}

// Bytes of the tiled genotypes of `count` columns and `planes` planes.
__host__ __device__ constexpr size_t fgls_x8t_bytes(int n, int count, int planes) {
  return (size_t)planes * (size_t)((count + 15) / 16) * 16 * (size_t)((n + 63) / 64) * 64;
}
// Offset of the hi plane in the tiled genotypes of `count` columns.
__host__ __device__ constexpr int64_t fgls_x8t_plane(int n, int count) {
  return (int64_t)fgls_x8t_bytes(n, count, 1);
}
// Bytes of 2-bit tiled genotypes (values 0..3, hard calls) of `count`
// columns: one 256-byte block per (16 SNP columns, 64-individual step), in
// the B-operand lane order of the int8 product; lane l = 16 g + i of a block
// holds, in one dword, the 16 genotypes k = 64 ks + 16 g + j (j < 16) of
// column 16 cg + i, genotype j = 4 q + b in bits 8 b + 2 q .. 8 b + 2 q + 1,
// so that (x >> 2 q) & 0x03030303 is dword q of the lane's int8 operand.
__host__ __device__ constexpr size_t fgls_x2t_bytes(int n, int count) {
  return (size_t)((count + 15) / 16) * 16 * (size_t)((n + 63) / 64) * 16;
}
// Column-major int8 genotypes (ld ldx8 >= ceil(n/64)*64, ldx8 % 16 == 0,
// 16-byte aligned) -> 2-bit tiles; values of rows < n outside 0..3 are
// counted in *bad and written as 0.
void fgls_launch_tile_u2(int n, int count, const int8_t* X8, int64_t ldx8, uint8_t* X2T, unsigned int* bad,
                         hipStream_t st);
// 2-bit genotype file columns (ceil(n/4) bytes at P + c ldp, individual r in
// bits 2 (r % 4) of byte r / 4) -> 2-bit tiles; code 3 in rows < n is
// counted in *bad and written as 0.
void fgls_launch_tile_u2_packed(int n, int count, const uint8_t* P, int64_t ldp, uint8_t* X2T,
                                unsigned int* bad, hipStream_t st);
// The int8 product on 2-bit tiles (fgls_launch_trmm_i8 with X8T in the 2-bit
// layout; genotypes only).
void fgls_launch_trmm_u2(int p, int n, int count, int RT32, int KS, const int8_t* A8,
                         const double* inv_scale, const uint8_t* X2T, const double* W, int64_t ldw,
                         double* part, int64_t ldp, int eig_ntr, int nslice, hipStream_t st);
// Column-major genotypes (ld ldx8 >= ceil(n/64)*64, ldx8 % 16 == 0, 16-byte
// aligned; `planes` planes of count columns, count ldx8 bytes apart) -> the
// tiled layout of the int8 product (fgls_i8.hip).
void fgls_launch_tile_i8(int n, int count, int planes, const int8_t* X8, int64_t ldx8, int8_t* X8T,
                         hipStream_t st);
// Launch of the int8 triangular product with fused reductions (fgls_i8.hip),
// for p = 1..16, on tiled genotypes X8T (ld8 = ceil(n/64)*64).  part receives
// (p+1) sums per SNP per 32-row tile.  With plane_off > 0 the genotypes are
// dosages as two int8 digit planes (lo at X8T, hi at X8T + plane_off,
// x = (128 hi + lo) * inv_xscale).  eig_ntr > 0: A8
// holds a full matrix and W has eig_ntr groups of p+1 weight columns (the
// eigenbasis variant; part then holds eig_ntr (p+1) sums per SNP and tile).
// nslice: the digits of a triangular A8, FGLS_NSLICE (k_slice_tri, six digits) or
// FGLS_TOP_NSLICE (five digits, the top digits at
// fgls_a8top_base); a full A8 (eig_ntr > 0) always has FGLS_NSLICE.
void fgls_launch_trmm_i8(int p, int n, int count, int RT32, int KS, const int8_t* A8,
                         const double* inv_scale, const int8_t* X8, int64_t ldx8, const double* W,
                         int64_t ldw, double* part, int64_t ldp, int64_t plane_off,
                         double inv_xscale, int eig_ntr, int nslice, hipStream_t st);

// The int8 product z = A x (triangular A, as for the reductions) written to
// Z (column-major, ld ldz >= RT32 * 32, 32-byte aligned); and, for a full A,
// y = A x reduced per SNP against Zin and W: part[(rt (p+1) + 0) ldp + col] =
// y' Zin[:, col], part[(rt (p+1) + 1 + k) ldp + col] = y' W[:, k] (k < p).
void fgls_launch_i8_store(int n, int count, int RT32, int KS, const int8_t* A8, const double* inv_scale,
                          const int8_t* X8T, int64_t ld8, double* Z, int64_t ldz, int nslice, hipStream_t st);
// The same store from 2-bit genotype tiles (fgls_x2t_bytes layout).
void fgls_launch_u2_store(int n, int count, int RT32, int KS, const int8_t* A8, const double* inv_scale,
                          const uint8_t* X2T, double* Z, int64_t ldz, int nslice, hipStream_t st);
void fgls_launch_i8_dot(int p, int n, int count, int RT32, int KS, const int8_t* A8,
                        const double* inv_scale, const int8_t* X8T, int64_t ld8, const double* W,
                        int64_t ldw, const double* Zin, int64_t ldz, double* part, int64_t ldp,
                        hipStream_t st);

#endif
