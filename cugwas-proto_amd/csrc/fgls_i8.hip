// fgls_i8.hip -- the mixed-precision hot kernel: int8 genotypes times the
// int8 digit slices of L^-1 (six, or five with a top digit on the diagonal
// 64-k steps) on v_mfma_i32_16x16x64_i8, exact int32 accumulation, FP64
// recombination and the per-SNP reductions.
//
// This translation unit is compiled with -mllvm -amdgpu-mfma-vgpr-form=1:
// the int32 accumulators then stay in place in VGPRs (with AGPR accumulators
// the allocator copies accumulators through scratch AGPRs around the MFMAs
// once more than a few dozen of them are live).
//
// Slice representation (built by k_slice_tri in fgls_kernels.hip): row i of
// L^-1 times sc_i = 127 / (row maximum) is sum_s a_s R^(-s), R = 254,
// a_s in [-127, 127], stored in 1 KB blocks (fgls_a8_offset).  For integer
// genotypes x, (A_s x)_i is an exact int32 (|a x| <= 127*127 per term, exact
// for n <= 133,000), and z_i = sum_s R^(-s) (A_s x)_i / sc_i.  The
// five-digit form (k_slice_tri with TOP; TOP below) scales row i by its largest
// entry outside its diagonal 64-k step instead, keeps five digits, and adds
// a top digit t (weight R) on that step only: z_i = (R (T x)_i +
// sum_{s<5} R^(-s) (A_s x)_i) / sc_i.

#include "fgls_device.h"

// buffer_load_dwordx4 ... lds: size bytes per lane from rsrc at voffset +
// soffset + offset into LDS at lds + size * lane
__device__ void fgls_buffer_load_lds(fgls_v4i rsrc, __attribute__((address_space(3))) void* lds,
                                     int size, int voffset, int soffset, int offset,
                                     int aux) __asm("llvm.amdgcn.raw.buffer.load.lds");

namespace {

#define FGLS_I8_P_SWITCH(p, FN, ...)                                                         \
  switch (p) {                                                                               \
    case 1: FN<1>(__VA_ARGS__); break;  case 2: FN<2>(__VA_ARGS__); break;                   \
    case 3: FN<3>(__VA_ARGS__); break;  case 4: FN<4>(__VA_ARGS__); break;                   \
    case 5: FN<5>(__VA_ARGS__); break;  case 6: FN<6>(__VA_ARGS__); break;                   \
    case 7: FN<7>(__VA_ARGS__); break;  case 8: FN<8>(__VA_ARGS__); break;                   \
    case 9: FN<9>(__VA_ARGS__); break;  case 10: FN<10>(__VA_ARGS__); break;                 \
    case 11: FN<11>(__VA_ARGS__); break; case 12: FN<12>(__VA_ARGS__); break;                \
    case 13: FN<13>(__VA_ARGS__); break; case 14: FN<14>(__VA_ARGS__); break;                \
    case 15: FN<15>(__VA_ARGS__); break; case 16: FN<16>(__VA_ARGS__); break;                \
    default: break;                                                                          \
  }

#define FGLS_LDS(p) ((__attribute__((address_space(3))) void*)(p))

__device__ __forceinline__ fgls_v4i mfma_i8(fgls_v4i a, fgls_v4i b, fgls_v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// s_waitcnt immediate that waits until at most n vector-memory operations of
// this wave are outstanding (expcnt and lgkmcnt are not waited for).
constexpr int vmcnt_imm(int n) { return 0x0F70 | (n & 0xF) | ((n >> 4) << 14); }

// Workgroup: 4 waves, tile I8_TM = 32 rows of L^-1 x I8_TN = 256 SNP columns;
// wave w owns columns [64w, 64w + 64) and all 32 rows: 6 slices x 2 x 4 MFMA
// tiles of 16 x 16 int32 accumulators (192 VGPRs, two waves per SIMD).
//
// Per 64-k step the tile's 2 x 6 slice blocks (12 KB) go from HBM straight
// into LDS (buffer_load_dwordx4 ... lds: no staging registers; each wave
// moves three 1 KB blocks) into one half of a double-buffered LDS area; every
// wave then reads all twelve blocks as MFMA A operands.  The genotype
// fragments (B operands, 4 x 1 KB per wave) go straight to registers.
// The slice DMA is the MUBUF form on purpose: after the FLAT form
// (global_load_lds) the compiler cannot order the LDS counter and waits for
// every outstanding LDS read (lgkmcnt(0)) before each fragment is used; after
// the MUBUF form it waits for the one fragment the MFMAs need (lgkmcnt(1)),
// 13.2 -> 12.7 ms per 65,536 SNPs at n = 10^4 with bitwise equal results.
// The buffer resource of a workgroup covers its two 16-row slice groups (one
// contiguous range of 2 KS S KB, far below 2^31 for every n the int8 path
// accepts); the per-lane voffset picks the wave's block inside it and the
// uniform soffset the step, both within that range at every step.
//
// Step k (reading half k & 1): wait for this wave's slices of step k,
// barrier (every wave's slices of step k are in LDS and every wave has
// finished reading the other half in step k - 1), issue the slices and then
// the genotype loads of step k + 1, wait for this wave's genotype fragments
// of step k (issued after its slices in step k - 1, so the barrier never
// waits for genotype loads), then 12 LDS fragment reads and 48 MFMAs (six
// digits; five: 10 and 40, and 2 and 8 more on the diagonal step), each
// read one fragment ahead of the four MFMAs that use it (vmcnt retires in
// issue order: the 4 genotype loads follow the slices, and the S/2 + 4 loads
// of step k + 1 follow them).  The loop runs two steps per trip so that the
// LDS halves and the genotype registers alternate without copies (a copy of
// a register that a load is filling would wait for that load).
// Prefetch addresses are clamped to the last step, so every load is in
// bounds.  The barrier carries no memory fence: a workgroup-scope release
// fence would also wait for the LDS loads just issued; the vmcnt wait before
// it orders them instead.
//
// MFMA v_mfma_i32_16x16x64_i8: A lane l holds 16 consecutive k of row l%16
// (k chunk l/16 of the 64), B lane l the same 16 k of column l%16; D element
// v of lane l is D[4*(l/16)+v][l%16].
//
// Genotypes X8T are tiled like the slices (fgls_x8t_offset): one 1 KB block
// per (16 SNP columns, 64-k step) in B-operand lane order, so a wave's
// fragment is one contiguous 1 KB load (column-major columns would make it 16
// half cache lines ld apart); KSX = ld8 / 64 steps per column group, rows
// >= n meet all-zero slice columns.  fgls_launch_tile_i8 builds X8T from the
// column-major genotypes of the C-ABI.
// Slice blocks of one 64-k step in LDS: block ti * S + s holds slice s of
// the tile's 16-row group ti; with the top digit (TOP), blocks 2 S + ti hold
// the top digits of the diagonal step.
template <int S, bool TOP = false>
struct SliceStage {
  fgls_v4i blk[2 * S + (TOP ? 2 : 0)][64];
};

// Slice blocks a wave moves per step: blocks w, w + 4, w + 8, ... < 2 S.
constexpr int slice_dma_max(int S) { return (2 * S + 3) / 4; }

template <int S, bool TOP>
__device__ __forceinline__ void slice_dma(fgls_v4i rs, const int (&vo)[slice_dma_max(S)], int off,
                                          SliceStage<S, TOP>& dst, int w) {
#pragma unroll
  for (int j = 0; j < slice_dma_max(S); ++j)
    if ((2 * S) % 4 == 0 || w + 4 * j < 2 * S)
      fgls_buffer_load_lds(rs, FGLS_LDS(&dst.blk[w + 4 * j][0]), 16, vo[j], off, 0, 0);
}

// Genotype (B) fragments: 16 int8 values per lane (fgls_v4i) from 1 KB
// blocks, or, from 2-bit tiles (U2, hard calls 0..3), one dword per lane from
// 256-byte blocks, unpacked in registers: dword q of the int8 operand is
// (x >> 2 q) & 0x03030303 (the layout of fgls_x2t_bytes).
template <bool U2> struct BFrag { using T = fgls_v4i; static constexpr int BLOCK = 1024, LANE = 16; };
template <> struct BFrag<true> { using T = unsigned int; static constexpr int BLOCK = 256, LANE = 4; };

template <bool U2>
__device__ __forceinline__ typename BFrag<U2>::T load_b(const int8_t* p) {
  return *reinterpret_cast<const typename BFrag<U2>::T*>(p);
}
__device__ __forceinline__ fgls_v4i unpack_b(fgls_v4i x) { return x; }
__device__ __forceinline__ fgls_v4i unpack_b(unsigned int x) {
  constexpr unsigned int m = 0x03030303u;
  return fgls_v4i{(int)(x & m), (int)((x >> 2) & m), (int)((x >> 4) & m), (int)((x >> 6) & m)};
}

// One 64-k step: rd = LDS half of step k, wr = half that receives step k + 1,
// b = genotype fragments of step k, bn receives those of step k + 1.
//
// TOP (five digits, S = 5, and a top digit on the diagonal step): the ten
// slice blocks of a step are moved by waves 0, 1 (three each) and 2, 3 (two
// each); for the diagonal step (the row tile's last) waves 2 and 3 also move
// the top-digit block of 16-row group w - 2 (from tsrc at tvo).  DIAG marks
// that last step: nothing further is issued, and the products of its two
// extra fragments are added to digit 0's sums times R (exact in int32 for
// the n the host admits to this form).
template <int S, bool U2, bool TOP = false, bool DIAG = false>
__device__ __forceinline__ void step_i8(fgls_v4i (&acc)[S][2][4], const typename BFrag<U2>::T (&b)[4],
                                        typename BFrag<U2>::T (&bn)[4], const SliceStage<S, TOP>& rd,
                                        SliceStage<S, TOP>& wr, fgls_v4i asrc,
                                        const int (&avo)[slice_dma_max(S)], fgls_v4i tsrc, int tvo,
                                        const int8_t* Xw, const int (&boff)[4], int k, int last,
                                        int w, int lane) {
  static_assert(!TOP || S == 5, "the top-digit form has five digits below the top one");
  static_assert(TOP || !DIAG, "only the top-digit form has a diagonal step");
  // this wave's slices of step k are in LDS (its genotype loads of step k,
  // issued after them, may still be in flight)
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(4));
  const int k1 = min(k + 1, last);
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (!DIAG) {
    slice_dma<S, TOP>(asrc, avo, k1 * S * 1024, wr, w);
    if constexpr (TOP) {
      if (w >= 2 && k1 == last) fgls_buffer_load_lds(tsrc, FGLS_LDS(&wr.blk[2 * S + w - 2][0]), 16, tvo, 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) bn[tj] = load_b<U2>(Xw + BFrag<U2>::BLOCK * k1 + boff[tj]);
    __builtin_amdgcn_sched_barrier(0);
    // the genotype fragments of step k (older than the slice and 4 genotype
    // loads just issued)
    if constexpr (!TOP) {
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(slice_dma_max(S) + 4));
    } else {
      if (w < 2 || k1 == last) __builtin_amdgcn_s_waitcnt(vmcnt_imm(3 + 4));
      else __builtin_amdgcn_s_waitcnt(vmcnt_imm(2 + 4));
    }
  } else {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
  }
  __builtin_amdgcn_sched_barrier(0);
  fgls_v4i bu[4];
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) bu[tj] = unpack_b(b[tj]);
  // A fragments one ahead: fragment f = 2 s + ti (block ti S + s); on the
  // diagonal step fragments 2 S + ti are the top digits (block 2 S + ti)
  constexpr int NF = 2 * S + (DIAG ? 2 : 0);
  auto blk_of = [](int f) { return f < 2 * S ? (f & 1) * S + (f >> 1) : f; };
  fgls_v4i af[2];
  af[0] = rd.blk[0][lane];
  __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    if (f + 1 < NF) af[(f + 1) & 1] = rd.blk[blk_of(f + 1)][lane];
    if (f < 2 * S) {
      const int s2 = f >> 1, ti = f & 1;
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) acc[s2][ti][tj] = mfma_i8(af[f & 1], bu[tj], acc[s2][ti][tj]);
    } else {   // the top digits (one step): R times their products join digit 0's sums
      const int ti = f - 2 * S;
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
        acc[0][ti][tj] += FGLS_SLICE_RADIX * mfma_i8(af[f & 1], bu[tj], fgls_v4i{0, 0, 0, 0});
    }
  }
  // issue order: one LDS read (the next fragment), then the four MFMAs of
  // the current one
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    if (f + 1 < NF) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
  }
  __builtin_amdgcn_sched_barrier(0);
}

// DOS (dosages): every SNP has two int8 digit columns, lo in [0, 127] and hi
// in [-127, 127], with x * scale = 128 hi + lo; the lo plane holds columns
// 0..count-1 of X8T and the hi plane (tiled the same way) starts plane_off
// bytes later.  A wave then
// covers 32 SNPs: fragments 0, 1 read their lo digits and fragments 2, 3 the
// hi digits of the same SNPs (the same MFMA work per fragment as four
// genotype fragments), and the epilogue forms acc_lo + 128 acc_hi in FP64.
//
// EIG (the eigenbasis variant): A8 holds the slices of a full matrix (every
// row tile runs all K steps), and the epilogue reduces z against ntr groups
// of NQ weight columns (group k at W + k*NQ*ldw: column 0 weighs z*z,
// columns 1..NQ-1 weigh z), writing part[((rt*ntr + k)*NQ + q)*ldp + col].
//
// MODE (genotypes only): I8_RED the fused reductions above; I8_STORE writes
// z (column-major, ld ldz) instead, for the triangular A; I8_DOT, for a full
// A, reduces per SNP y'Zin[:, col] and y'W[:, k] (Zin column-major, ld ldz)
// into part as I8_RED does z'z and z'W[:, k].
enum { I8_RED = 0, I8_STORE = 1, I8_DOT = 2 };

// XCD group shape of each variant (row tiles x column tiles that one XCD
// takes together, fgls_tile_of_block): 16 x 8 on 2-bit tiles (1-1.5 %
// faster than 16 x 4 with five or six digits), 32 x 2 for the five-digit
// product on 8-bit tiles and dosage planes (4 % faster than 16 x 4: 32 row
// tiles share each genotype panel in the XCD's L2; 64 x 1, 48 x 1, 32 x 1
// and 16 x 2 measured 1-4 % slower; with six digits 16 x 4 is faster;
// -DFGLS_I8_GR8= -DFGLS_I8_GC8= set it at build time), I8_GR x I8_GC
// otherwise and for full matrices (FULL).
#ifndef FGLS_I8_GR8
#define FGLS_I8_GR8 32
#endif
#ifndef FGLS_I8_GC8
#define FGLS_I8_GC8 2
#endif
template <bool TOP, bool U2, bool FULL>
struct I8Group {
  static constexpr int R = FULL ? I8_GR : U2 ? 16 : TOP ? FGLS_I8_GR8 : I8_GR;
  static constexpr int C = FULL ? I8_GC : U2 ? 8 : TOP ? FGLS_I8_GC8 : I8_GC;
};
template <bool TOP, bool U2, bool FULL>
int i8_grid(int RT32, int CT) {
  return fgls_tile_grid(RT32, CT, I8Group<TOP, U2, FULL>::R, I8Group<TOP, U2, FULL>::C);
}

// TOP: A8 holds five digits per entry (fgls_a8_offset with S = 5) and A8top
// the top digits of every 16-row group's diagonal 64-k step (the group's
// last step: fgls_a8top_offset); triangular A only.
template <int NQ, int S, bool DOS = false, bool EIG = false, int MODE = I8_RED, bool U2 = false,
          bool TOP = false>
__global__ __launch_bounds__(256, 2) void k_trmm_i8(
    int n, int count, int RT32, int CT, int KS, const int8_t* __restrict__ A8,
    const double* __restrict__ inv_scale, const int8_t* __restrict__ X8T, int64_t ld8,
    const double* __restrict__ W, int64_t ldw, double* __restrict__ part, int64_t ldp,
    int64_t plane_off = 0, double inv_xscale = 1.0, int ntr = 1, double* __restrict__ Zio = nullptr,
    int64_t ldz = 0, const int8_t* __restrict__ A8top = nullptr) {
  static_assert(TOP ? S == 5 : S % 2 == 0, "the 2 x S slice blocks are spread over 4 waves");
  static_assert(!(TOP && (EIG || MODE == I8_DOT)), "the top digit is for the triangular L^-1");
  static_assert(!(U2 && DOS), "2-bit tiles hold genotypes, not dosage digit planes");
  using BT = typename BFrag<U2>::T;
  constexpr int BB = BFrag<U2>::BLOCK, BL = BFrag<U2>::LANE;
  constexpr int WN = DOS ? 32 : 64;   // SNP columns per wave
  __shared__ SliceStage<S, TOP> st0, st1;
  int rt, ct;
  using Grp = I8Group<TOP, U2, EIG || MODE == I8_DOT>;
  if (!fgls_tile_of_block<Grp::R, Grp::C>(blockIdx.x, RT32, CT, rt, ct)) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int r0 = rt * I8_TM;
  const int c0 = ct * (4 * WN) + w * WN;
  const int n64 = (n + 63) & ~63;
  constexpr bool FULL = EIG || MODE == I8_DOT;   // every row tile runs all K steps
  const int nsteps = (r0 < n64) ? (FULL ? n64 : min((r0 + I8_TM + 63) & ~63, n64)) / 64 : 0;

  // wave w moves blocks w, w + 4, w + 8 of each step: block = ti * S + s
  // buffer resource over the tile's two 16-row slice groups: 48-bit base,
  // stride 0, size in bytes, raw dword format (DATA_FORMAT 32)
  const size_t a0 = fgls_a8_offset(r0, 0, 0, S, KS);
  const uint64_t abase = (uint64_t)(A8 + a0);
  const fgls_v4i asrc = {(int)(uint32_t)abase, (int)(uint32_t)(abase >> 32) & 0xFFFF,
                         2 * KS * S * 1024, 0x00020000};
  int avo[slice_dma_max(S)];
#pragma unroll
  for (int j = 0; j < slice_dma_max(S); ++j) {
    const int blk = min(w + 4 * j, 2 * S - 1), ti = blk / S, s2 = blk % S;
    avo[j] = (int)(fgls_a8_offset(r0 + 16 * ti, s2, 0, S, KS) - a0) + 16 * lane;
  }
  // TOP: the two top-digit blocks of the tile (2 KB from its first 16-row
  // group's); waves 2 and 3 move group w - 2's
  fgls_v4i tsrc = {0, 0, 0, 0};
  int tvo = 0;
  if constexpr (TOP) {
    const uint64_t tbase = (uint64_t)(A8top + fgls_a8top_offset(r0, 0));
    tsrc = fgls_v4i{(int)(uint32_t)tbase, (int)(uint32_t)(tbase >> 32) & 0xFFFF, 2 * 1024, 0x00020000};
    tvo = max(w - 2, 0) * 1024 + 16 * lane;
  }
  // genotype fragments of this wave: Xw + boff[tj] (+ 1024 k) for fragment tj;
  // column groups past the last one re-read it (plane_off < 2^31 is checked
  // by the host)
  const int KSX = (int)(ld8 / 64);
  const int CG = (count + 15) / 16;
  const int cgb = min(c0 / 16, CG - 1);
  const int8_t* Xw = X8T + (size_t)cgb * KSX * BB;
  int boff[4];
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    const int cj = DOS ? (tj & 1) : tj;   // DOS: 0, 1 lo digits; 2, 3 hi digits
    boff[tj] = (min(c0 / 16 + cj, CG - 1) - cgb) * KSX * BB + BL * lane +
               ((DOS && tj >= 2) ? (int)plane_off : 0);
  }

  fgls_v4i acc[S][2][4];
#pragma unroll
  for (int s2 = 0; s2 < S; ++s2)
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) acc[s2][ti][tj] = fgls_v4i{0, 0, 0, 0};

  if (nsteps > 0) {
    const int last = nsteps - 1;
    BT b0[4], b1[4];
    if constexpr (TOP) {
      // the diagonal step always reads (b0, st0): with an odd number of steps
      // before it, step 0 goes to (b1, st1) and runs alone first
      auto prologue = [&](BT (&b)[4], SliceStage<S, TOP>& st) {
        slice_dma<S, TOP>(asrc, avo, 0, st, w);
        if (w >= 2 && last == 0) fgls_buffer_load_lds(tsrc, FGLS_LDS(&st.blk[2 * S + w - 2][0]), 16, tvo, 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) b[tj] = load_b<U2>(Xw + boff[tj]);
        __builtin_amdgcn_sched_barrier(0);
      };
      int k = 0;
      if (last & 1) {
        prologue(b1, st1);
        step_i8<S, U2, true>(acc, b1, b0, st1, st0, asrc, avo, tsrc, tvo, Xw, boff, 0, last, w, lane);
        k = 1;
      } else {
        prologue(b0, st0);
      }
      for (; k < last; k += 2) {
        step_i8<S, U2, true>(acc, b0, b1, st0, st1, asrc, avo, tsrc, tvo, Xw, boff, k, last, w, lane);
        step_i8<S, U2, true>(acc, b1, b0, st1, st0, asrc, avo, tsrc, tvo, Xw, boff, k + 1, last, w,
                             lane);
      }
      // the diagonal step waits for all of its loads: nothing is in flight after it
      step_i8<S, U2, true, true>(acc, b0, b1, st0, st1, asrc, avo, tsrc, tvo, Xw, boff, last, last, w,
                                 lane);
    } else {
      slice_dma<S, TOP>(asrc, avo, 0, st0, w);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) b0[tj] = load_b<U2>(Xw + boff[tj]);
      __builtin_amdgcn_sched_barrier(0);
      int k = 0;
      for (; k + 2 <= nsteps; k += 2) {
        step_i8<S, U2>(acc, b0, b1, st0, st1, asrc, avo, tsrc, tvo, Xw, boff, k, last, w, lane);
        step_i8<S, U2>(acc, b1, b0, st1, st0, asrc, avo, tsrc, tvo, Xw, boff, k + 1, last, w, lane);
      }
      if (k < nsteps)
        step_i8<S, U2>(acc, b0, b1, st0, st1, asrc, avo, tsrc, tvo, Xw, boff, k, last, w, lane);
      // no LDS load may still be in flight when the workgroup ends
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    }
  }

  if constexpr (MODE != I8_RED) {
    static_assert(!DOS && !EIG, "store / dot epilogues take genotypes");
    // z(row r0+16ti+4lg+v, SNP c0+16tj+li): each lane holds 4 consecutive
    // rows of a column per (ti, tj), moved as one 32-byte access
    double z[4][2][4];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double isc = inv_scale[r0 + 16 * ti + 4 * lg + v];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
          double t = (double)acc[S - 1][ti][tj][v];
#pragma unroll
          for (int s2 = S - 2; s2 >= 0; --s2)
            t = fma(t, 1.0 / FGLS_SLICE_RADIX, (double)acc[s2][ti][tj][v]);
          z[tj][ti][v] = t * isc;
        }
      }
    if constexpr (MODE == I8_STORE) {
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        const int col = c0 + 16 * tj + li;
        if (col < count) {
#pragma unroll
          for (int ti = 0; ti < 2; ++ti)
            *reinterpret_cast<fgls_v4d*>(Zio + (size_t)col * ldz + r0 + 16 * ti + 4 * lg) =
                fgls_v4d{z[tj][ti][0], z[tj][ti][1], z[tj][ti][2], z[tj][ti][3]};
        }
      }
    } else {
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        int oz;   // opaque zero: W is re-read per fragment instead of held live
        asm volatile("v_mov_b32 %0, 0" : "=v"(oz));
        const int col = c0 + 16 * tj + li;
        const double* zc = Zio + (size_t)min(col, count - 1) * ldz;
        double q[NQ];
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) q[qi] = 0.0;
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) {
          const fgls_v4d zin = *reinterpret_cast<const fgls_v4d*>(zc + r0 + 16 * ti + 4 * lg);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = r0 + 16 * ti + 4 * lg + v + oz;
            const double yy = z[tj][ti][v];
            q[0] = fma(yy, zin[v], q[0]);
#pragma unroll
            for (int k = 0; k < NQ - 1; ++k) q[1 + k] = fma(yy, W[(size_t)k * ldw + row], q[1 + k]);
          }
        }
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) {
          q[qi] += __shfl_xor(q[qi], 16, 64);
          q[qi] += __shfl_xor(q[qi], 32, 64);
        }
        if (lg == 0 && col < count) {
#pragma unroll
          for (int qi = 0; qi < NQ; ++qi) part[((size_t)rt * NQ + qi) * ldp + col] = q[qi];
        }
      }
    }
  } else if constexpr (EIG) {
    // z(row r0+16ti+4lg+v, SNP c0+16tj+li) for the NF fragments of the wave
    // (dosages: lo and hi digits combined), then one pass per weight group
    constexpr int NF = DOS ? 2 : 4;
    double z[NF][2][4];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double isc = inv_scale[r0 + 16 * ti + 4 * lg + v] * inv_xscale;
#pragma unroll
        for (int tj = 0; tj < NF; ++tj) {
          double t;
          if constexpr (DOS) {
            t = fma(128.0, (double)acc[S - 1][ti][tj + 2][v], (double)acc[S - 1][ti][tj][v]);
#pragma unroll
            for (int s2 = S - 2; s2 >= 0; --s2)
              t = fma(t, 1.0 / FGLS_SLICE_RADIX,
                      fma(128.0, (double)acc[s2][ti][tj + 2][v], (double)acc[s2][ti][tj][v]));
          } else {
            t = (double)acc[S - 1][ti][tj][v];
#pragma unroll
            for (int s2 = S - 2; s2 >= 0; --s2)
              t = fma(t, 1.0 / FGLS_SLICE_RADIX, (double)acc[s2][ti][tj][v]);
          }
          z[tj][ti][v] = t * isc;
        }
      }
    // group k's weights of the tile's 32 rows go through LDS (one load per
    // thread and group) instead of NQ dependent global loads per row and lane
    __shared__ double wsh[NQ][I8_TM];
    for (int k = 0; k < ntr; ++k) {
      const double* Wk = W + (size_t)k * NQ * ldw;
      __syncthreads();   // every wave is done with group k - 1
      for (int idx = tid; idx < NQ * I8_TM; idx += 256)
        wsh[idx / I8_TM][idx % I8_TM] = Wk[(size_t)(idx / I8_TM) * ldw + r0 + idx % I8_TM];
      __syncthreads();
#pragma unroll
      for (int tj = 0; tj < NF; ++tj) {
        double q[NQ];
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) q[qi] = 0.0;
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rl = 16 * ti + 4 * lg + v;   // row r0 + rl
            const double zz = z[tj][ti][v];
            q[0] = fma(zz * wsh[0][rl], zz, q[0]);
#pragma unroll
            for (int j = 1; j < NQ; ++j) q[j] = fma(zz, wsh[j][rl], q[j]);
          }
        }
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) {
          q[qi] += __shfl_xor(q[qi], 16, 64);
          q[qi] += __shfl_xor(q[qi], 32, 64);
        }
        const int col = c0 + 16 * tj + li;
        if (lg == 0 && col < count) {
#pragma unroll
          for (int qi = 0; qi < NQ; ++qi)
            part[(((size_t)rt * ntr + k) * NQ + qi) * ldp + col] = q[qi];
        }
      }
    }
  } else if constexpr (DOS) {
    // z(row r0+16ti+4lg+v, SNP c0+16tj+li), tj < 2:
    //   inv_scale[row] / scale * sum_s R^(-s) (acc[s][tj] + 128 acc[s][tj+2])
    double z[2][2][4];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double isc = inv_scale[r0 + 16 * ti + 4 * lg + v] * inv_xscale;
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
          double t = fma(128.0, (double)acc[S - 1][ti][tj + 2][v], (double)acc[S - 1][ti][tj][v]);
#pragma unroll
          for (int s2 = S - 2; s2 >= 0; --s2)
            t = fma(t, 1.0 / FGLS_SLICE_RADIX,
                    fma(128.0, (double)acc[s2][ti][tj + 2][v], (double)acc[s2][ti][tj][v]));
          z[tj][ti][v] = t * isc;
        }
      }
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      int oz;
      asm volatile("v_mov_b32 %0, 0" : "=v"(oz));
      double q[NQ];
#pragma unroll
      for (int qi = 0; qi < NQ; ++qi) q[qi] = 0.0;
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + 16 * ti + 4 * lg + v + oz;
          const double zz = z[tj][ti][v];
          q[0] = fma(zz, zz, q[0]);
#pragma unroll
          for (int k = 0; k < NQ - 1; ++k) q[1 + k] = fma(zz, W[(size_t)k * ldw + row], q[1 + k]);
        }
      }
#pragma unroll
      for (int qi = 0; qi < NQ; ++qi) {
        q[qi] += __shfl_xor(q[qi], 16, 64);
        q[qi] += __shfl_xor(q[qi], 32, 64);
      }
      const int col = c0 + 16 * tj + li;
      if (lg == 0 && col < count) {
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) part[((size_t)rt * NQ + qi) * ldp + col] = q[qi];
      }
    }
  } else if constexpr (NQ <= 5) {   // p <= 4: sums interleaved with the conversion (fastest here)
    // z(row r0+16ti+4lg+v, col c0+16tj+li) = inv_scale[row] * sum_s R^(-s) acc[s];
    // one 16-column fragment at a time, so that the accumulators of finished
    // fragments are free and the FP64 sums stay within the register budget.
    // The tile's 32 row scales and weight rows go through LDS once (one load
    // per thread) instead of 40 dependent global loads per fragment and lane.
    __shared__ double wsh[NQ][I8_TM];
    for (int idx = tid; idx < NQ * I8_TM; idx += 256) {
      const int qi = idx / I8_TM, r = idx % I8_TM;
      wsh[qi][r] = (qi == 0) ? inv_scale[r0 + r] : W[(size_t)(qi - 1) * ldw + r0 + r];
    }
    __syncthreads();
  #pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
      double q[NQ];
  #pragma unroll
      for (int qi = 0; qi < NQ; ++qi) q[qi] = 0.0;
  #pragma unroll
      for (int ti = 0; ti < 2; ++ti) {
  #pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rl = 16 * ti + 4 * lg + v;   // row r0 + rl
          double t = (double)acc[S - 1][ti][tj][v];
  #pragma unroll
          for (int s2 = S - 2; s2 >= 0; --s2)
            t = fma(t, 1.0 / FGLS_SLICE_RADIX, (double)acc[s2][ti][tj][v]);
          const double z = t * wsh[0][rl];
          q[0] = fma(z, z, q[0]);
  #pragma unroll
          for (int k = 0; k < NQ - 1; ++k) q[1 + k] = fma(z, wsh[1 + k][rl], q[1 + k]);
        }
      }
  #pragma unroll
      for (int qi = 0; qi < NQ; ++qi) {
        q[qi] += __shfl_xor(q[qi], 16, 64);
        q[qi] += __shfl_xor(q[qi], 32, 64);
      }
      const int col = c0 + 16 * tj + li;
      if (lg == 0 && col < count) {
  #pragma unroll
        for (int qi = 0; qi < NQ; ++qi) part[((size_t)rt * NQ + qi) * ldp + col] = q[qi];
      }
    }
  } else {
    // z(row r0+16ti+4lg+v, col c0+16tj+li) = inv_scale[row] * sum_s R^(-s) acc[s].
    // All 32 z of a lane are formed first (the int32 accumulators die as they
    // are read: 192 registers become 64), then the per-column sums; with the
    // sums interleaved with the conversion, p >= 8 spilled inside the loop.
    double z[4][2][4];
  #pragma unroll
    for (int ti = 0; ti < 2; ++ti)
  #pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double isc = inv_scale[r0 + 16 * ti + 4 * lg + v];
  #pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
          double t = (double)acc[S - 1][ti][tj][v];
  #pragma unroll
          for (int s2 = S - 2; s2 >= 0; --s2)
            t = fma(t, 1.0 / FGLS_SLICE_RADIX, (double)acc[s2][ti][tj][v]);
          z[tj][ti][v] = t * isc;
        }
      }
  #pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
      // W is re-read (from cache) for every fragment: an opaque zero in the
      // row index keeps the compiler from holding 8 (p + 1) values of W live
      // across the four fragments
      int oz;
      asm volatile("v_mov_b32 %0, 0" : "=v"(oz));
      double q[NQ];
  #pragma unroll
      for (int qi = 0; qi < NQ; ++qi) q[qi] = 0.0;
  #pragma unroll
      for (int ti = 0; ti < 2; ++ti) {
  #pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + 16 * ti + 4 * lg + v + oz;
          const double zz = z[tj][ti][v];
          q[0] = fma(zz, zz, q[0]);
  #pragma unroll
          for (int k = 0; k < NQ - 1; ++k) q[1 + k] = fma(zz, W[(size_t)k * ldw + row], q[1 + k]);
        }
      }
  #pragma unroll
      for (int qi = 0; qi < NQ; ++qi) {
        q[qi] += __shfl_xor(q[qi], 16, 64);
        q[qi] += __shfl_xor(q[qi], 32, 64);
      }
      const int col = c0 + 16 * tj + li;
      if (lg == 0 && col < count) {
  #pragma unroll
        for (int qi = 0; qi < NQ; ++qi) part[((size_t)rt * NQ + qi) * ldp + col] = q[qi];
      }
    }
  }
}

// Column-major genotypes (ld ldx8; `planes` planes of count columns, plane q
// at X8 + q count ldx8) -> tiled X8T (plane q at X8T + q CG KSX 1024, CG =
// ceil(count/16), KSX = ld8/64): one workgroup per (column group, plane); its
// output is contiguous, each thread writes one 16-byte lane chunk.  Columns
// past count repeat the last column.
__global__ __launch_bounds__(256) void k_tile_i8(int count, int KSX, const int8_t* __restrict__ X8,
                                                 int64_t ldx8, int8_t* __restrict__ X8T) {
  const int cg = blockIdx.x, pl = blockIdx.y, CG = gridDim.x;
  const int8_t* src = X8 + (size_t)pl * count * ldx8;
  int8_t* dst = X8T + ((size_t)pl * CG + cg) * KSX * 1024;
  for (int t = threadIdx.x; t < KSX * 64; t += 256) {
    const int ks = t >> 6, g = (t >> 4) & 3, i = t & 15;
    const int col = min(cg * 16 + i, count - 1);
    *reinterpret_cast<fgls_v4i*>(dst + (size_t)t * 16) =
        *reinterpret_cast<const fgls_v4i*>(src + (size_t)col * ldx8 + 64 * ks + 16 * g);
  }
}

// Column-major int8 genotypes -> 2-bit tiles (fgls_x2t_bytes): one workgroup
// per 16-column group, thread t = 64 ks + 16 g + i packs the 16 genotypes
// k = 64 ks + 16 g + j of column 16 cg + i into one dword (genotype j = 4 q + b
// at bit 8 b + 2 q).  Values of rows < n outside 0..3 are counted in *bad
// and written as 0; rows >= n are written as 0; columns past count repeat
// the last column.
__global__ __launch_bounds__(256) void k_tile_u2(int n, int count, int KSX, const int8_t* __restrict__ X8,
                                                 int64_t ldx8, unsigned int* __restrict__ X2T,
                                                 unsigned int* __restrict__ bad) {
  const int cg = blockIdx.x;
  unsigned int* dst = X2T + (size_t)cg * KSX * 64;
  unsigned int nb = 0;
  for (int t = threadIdx.x; t < KSX * 64; t += 256) {
    const int ks = t >> 6, g = (t >> 4) & 3, i = t & 15;
    const int col = min(cg * 16 + i, count - 1);
    const int r0 = 64 * ks + 16 * g;
    const fgls_v4i v = *reinterpret_cast<const fgls_v4i*>(X8 + (size_t)col * ldx8 + r0);
    unsigned int x = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int gv = (int)(signed char)((v[q] >> (8 * b)) & 0xff);
        unsigned int c = 0;
        if (r0 + 4 * q + b < n) {
          if (gv < 0 || gv > 3) ++nb;
          else c = (unsigned int)gv;
        }
        x |= c << (8 * b + 2 * q);
      }
    dst[t] = x;
  }
  if (nb) atomicAdd(bad, nb);
}

// 2-bit genotype file columns (ceil(n/4) bytes per column at P + c ldp; the
// file's genotype r in bits 2 (r % 4) of byte r / 4, so the 16 genotypes
// j = 0..15 of a dword sit at bits 2 j) -> 2-bit tiles, whose lane dword
// holds genotype j at bit 8 (j % 4) + 2 (j / 4): a transpose of the dword's
// 4 x 4 two-bit fields.  One workgroup per 16-column group and 16 steps;
// thread t takes column t % 16 and step 16 blockIdx.y + t / 16: it reads the
// column's 16 bytes of the step as 4 dwords (bytes when the columns are not
// 4-byte aligned or at the end of a column) and writes lanes 16 g + t % 16
// of the step's block, so the 16 threads of a column group write 64
// contiguous bytes.  Code 3 in rows < n is counted in *bad and written as 0;
// rows >= n and bytes past the column are 0; columns past count repeat the
// last column.
__global__ __launch_bounds__(256) void k_tile_u2_packed(int n, int count, int KSX, const uint8_t* __restrict__ P,
                                                        int64_t ldp, unsigned int* __restrict__ X2T,
                                                        unsigned int* __restrict__ bad) {
  const int cg = blockIdx.x;
  const int nbytes = (n + 3) / 4;
  const bool al4 = ((ldp & 3) == 0) && ((reinterpret_cast<uintptr_t>(P) & 3) == 0);
  const int c = threadIdx.x & 15;
  const int col = min(cg * 16 + c, count - 1);
  const uint8_t* src = P + (size_t)col * ldp;
  unsigned int* dst = X2T + (size_t)cg * KSX * 64;
  unsigned int nb = 0;
  const int ks = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (ks < KSX) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int bi = 16 * ks + 4 * g;   // bytes of genotypes r = 64 ks + 16 g + 0..15
      unsigned int x = 0;
      if (al4 && bi + 3 < nbytes) {
        x = *reinterpret_cast<const unsigned int*>(src + bi);
      } else {
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (bi + m < nbytes) x |= (unsigned int)src[bi + m] << (8 * m);
      }
      // fields of rows >= n cleared, code 3 counted and cleared
      const int valid = n - 4 * bi;
      if (valid < 16) x &= valid <= 0 ? 0u : (1u << (2 * valid)) - 1u;
      const unsigned int three = x & (x >> 1) & 0x55555555u;
      nb += __builtin_popcount(three);
      x &= ~(three | (three << 1));
      // transpose the 4 x 4 two-bit fields (field (m, q) at bit 8 m + 2 q
      // goes to bit 8 q + 2 m): swap the off-diagonal 2 x 2 blocks, then
      // the off-diagonal fields inside each block
      unsigned int t = ((x >> 12) ^ x) & 0x0000F0F0u;
      x ^= t ^ (t << 12);
      t = ((x >> 6) ^ x) & 0x00CC00CCu;
      x ^= t ^ (t << 6);
      dst[(size_t)ks * 64 + 16 * g + c] = x;
    }
  }
  if (nb) atomicAdd(bad, nb);
}

template <int P, bool EIG>
void launch(int n, int count, int RT32, int KS, const int8_t* A8, const double* isc,
            const int8_t* X8, int64_t ldx8, const double* W, int64_t ldw, double* part, int64_t ldp,
            int64_t plane_off, double inv_xscale, int ntr, int nslice, hipStream_t st) {
  constexpr int T = FGLS_TOP_NSLICE;
  const int8_t* top = A8 + fgls_a8top_base(KS);
  if (plane_off > 0) {
    const int CT = (count + 127) / 128;
    if (!EIG && nslice == T)
      hipLaunchKernelGGL((k_trmm_i8<P + 1, T, true, false, I8_RED, false, true>), dim3(i8_grid<true, false, false>(RT32, CT)), dim3(256), 0, st, n,
                         count, RT32, CT, KS, A8, isc, X8, ldx8, W, ldw, part, ldp, plane_off, inv_xscale, 1,
                         (double*)nullptr, (int64_t)0, top);
    else
      hipLaunchKernelGGL((k_trmm_i8<P + 1, FGLS_NSLICE, true, EIG>), dim3(i8_grid<false, false, EIG>(RT32, CT)),
                         dim3(256), 0, st, n, count, RT32, CT, KS, A8, isc, X8, ldx8, W, ldw, part, ldp, plane_off,
                         inv_xscale, ntr);
  } else {
    const int CT = (count + I8_TN - 1) / I8_TN;
    if (!EIG && nslice == T)
      hipLaunchKernelGGL((k_trmm_i8<P + 1, T, false, false, I8_RED, false, true>), dim3(i8_grid<true, false, false>(RT32, CT)), dim3(256), 0, st,
                         n, count, RT32, CT, KS, A8, isc, X8, ldx8, W, ldw, part, ldp, (int64_t)0, 1.0, 1,
                         (double*)nullptr, (int64_t)0, top);
    else
      hipLaunchKernelGGL((k_trmm_i8<P + 1, FGLS_NSLICE, false, EIG>), dim3(i8_grid<false, false, EIG>(RT32, CT)),
                         dim3(256), 0, st, n, count, RT32, CT, KS, A8, isc, X8, ldx8, W, ldw, part, ldp, (int64_t)0,
                         1.0, ntr);
  }
}

template <int P>
void launch_any(int n, int count, int RT32, int KS, const int8_t* A8, const double* isc,
                const int8_t* X8, int64_t ldx8, const double* W, int64_t ldw, double* part,
                int64_t ldp, int64_t plane_off, double inv_xscale, int eig_ntr, int nslice, hipStream_t st) {
  if (eig_ntr > 0)
    launch<P, true>(n, count, RT32, KS, A8, isc, X8, ldx8, W, ldw, part, ldp, plane_off,
                    inv_xscale, eig_ntr, FGLS_NSLICE, st);
  else
    launch<P, false>(n, count, RT32, KS, A8, isc, X8, ldx8, W, ldw, part, ldp, plane_off,
                     inv_xscale, 1, nslice, st);
}

}  // namespace

namespace {

template <int P>
void launch_u2(int n, int count, int RT32, int KS, const int8_t* A8, const double* isc, const uint8_t* X2T,
               const double* W, int64_t ldw, double* part, int64_t ldp, int eig_ntr, int nslice,
               hipStream_t st) {
  const int CT = (count + I8_TN - 1) / I8_TN;
  const int64_t ld8 = ((int64_t)n + 63) / 64 * 64;
  if (eig_ntr > 0)
    hipLaunchKernelGGL((k_trmm_i8<P + 1, FGLS_NSLICE, false, true, I8_RED, true>), dim3(i8_grid<false, true, true>(RT32, CT)), dim3(256), 0, st,
                       n, count, RT32, CT, KS, A8, isc, (const int8_t*)X2T, ld8, W, ldw, part, ldp, (int64_t)0,
                       1.0, eig_ntr);
  else if (nslice == FGLS_TOP_NSLICE)
    hipLaunchKernelGGL((k_trmm_i8<P + 1, FGLS_TOP_NSLICE, false, false, I8_RED, true, true>),
                       dim3(i8_grid<true, true, false>(RT32, CT)), dim3(256), 0, st, n, count, RT32, CT, KS, A8,
                       isc, (const int8_t*)X2T, ld8, W, ldw, part,
                       ldp, (int64_t)0, 1.0, 1, (double*)nullptr, (int64_t)0, A8 + fgls_a8top_base(KS));
  else
    hipLaunchKernelGGL((k_trmm_i8<P + 1, FGLS_NSLICE, false, false, I8_RED, true>),
                       dim3(i8_grid<false, true, false>(RT32, CT)), dim3(256), 0, st, n, count, RT32, CT, KS, A8,
                       isc, (const int8_t*)X2T, ld8, W, ldw, part, ldp, (int64_t)0, 1.0, 1);
}

template <int P>
void launch_dot(int n, int count, int RT32, int KS, const int8_t* A8, const double* isc,
                const int8_t* X8T, int64_t ld8, const double* W, int64_t ldw, const double* Zin,
                int64_t ldz, double* part, int64_t ldp, hipStream_t st) {
  const int CT = (count + I8_TN - 1) / I8_TN;
  hipLaunchKernelGGL((k_trmm_i8<P + 1, FGLS_NSLICE, false, false, I8_DOT>), dim3(i8_grid<false, false, true>(RT32, CT)), dim3(256), 0, st,
                     n, count, RT32, CT, KS, A8, isc, X8T, ld8, W, ldw, part, ldp, (int64_t)0, 1.0, 1,
                     const_cast<double*>(Zin), ldz);
}

template <bool U2>
void launch_store(int n, int count, int RT32, int KS, const int8_t* A8, const double* inv_scale,
                  const int8_t* X, double* Z, int64_t ldz, int nslice, hipStream_t st) {
  const int CT = (count + I8_TN - 1) / I8_TN;
  const int64_t ld8 = ((int64_t)n + 63) / 64 * 64;
  if (nslice == FGLS_TOP_NSLICE)
    hipLaunchKernelGGL((k_trmm_i8<2, FGLS_TOP_NSLICE, false, false, I8_STORE, U2, true>),
                       dim3(i8_grid<true, U2, false>(RT32, CT)), dim3(256), 0,
                       st, n, count, RT32, CT, KS, A8, inv_scale, X, ld8, (const double*)nullptr, (int64_t)0,
                       (double*)nullptr, (int64_t)0, (int64_t)0, 1.0, 1, Z, ldz, A8 + fgls_a8top_base(KS));
  else
    hipLaunchKernelGGL((k_trmm_i8<2, FGLS_NSLICE, false, false, I8_STORE, U2>), dim3(i8_grid<false, U2, false>(RT32, CT)),
                       dim3(256), 0, st, n,
                       count, RT32, CT, KS, A8, inv_scale, X, ld8, (const double*)nullptr, (int64_t)0,
                       (double*)nullptr, (int64_t)0, (int64_t)0, 1.0, 1, Z, ldz);
}

}  // namespace

void fgls_launch_i8_store(int n, int count, int RT32, int KS, const int8_t* A8, const double* inv_scale,
                          const int8_t* X8T, int64_t ld8, double* Z, int64_t ldz, int nslice, hipStream_t st) {
  launch_store<false>(n, count, RT32, KS, A8, inv_scale, X8T, Z, ldz, nslice, st);
}

void fgls_launch_u2_store(int n, int count, int RT32, int KS, const int8_t* A8, const double* inv_scale,
                          const uint8_t* X2T, double* Z, int64_t ldz, int nslice, hipStream_t st) {
  launch_store<true>(n, count, RT32, KS, A8, inv_scale, (const int8_t*)X2T, Z, ldz, nslice, st);
}

void fgls_launch_i8_dot(int p, int n, int count, int RT32, int KS, const int8_t* A8,
                        const double* inv_scale, const int8_t* X8T, int64_t ld8, const double* W,
                        int64_t ldw, const double* Zin, int64_t ldz, double* part, int64_t ldp,
                        hipStream_t st) {
  FGLS_I8_P_SWITCH(p, launch_dot, n, count, RT32, KS, A8, inv_scale, X8T, ld8, W, ldw, Zin, ldz, part,
                   ldp, st);
}

void fgls_launch_trmm_u2(int p, int n, int count, int RT32, int KS, const int8_t* A8,
                         const double* inv_scale, const uint8_t* X2T, const double* W, int64_t ldw,
                         double* part, int64_t ldp, int eig_ntr, int nslice, hipStream_t st) {
  FGLS_I8_P_SWITCH(p, launch_u2, n, count, RT32, KS, A8, inv_scale, X2T, W, ldw, part, ldp, eig_ntr, nslice,
                   st);
}

void fgls_launch_tile_u2(int n, int count, const int8_t* X8, int64_t ldx8, uint8_t* X2T, unsigned int* bad,
                         hipStream_t st) {
  const int KSX = (n + 63) / 64;
  hipLaunchKernelGGL(k_tile_u2, dim3((count + 15) / 16), dim3(256), 0, st, n, count, KSX, X8, ldx8,
                     reinterpret_cast<unsigned int*>(X2T), bad);
}

void fgls_launch_tile_u2_packed(int n, int count, const uint8_t* P, int64_t ldp, uint8_t* X2T,
                                unsigned int* bad, hipStream_t st) {
  const int KSX = (n + 63) / 64;
  hipLaunchKernelGGL(k_tile_u2_packed, dim3((count + 15) / 16, (KSX + 15) / 16), dim3(256), 0, st, n, count, KSX, P, ldp,
                     reinterpret_cast<unsigned int*>(X2T), bad);
}

void fgls_launch_tile_i8(int n, int count, int planes, const int8_t* X8, int64_t ldx8, int8_t* X8T,
                         hipStream_t st) {
  const int KSX = (n + 63) / 64;
  hipLaunchKernelGGL(k_tile_i8, dim3((count + 15) / 16, planes), dim3(256), 0, st, count, KSX, X8,
                     ldx8, X8T);
}

void fgls_launch_trmm_i8(int p, int n, int count, int RT32, int KS, const int8_t* A8,
                         const double* inv_scale, const int8_t* X8, int64_t ldx8, const double* W,
                         int64_t ldw, double* part, int64_t ldp, int64_t plane_off,
                         double inv_xscale, int eig_ntr, int nslice, hipStream_t st) {
#define FGLS_I8_CASE(P)                                                                      \
  case P:                                                                                    \
    launch_any<P>(n, count, RT32, KS, A8, inv_scale, X8, ldx8, W, ldw, part, ldp, plane_off, \
                  inv_xscale, eig_ntr, nslice, st);                                          \
    break;
  switch (p) {
    FGLS_I8_CASE(1) FGLS_I8_CASE(2) FGLS_I8_CASE(3) FGLS_I8_CASE(4) FGLS_I8_CASE(5)
    FGLS_I8_CASE(6) FGLS_I8_CASE(7) FGLS_I8_CASE(8) FGLS_I8_CASE(9) FGLS_I8_CASE(10)
    FGLS_I8_CASE(11) FGLS_I8_CASE(12) FGLS_I8_CASE(13) FGLS_I8_CASE(14) FGLS_I8_CASE(15)
    FGLS_I8_CASE(16)
    default: break;
  }
#undef FGLS_I8_CASE
}
