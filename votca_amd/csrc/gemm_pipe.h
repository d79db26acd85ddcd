// Split-BF16 ("bf16x3") GEMM for gfx950 with a multi-stage direct-to-LDS pipeline.
//
//   C(m,n) = epilogue( sum_k (Ah + Al)(m,k) (Bh + Bl)(n,k) ),  products hi*hi + hi*lo + lo*hi
//
// Both operands are K-contiguous bf16 planes (hi and lo) with 16-byte aligned rows.
// Block tile BM x BN x 32, 256 threads = 2 x 2 waves, wave tile (BM/2) x (BN/2) built from
// v_mfma_f32_32x32x16_bf16 tiles with FP32 accumulators.
//
// Operand staging: global_load_lds_dwordx4 moves 16 bytes per lane straight from HBM/L2 into
// LDS (lane l lands at base + 16 l), so no VGPRs hold in-flight tiles and STAGES k-steps can be
// in flight at once. Each 1 KB LDS fragment is stored in MFMA operand order (lane l <-> row
// l % 32, k-chunk l / 32), which makes every operand read a linear, conflict-free ds_read_b128.
// Rows past M/N and k-chunks past the K range are redirected to a zero page, so the tile edges
// need no masking in the MFMA loop and never read outside the operands.
#pragma once
#include "gemm_kernels.h"

namespace xg {

typedef float f16v __attribute__((ext_vector_type(16)));

// 16-byte-aligned zeros that out-of-range lanes load from (never written)
__device__ __attribute__((aligned(64))) unsigned int g_zero_page[16];

typedef __attribute__((address_space(3))) uint4 lds_u4;

// 16 bytes per lane from global memory into LDS at (uniform) byte offset `lds_off` + 16 * lane.
// Issued as inline assembly so that the compiler's wait-count pass does not tie every later LDS
// read to these loads; completion is tracked explicitly with s_waitcnt vmcnt in the main loop.
__device__ __forceinline__ void dma16(const void* g, unsigned lds_off) {
  const unsigned o = __builtin_amdgcn_readfirstlane(lds_off);
  // one wait state between the M0 write and the LDS-DMA load that reads it
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(o), "v"(g) : "memory", "m0");
}

template <int BM, int BN, int STAGES, class Epi>
__global__ __launch_bounds__(256) void gemm_bf16x3_pipe_kernel(GemmDims d, Opnd<unsigned short> A,
                                                               Opnd<unsigned short> B, Epi epi) {
  constexpr int BK = 32;
  constexpr int FA = BM / 32 * 2;        // 1 KB fragments per A plane per stage (row groups x 2 k-halves)
  constexpr int FB = BN / 32 * 2;
  constexpr int FST = 2 * FA + 2 * FB;   // fragments per stage
  constexpr int PA = FA / 4, PB = FB / 4;  // fragments per wave per plane
  constexpr int LPW = 2 * PA + 2 * PB;     // direct loads per wave per stage
  constexpr int TM = BM / 64, TN = BN / 64;
  static_assert(FA % 4 == 0 && FB % 4 == 0, "BM and BN must be multiples of 64");
  static_assert(STAGES >= 2 && (STAGES - 2) * LPW <= 63, "vmcnt range");
  __shared__ __attribute__((aligned(1024))) uint4 lds[STAGES * FST * 64];

  int tm, tn;
  tile_of_block(d, tm, tn);
  const int b = blockIdx.z, ks = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = ks * d.kchunk;
  const int kend = min(d.K, kbeg + d.kchunk);
  const int nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;

  // per-lane row pointers (k offset of this lane's chunk included) of the fragments this wave loads
  const unsigned short* pa[2][PA];
  const unsigned short* pb[2][PB];
  const int kl = 8 * (lane >> 5);
#pragma unroll
  for (int i = 0; i < PA; ++i) {
    const int f = wave * PA + i, rg = f >> 1, sub = f & 1;
    const int row = m0 + rg * 32 + (lane & 31);
    const long off = (long)b * A.bs + (long)row * A.ld + sub * 16 + kl;
    pa[0][i] = row < d.M ? A.p + off : nullptr;
    pa[1][i] = row < d.M ? A.lo + off : nullptr;
  }
#pragma unroll
  for (int i = 0; i < PB; ++i) {
    const int f = wave * PB + i, rg = f >> 1, sub = f & 1;
    const int row = n0 + rg * 32 + (lane & 31);
    const long off = (long)b * B.bs + (long)row * B.ld + sub * 16 + kl;
    pb[0][i] = row < d.N ? B.p + off : nullptr;
    pb[1][i] = row < d.N ? B.lo + off : nullptr;
  }
  const unsigned short* zp = reinterpret_cast<const unsigned short*>(g_zero_page);

  const unsigned lds0 = (unsigned)(uintptr_t)(lds_u4*)lds;
  auto issue = [&](int stage, int k0) {
    const unsigned base = lds0 + (unsigned)(stage * FST * 1024);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        const int f = wave * PA + i;
        const int kc = k0 + (f & 1) * 16 + kl;
        const unsigned short* g = (pa[h][i] && kc < kend) ? pa[h][i] + k0 : zp;
        dma16(g, base + (unsigned)((h * FA + f) * 1024));
      }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        const int f = wave * PB + i;
        const int kc = k0 + (f & 1) * 16 + kl;
        const unsigned short* g = (pb[h][i] && kc < kend) ? pb[h][i] + k0 : zp;
        dma16(g, base + (unsigned)((2 * FA + h * FB + f) * 1024));
      }
  };

  f16v acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

#pragma unroll
  for (int s = 0; s < STAGES - 1; ++s)
    if (s < nk) issue(s, kbeg + s * BK);

  for (int kt = 0; kt < nk; ++kt) {
    // the loads of step kt are complete once at most the later steps' loads are outstanding
    if (kt + STAGES - 2 < nk)
      asm volatile("s_waitcnt vmcnt(%0)" ::"n"((STAGES - 2) * LPW) : "memory");
    else
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // every wave's fragments of step kt are in LDS, and every wave has finished reading step kt-1,
    // whose buffer the next issue overwrites
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    if (kt + STAGES - 1 < nk) issue((kt + STAGES - 1) % STAGES, kbeg + (kt + STAGES - 1) * BK);
    const uint4* sAh = lds + (kt % STAGES) * FST * 64;
    const uint4* sAl = sAh + FA * 64;
    const uint4* sBh = sAl + FA * 64;
    const uint4* sBl = sBh + FB * 64;
#pragma unroll
    for (int sub = 0; sub < 2; ++sub) {
      bf8v bh[TN], bl[TN];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int f = ((wn * TN + j) * 2 + sub) * 64 + lane;
        uint4 vh = sBh[f], vl = sBl[f];
        bh[j] = __builtin_bit_cast(bf8v, vh);
        bl[j] = __builtin_bit_cast(bf8v, vl);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int f = ((wm * TM + i) * 2 + sub) * 64 + lane;
        uint4 vh = sAh[f], vl = sAl[f];
        const bf8v ah = __builtin_bit_cast(bf8v, vh);
        const bf8v al = __builtin_bit_cast(bf8v, vl);
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[j], acc[i][j], 0, 0, 0);
      }
    }
  }
  // D layout of the 32x32 MFMA: lane l holds column l % 32, register r holds row
  // 8 (r / 4) + 4 (l / 32) + r % 4
  epilogue_store<TM, TN, 16>(
      d, epi, b, ks, acc,
      [&](int i, int r) { return m0 + (wm * TM + i) * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); },
      [&](int j) { return n0 + (wn * TN + j) * 32 + (lane & 31); });
}

// Pipelined split-BF16 GEMM is usable when both operands are K-contiguous split planes with
// 16-byte aligned rows (ld, batch stride and K multiples of 8 bf16).
inline bool pipe_ok(const Opnd<unsigned short>& A, const Opnd<unsigned short>& B, int K) {
  auto al = [](const void* p) { return p != nullptr && (reinterpret_cast<uintptr_t>(p) % 16) == 0; };
  return al(A.p) && al(A.lo) && al(B.p) && al(B.lo) && A.ld % 8 == 0 && B.ld % 8 == 0 && A.bs % 8 == 0 &&
         B.bs % 8 == 0 && K % 8 == 0;
}

template <int BM, int BN, int STAGES, class Epi>
inline hipError_t launch_gemm_bf16x3_pipe(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                          Opnd<unsigned short> B, Epi epi) {
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  hipLaunchKernelGGL((gemm_bf16x3_pipe_kernel<BM, BN, STAGES, Epi>), grid, dim3(256), 0, st, d, A, B, epi);
  return hipGetLastError();
}

}  // namespace xg
