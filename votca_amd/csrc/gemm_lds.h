// BF16 GEMM for gfx950 with a multi-stage direct-to-LDS pipeline and 128-row wave tiles.
//
//   C(m,n) = epilogue( sum_k A(m,k) B(n,k) )   (both operands K-contiguous bf16 rows)
//   SPLIT:   operands a = a_hi + a_lo, products hi*hi + hi*lo + lo*hi (FP32 accumulation)
//
// Block tile BM x BN (BM = 256, BN = 256 or 128), 256 threads = 2 x 2 waves, wave tile
// (BM/2) x (BN/2) of v_mfma_f32_16x16x32_bf16 tiles (a 128-row wave tile halves the LDS bytes
// read per MFMA against the 64 x 64 tile of gemm_bf16_kernel). One wave per SIMD: latency is
// hidden by STAGES K steps in flight, not by occupancy.
//
// Staging: global_load_lds_dwordx4 moves 16 bytes per lane from L2/HBM straight into LDS (lane l
// lands at base + 16 l), so in-flight tiles hold no registers. Every LDS row is 128 bytes: 64 k of
// one operand row (plain), or 32 k of the hi plane followed by the same 32 k of the lo plane
// (SPLIT). One load instruction fills 8 rows; the 8 lanes of a row fetch that row's 128 contiguous
// bytes (64 per plane when SPLIT), so every instruction is 8 full-line row segments. The 16-byte
// chunk c of row r sits at slot c ^ ((r & 7) ^ ((r >> 3) & 1)): MFMA fragment reads (16 rows, one
// chunk) then touch 16 distinct 16-byte bank groups, conflict-free for 8- and 16-lane LDS passes.
// Rows past M / N and chunks past the K range load from a zero page.
#pragma once
#include "gemm_kernels.h"
#include "gemm_pipe.h"

namespace xg {

__device__ __forceinline__ int lds_slot(int row, int c) { return c ^ ((row & 7) ^ ((row >> 3) & 1)); }

template <int BM, int BN, int STAGES, bool SPLIT, class Epi, int SCHED = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void gemm_bf16_lds_kernel(
    GemmDims d, Opnd<unsigned short> A, Opnd<unsigned short> B, Epi epi) {
  constexpr int KSTEP = SPLIT ? 32 : 64;  // k per stage
  constexpr int ROWB = 128;               // bytes per LDS row
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int ROWS = BM + BN;           // LDS rows per stage: A rows, then B rows
  constexpr int STAGE_BYTES = ROWS * ROWB;
  constexpr int LPW = STAGE_BYTES / 1024 / 4;  // load instructions per wave per stage
  static_assert(STAGE_BYTES % 4096 == 0, "a stage must split evenly over the four waves");
  static_assert(BM % 128 == 0 && BN % 128 == 0, "8-row load groups must not straddle A and B");
  static_assert(STAGES >= 2 && (STAGES - 2) * LPW <= 63, "vmcnt range");
  __shared__ __attribute__((aligned(1024))) unsigned char lds[STAGES * STAGE_BYTES];

  int tm, tn;
  tile_of_block(d, tm, tn);
  const int b = blockIdx.z, ks = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = ks * d.kchunk;
  const int kend = min(d.K, kbeg + d.kchunk);
  const int nk = kend > kbeg ? (kend - kbeg + KSTEP - 1) / KSTEP : 0;

  // this lane's source row and k offset for each of the wave's load instructions
  const unsigned short* src[LPW];
  int kof[LPW];
#pragma unroll
  for (int j = 0; j < LPW; ++j) {
    const int rr = 8 * (wave * LPW + j) + (lane >> 3);  // LDS row within the stage
    const int c = lds_slot(rr, lane & 7);               // the chunk this lane's slot holds
    const bool lo = SPLIT && c >= 4;
    kof[j] = SPLIT ? 8 * (c & 3) : 8 * c;
    const unsigned short* p = nullptr;
    if (rr < BM) {
      const int g = m0 + rr;
      if (g < d.M) p = (lo ? A.lo : A.p) + (long)b * A.bs + (long)g * A.ld;
    } else {
      const int g = n0 + rr - BM;
      if (g < d.N) p = (lo ? B.lo : B.p) + (long)b * B.bs + (long)g * B.ld;
    }
    src[j] = p;
  }
  const unsigned short* zp = reinterpret_cast<const unsigned short*>(g_zero_page);
  const unsigned lds0 = (unsigned)(uintptr_t)(lds_u4*)lds;

  auto issue = [&](int stage, int k0) {
    const unsigned base = lds0 + (unsigned)(stage * STAGE_BYTES) + (unsigned)(wave * LPW * 1024);
#pragma unroll
    for (int j = 0; j < LPW; ++j) {
      const int k = k0 + kof[j];
      const unsigned short* g = (src[j] && k < kend) ? src[j] + k : zp;
      dma16(g, base + (unsigned)(j * 1024));
    }
  };

  f4v acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (f4v){0.f, 0.f, 0.f, 0.f};

  auto frag = [&](const unsigned char* base, int row, int c) -> bf8v {
    return *reinterpret_cast<const bf8v*>(base + row * ROWB + lds_slot(row, c) * 16);
  };

  // All fragments of a K step are read first, B then A in the order the MFMAs consume them, so
  // the waits on LDS retire progressively while the MFMA chain runs (one register set per
  // fragment: with one wave per SIMD nothing else hides the LDS latency).
  constexpr int NSUB = SPLIT ? 1 : 2;  // 32-k sub-steps per stage
  auto compute = [&](int stage) {
    const unsigned char* sA = lds + stage * STAGE_BYTES;
    const unsigned char* sB = sA + BM * ROWB;
    const int r16 = lane & 15, q = lane >> 4;
    bf8v bh[NSUB][TN], bl[NSUB][TN], ah[NSUB][TM], al[NSUB][TM];
#pragma unroll
    for (int sub = 0; sub < NSUB; ++sub) {
      const int c = 4 * sub + q;  // plain: k half `sub`; SPLIT: hi chunk c, lo chunk c + 4
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        bh[sub][j] = frag(sB, wn0 + j * 16 + r16, c);
        if (SPLIT) bl[sub][j] = frag(sB, wn0 + j * 16 + r16, c + 4);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        ah[sub][i] = frag(sA, wm0 + i * 16 + r16, c);
        if (SPLIT) al[sub][i] = frag(sA, wm0 + i * 16 + r16, c + 4);
      }
    }
    // SCHED 0: nothing crosses this point, the reads all issue before the MFMA chain (left to itself
    // the scheduler interleaves one fragment read per group of MFMAs and waits on each);
    // SCHED 1 / 2: the compiler's MFMA / LDS-read interleaving strategies
    if constexpr (SCHED == 0) __builtin_amdgcn_sched_barrier(0);
    if constexpr (SCHED == 1) __builtin_amdgcn_iglp_opt(0);
    if constexpr (SCHED == 2) __builtin_amdgcn_iglp_opt(1);
#pragma unroll
    for (int sub = 0; sub < NSUB; ++sub)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[sub][i], bh[sub][j], acc[i][j], 0, 0, 0);
          if (SPLIT) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[sub][i], bl[sub][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[sub][i], bh[sub][j], acc[i][j], 0, 0, 0);
          }
        }
  };

#pragma unroll
  for (int s = 0; s < STAGES - 1; ++s)
    if (s < nk) issue(s, kbeg + s * KSTEP);

  for (int t = 0; t < nk; ++t) {
    // this wave's loads of step t are done once at most those of the later issued steps remain
    if constexpr (STAGES >= 3) {
      if (nk - 1 - t >= 1)
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"((STAGES - 2) * LPW) : "memory");
      else
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    // every wave's loads of step t are in LDS and every wave has finished reading step t - 1,
    // whose buffer the next issue overwrites
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    if (t + STAGES - 1 < nk) issue((t + STAGES - 1) % STAGES, kbeg + (t + STAGES - 1) * KSTEP);
    compute(t % STAGES);
  }
  epilogue_store<TM, TN, 4>(
      d, epi, b, ks, acc, [&](int i, int r) { return m0 + wm0 + i * 16 + 4 * (lane >> 4) + r; },
      [&](int j) { return n0 + wn0 + j * 16 + (lane & 15); });
}

// Usable when both operands are K-contiguous bf16 planes (hi and, for SPLIT, lo) with 16-byte
// aligned rows: ld, batch stride and K multiples of 8, K chunks multiples of the stage's k.
inline bool lds_ok(const Opnd<unsigned short>& A, const Opnd<unsigned short>& B, int K, bool split) {
  auto al = [](const void* p) { return p != nullptr && (reinterpret_cast<uintptr_t>(p) % 16) == 0; };
  return al(A.p) && al(B.p) && (!split || (al(A.lo) && al(B.lo))) && A.ld % 8 == 0 && B.ld % 8 == 0 &&
         A.bs % 8 == 0 && B.bs % 8 == 0 && K % 8 == 0;
}

template <int BM, int BN, int STAGES, class Epi, int SCHED = 0>
inline hipError_t launch_gemm_bf16_lds(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                       Opnd<unsigned short> B, bool split, Epi epi) {
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  if (split)
    hipLaunchKernelGGL((gemm_bf16_lds_kernel<BM, BN, STAGES, true, Epi, SCHED>), grid, dim3(256), 0, st, d, A, B, epi);
  else
    hipLaunchKernelGGL((gemm_bf16_lds_kernel<BM, BN, STAGES, false, Epi, SCHED>), grid, dim3(256), 0, st, d, A, B, epi);
  return hipGetLastError();
}

}  // namespace xg
