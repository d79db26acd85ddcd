// MFMA GEMM kernels for gfx950 (CDNA4): FP64 (v_mfma_f64_16x16x4f64) and
// BF16 with FP32 accumulation (v_mfma_f32_16x16x32_bf16), the latter optionally
// with split operands a = a_hi + a_lo ("bf16x3": hi*hi + hi*lo + lo*hi).
//
// C[b](m,n) = epilogue( sum_k A[b](m,k) * s[b](k) * B[b](k,n) )
//
// Operand addressing (one of the two strides is 1, chosen at compile time):
//   A(m,k) = A + b*a.bs + (AK ? m*a.ld + k : k*a.ld + m)
//   B(k,n) = B + b*b.bs + (BK ? n*b.ld + k : k*b.ld + n)
// AK/BK = "operand is contiguous along K". Every hot contraction of the GW-BSE path
// is laid out K-contiguous in HBM so both operands stream 16-byte vectors.
//
// Grid: x = output tile (optionally only tiles on/below the diagonal for symmetric
// rank-k updates such as eps = 1 + M^T D M), y = K split, z = batch.
// The epilogue functor receives (batch, ksplit, m, n, value) for each element in range.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <cstdlib>
#include <cstring>

namespace xg {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short us8v __attribute__((ext_vector_type(8)));

template <typename T>
struct Opnd {
  const T* p;     // hi (or full-precision) part
  const T* lo;    // lo part for split-bf16 operands (may be null)
  long ld;        // stride of the non-contiguous index
  long bs;        // batch stride
};

struct GemmDims {
  int M, N, K;
  int batch;
  int ksplit;    // number of K chunks (grid.y)
  int kchunk;    // K elements per chunk (multiple of the kernel's BK)
  int tiles_m, tiles_n;
  int lower;     // 1: only tiles with tile_n <= tile_m are computed (symmetric result)
  int nfast;     // 1: consecutive logical tiles walk along N (reuse of the A row block)
  int group;     // > 0: grouped order: `group` row blocks x all column blocks at a time (full outputs),
                 // super-blocks of group x group tiles (lower triangle)
};

// Lower-triangle tile `tile` (of tiles_m tile rows) in super-blocks of G x G tiles: super-row by
// super-row, the full blocks left of the diagonal first, then the diagonal block's triangle, so
// consecutive tiles share G row panels and at most G column panels.
__device__ __forceinline__ void lower_group_tile(int tile, int tiles_m, int G, int& tm, int& tn) {
  int cum = 0;
  for (int R = 0; R * G < tiles_m; ++R) {
    const int h = min(G, tiles_m - R * G);
    const int full = R * G * h;
    const int cnt = full + h * (h + 1) / 2;
    if (tile < cum + cnt) {
      const int u = tile - cum;
      if (u < full) {
        const int C = u / (G * h), w = u - C * (G * h);
        tm = R * G + w % h;
        tn = C * G + w / h;
      } else {
        const int v = u - full;
        int i = (int)((sqrt(8.0 * (double)v + 1.0) - 1.0) * 0.5);
        while ((i + 1) * (i + 2) / 2 <= v) ++i;
        while (i * (i + 1) / 2 > v) --i;
        tm = R * G + i;
        tn = R * G + (v - i * (i + 1) / 2);
      }
      return;
    }
    cum += cnt;
  }
  tm = tn = 0;  // not reached for tile < tiles_m (tiles_m + 1) / 2
}

// Block -> (tm, tn). Workgroups are dealt round-robin to the 8 XCDs (block b runs on
// XCD b % 8); the remap gives every XCD a contiguous range of logical tiles, so tiles that
// share an operand block run on the same XCD and hit its L2.
__device__ __forceinline__ void tile_of_block(const GemmDims& d, int& tm, int& tn) {
  const int nb = gridDim.x;
  const int bid = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = bid & 7;
  const int tile = x * q + min(x, r) + (bid >> 3);
  if (d.lower && d.group > 0) {
    lower_group_tile(tile, d.tiles_m, d.group, tm, tn);
  } else if (d.lower) {
    tm = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
    while ((tm + 1) * (tm + 2) / 2 <= tile) ++tm;
    while (tm * (tm + 1) / 2 > tile) --tm;
    tn = tile - tm * (tm + 1) / 2;
  } else if (d.group > 0) {
    // `group` row blocks sweep the column blocks together: the blocks resident on one XCD at a
    // time share both their A row panels and their B column panels in that XCD's L2
    const int span = d.group * d.tiles_n;
    const int g = tile / span, w = tile - g * span;
    const int gm = min(d.group, d.tiles_m - g * d.group);
    tn = w / gm;
    tm = g * d.group + (w - tn * gm);
  } else if (d.nfast) {
    tn = tile % d.tiles_n;
    tm = tile / d.tiles_n;
  } else {
    tm = tile % d.tiles_m;
    tn = tile / d.tiles_m;
  }
}

__device__ __forceinline__ unsigned short bf16_bits_rn(float f) {
  // round-to-nearest-even float -> bf16 (finite inputs)
  unsigned int u = __float_as_uint(f);
  unsigned int r = 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)((u + r) >> 16);
}
__device__ __forceinline__ float bf16_bits_to_f(unsigned short h) {
  return __uint_as_float(((unsigned int)h) << 16);
}

// Epilogue protocol. Every epilogue is callable as epi(b, ks, m, n, v). A "separable" one
// (kSeparable = true) also splits its output address into a row term and a column term:
//   typename Epi::Row r = epi.row(b, ks, m);  long c = epi.col(b, n);  epi.put(r, c, v);
// so a tile computes the (division-heavy) index arithmetic once per row and once per column
// instead of once per element.
template <class E, class = void>
struct epi_separable { static constexpr bool value = false; };
template <class E>
struct epi_separable<E, decltype((void)E::kSeparable)> { static constexpr bool value = E::kSeparable; };

// Store a (TM x TN) grid of MFMA result fragments through an epilogue. row_of(i, r) and
// col_of(j) give the global row / column of fragment element (i, j, r).
template <int TM, int TN, int NR, class Epi, class Acc, class RowOf, class ColOf>
__device__ __forceinline__ void epilogue_store(const GemmDims& d, const Epi& epi, int b, int ks, const Acc& acc,
                                               RowOf row_of, ColOf col_of) {
  if constexpr (epi_separable<Epi>::value) {
    long co[TN];
    bool cv[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int gn = col_of(j);
      cv[j] = gn < d.N;
      co[j] = cv[j] ? epi.col(b, gn) : 0;
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int gm = row_of(i, r);
        if (gm < d.M) {
          const typename Epi::Row ro = epi.row(b, ks, gm);
#pragma unroll
          for (int j = 0; j < TN; ++j)
            if (cv[j]) epi.put(ro, co[j], acc[i][j][r]);
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int gm = row_of(i, r), gn = col_of(j);
          if (gm < d.M && gn < d.N) epi(b, ks, gm, gn, acc[i][j][r]);
        }
  }
}

// ------------------------------------------------------------------ FP64 kernel
// Block tile BM x BN x 16, 256 threads = 2x2 waves, wave tile (BM/2) x (BN/2).
// VEC: operands are K-contiguous with ld, base and K bounds multiples of 2 (16-byte loads).
template <int BM, int BN, bool AK, bool BK_, bool HAS_S, bool VEC, class Epi>
__device__ __forceinline__ void gemm_f64_tile(const GemmDims& d, const Opnd<double>& A, const Opnd<double>& B,
                                              const double* __restrict__ s, long s_bs, const Epi& epi, int tm,
                                              int tn, int b, int ks) {
  constexpr int BK = 16;
  constexpr int PA = BM + 1;  // padded LDS row (doubles)
  constexpr int PB = BN + 1;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int LA = BM * BK / 256;  // elements per thread per tile
  constexpr int LB = BN * BK / 256;
  __shared__ double As[BK * PA];
  __shared__ double Bs[BK * PB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM;
  const int wn0 = (wave & 1) * WN;

  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = ks * d.kchunk;
  const int kend = min(d.K, kbeg + d.kchunk);

  const double* Ab = A.p + (long)b * A.bs;
  const double* Bb = B.p + (long)b * B.bs;
  const double* sb = HAS_S ? (s + (long)b * s_bs) : nullptr;

  d4v acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (d4v){0.0, 0.0, 0.0, 0.0};

  double ra[LA], rb[LB];
  typedef double d2v __attribute__((ext_vector_type(2)));

  auto load = [&](int k0) {
    if (VEC) {
      // LA/2 16-byte vectors per thread: vector id -> (row, k pair)
#pragma unroll
      for (int i = 0; i < LA / 2; ++i) {
        int vid = tid + i * 256;
        int mm = vid / (BK / 2), kq = 2 * (vid % (BK / 2));
        int gm = m0 + mm, gk = k0 + kq;
        d2v v = {0.0, 0.0};
        if (gm < d.M && gk < kend) v = *reinterpret_cast<const d2v*>(Ab + (long)gm * A.ld + gk);
        ra[2 * i] = v[0];
        ra[2 * i + 1] = v[1];
      }
#pragma unroll
      for (int i = 0; i < LB / 2; ++i) {
        int vid = tid + i * 256;
        int nn = vid / (BK / 2), kq = 2 * (vid % (BK / 2));
        int gn = n0 + nn, gk = k0 + kq;
        d2v v = {0.0, 0.0};
        if (gn < d.N && gk < kend) {
          v = *reinterpret_cast<const d2v*>(Bb + (long)gn * B.ld + gk);
          if (HAS_S) { v[0] *= sb[gk]; v[1] *= sb[gk + 1]; }
        }
        rb[2 * i] = v[0];
        rb[2 * i + 1] = v[1];
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      int idx = tid + i * 256;
      int mm, kk;
      if (AK) { mm = idx / BK; kk = idx % BK; } else { mm = idx % BM; kk = idx / BM; }
      int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < d.M && gk < kend) v = AK ? Ab[(long)gm * A.ld + gk] : Ab[(long)gk * A.ld + gm];
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      int idx = tid + i * 256;
      int nn, kk;
      if (BK_) { nn = idx / BK; kk = idx % BK; } else { nn = idx % BN; kk = idx / BN; }
      int gn = n0 + nn, gk = k0 + kk;
      double v = 0.0;
      if (gn < d.N && gk < kend) {
        v = BK_ ? Bb[(long)gn * B.ld + gk] : Bb[(long)gk * B.ld + gn];
        if (HAS_S) v *= sb[gk];
      }
      rb[i] = v;
    }
  };
  auto store = [&]() {
    if (VEC) {
#pragma unroll
      for (int i = 0; i < LA / 2; ++i) {
        int vid = tid + i * 256;
        int mm = vid / (BK / 2), kq = 2 * (vid % (BK / 2));
        As[kq * PA + mm] = ra[2 * i];
        As[(kq + 1) * PA + mm] = ra[2 * i + 1];
      }
#pragma unroll
      for (int i = 0; i < LB / 2; ++i) {
        int vid = tid + i * 256;
        int nn = vid / (BK / 2), kq = 2 * (vid % (BK / 2));
        Bs[kq * PB + nn] = rb[2 * i];
        Bs[(kq + 1) * PB + nn] = rb[2 * i + 1];
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      int idx = tid + i * 256;
      int mm, kk;
      if (AK) { mm = idx / BK; kk = idx % BK; } else { mm = idx % BM; kk = idx / BM; }
      As[kk * PA + mm] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      int idx = tid + i * 256;
      int nn, kk;
      if (BK_) { nn = idx / BK; kk = idx % BK; } else { nn = idx % BN; kk = idx / BN; }
      Bs[kk * PB + nn] = rb[i];
    }
  };

  if (kbeg < kend) {
    load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      __syncthreads();
      store();
      __syncthreads();
      if (k0 + BK < kend) load(k0 + BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        double af[TM], bf[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = As[kr * PA + wm0 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < TN; ++j) bf[j] = Bs[kr * PB + wn0 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
    }
  }
  // D layout of 16x16x4 f64 (measured on gfx950): lane l, register r holds
  // D[(l/16) + 4r][l%16]; A[i][k] sits in lane i + 16k and B[k][j] in lane j + 16k.
  epilogue_store<TM, TN, 4>(
      d, epi, b, ks, acc, [&](int i, int r) { return m0 + wm0 + i * 16 + (lane >> 4) + 4 * r; },
      [&](int j) { return n0 + wn0 + j * 16 + (lane & 15); });
  __syncthreads();  // LDS is reused by the next tile of a persistent caller
}

template <int BM, int BN, bool AK, bool BK_, bool HAS_S, bool VEC, class Epi, int MINW = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MINW))) void gemm_f64_kernel(
    GemmDims d, Opnd<double> A, Opnd<double> B, const double* __restrict__ s, long s_bs, Epi epi) {
  int tm, tn;
  tile_of_block(d, tm, tn);
  gemm_f64_tile<BM, BN, AK, BK_, HAS_S, VEC, Epi>(d, A, B, s, s_bs, epi, tm, tn, blockIdx.z, blockIdx.y);
}

// ------------------------------------------------------------------ BF16 kernel
// Operands are bf16 bit patterns (unsigned short). SPLIT: use hi/lo parts of both
// operands (3 MFMAs per fragment pair). HAS_S: B(k,n) is rescaled by s[k] (float)
// on load and re-split into hi/lo, so the weight costs no extra precision.
// Block tile BM x BN x BKT, NT threads = (NT/128) x 2 waves.
// VEC: operands are K-contiguous with ld, base and K bounds multiples of 8 (16-byte loads).
template <int BM, int BN, bool AK, bool BK_, bool SPLIT, bool HAS_S, bool VEC, class Epi, int BKT = 32, int NT = 256,
          bool DB = false, int PAD = 8>
__device__ __forceinline__ void gemm_bf16_tile(const GemmDims& d, const Opnd<unsigned short>& A,
                                               const Opnd<unsigned short>& B, const float* __restrict__ s,
                                               long s_bs, const Epi& epi, int tm, int tn, int b, int ks) {
  // BK = 64 (VEC only) gives every row a full 128-byte line per k-step: 8 lanes per row, 8 rows
  // per load instruction; the padded row (72 elements = 144 B) keeps MFMA fragment reads
  // conflict-free. BK = 128 (plain BF16: one plane per operand) keeps the same LDS and register
  // footprint as split BK = 64 and amortises each barrier over twice the MFMAs; its 272-byte
  // padded rows are conflict-free as well.
  constexpr int BK = BKT;
  static_assert(BK == 32 || ((BK == 64 || BK == 128) && VEC), "BK 64 / 128 need vector loads");
  static_assert(PAD % 8 == 0, "LDS rows must stay 16-byte aligned");
  constexpr int PK = BK + PAD;  // padded LDS row (bf16 elements)
  static_assert(NT == 256 || NT == 512, "256 or 512 threads");
  constexpr int WM = BM / (NT / 128), WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int NS = SPLIT ? 2 : 1;
  constexpr int LA = BM * BK / NT;
  constexpr int LB = BN * BK / NT;
  // DB: two LDS stages; the next panel is stored into the idle stage after the MFMAs of the
  // current one, so a K step needs one barrier instead of two
  constexpr int NB = DB ? 2 : 1;
  __shared__ __attribute__((aligned(16))) unsigned short As_[NB][NS][BM * PK];
  __shared__ __attribute__((aligned(16))) unsigned short Bs_[NB][NS][BN * PK];
  int stage = 0;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM;
  const int wn0 = (wave & 1) * WN;

  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = ks * d.kchunk;
  const int kend = min(d.K, kbeg + d.kchunk);

  const unsigned short* Ah = A.p + (long)b * A.bs;
  const unsigned short* Al = SPLIT ? A.lo + (long)b * A.bs : nullptr;
  const unsigned short* Bh = B.p + (long)b * B.bs;
  const unsigned short* Bl = SPLIT || HAS_S ? (B.lo ? B.lo + (long)b * B.bs : nullptr) : nullptr;
  const float* sb = HAS_S ? (s + (long)b * s_bs) : nullptr;

  f4v acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (f4v){0.f, 0.f, 0.f, 0.f};

  unsigned short ra[NS][LA], rb[NS][LB];
  constexpr int VA = BM * BK / 8 / NT;  // 16-byte vectors per thread (VEC path)
  constexpr int VB = BN * BK / 8 / NT;
  us8v va[NS][VA > 0 ? VA : 1], vb[NS][VB > 0 ? VB : 1];

  auto load_vec = [&](int k0) {
#pragma unroll
    for (int i = 0; i < VA; ++i) {
      int vid = tid + i * NT;
      int mm = vid / (BK / 8), kq = 8 * (vid % (BK / 8));
      int gm = m0 + mm, gk = k0 + kq;
      us8v h = {0, 0, 0, 0, 0, 0, 0, 0}, l = h;
      if (gm < d.M && gk < kend) {
        long off = (long)gm * A.ld + gk;
        h = *reinterpret_cast<const us8v*>(Ah + off);
        if (SPLIT) l = *reinterpret_cast<const us8v*>(Al + off);
      }
      va[0][i] = h;
      if (SPLIT) va[NS - 1][i] = l;
    }
#pragma unroll
    for (int i = 0; i < VB; ++i) {
      int vid = tid + i * NT;
      int nn = vid / (BK / 8), kq = 8 * (vid % (BK / 8));
      int gn = n0 + nn, gk = k0 + kq;
      us8v h = {0, 0, 0, 0, 0, 0, 0, 0}, l = h;
      if (gn < d.N && gk < kend) {
        long off = (long)gn * B.ld + gk;
        h = *reinterpret_cast<const us8v*>(Bh + off);
        if (Bl) l = *reinterpret_cast<const us8v*>(Bl + off);
        if (HAS_S) {
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            float v = (bf16_bits_to_f(h[e]) + bf16_bits_to_f(l[e])) * sb[gk + e];
            unsigned short hh = bf16_bits_rn(v);
            h[e] = hh;
            l[e] = bf16_bits_rn(v - bf16_bits_to_f(hh));
          }
        }
      }
      vb[0][i] = h;
      if (SPLIT) vb[NS - 1][i] = l;
    }
  };
  auto store_vec = [&]() {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int i = 0; i < VA; ++i) {
        int vid = tid + i * NT;
        int mm = vid / (BK / 8), kq = 8 * (vid % (BK / 8));
        *reinterpret_cast<us8v*>(&As_[stage][q][mm * PK + kq]) = va[q][i];
      }
#pragma unroll
      for (int i = 0; i < VB; ++i) {
        int vid = tid + i * NT;
        int nn = vid / (BK / 8), kq = 8 * (vid % (BK / 8));
        *reinterpret_cast<us8v*>(&Bs_[stage][q][nn * PK + kq]) = vb[q][i];
      }
    }
  };

  auto load = [&](int k0) {
    if (VEC) { load_vec(k0); return; }
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      int idx = tid + i * NT;
      int mm, kk;
      if (AK) { mm = idx / BK; kk = idx % BK; } else { mm = idx % BM; kk = idx / BM; }
      int gm = m0 + mm, gk = k0 + kk;
      unsigned short h = 0, l = 0;
      if (gm < d.M && gk < kend) {
        long off = AK ? (long)gm * A.ld + gk : (long)gk * A.ld + gm;
        h = Ah[off];
        if (SPLIT) l = Al[off];
      }
      ra[0][i] = h;
      if (SPLIT) ra[NS - 1][i] = l;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      int idx = tid + i * NT;
      int nn, kk;
      if (BK_) { nn = idx / BK; kk = idx % BK; } else { nn = idx % BN; kk = idx / BN; }
      int gn = n0 + nn, gk = k0 + kk;
      unsigned short h = 0, l = 0;
      if (gn < d.N && gk < kend) {
        long off = BK_ ? (long)gn * B.ld + gk : (long)gk * B.ld + gn;
        h = Bh[off];
        if (Bl) l = Bl[off];
        if (HAS_S) {
          float v = (bf16_bits_to_f(h) + bf16_bits_to_f(l)) * sb[gk];
          h = bf16_bits_rn(v);
          l = bf16_bits_rn(v - bf16_bits_to_f(h));
        }
      }
      rb[0][i] = h;
      if (SPLIT) rb[NS - 1][i] = l;
    }
  };
  auto store = [&]() {
    if (VEC) { store_vec(); return; }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int i = 0; i < LA; ++i) {
        int idx = tid + i * NT;
        int mm, kk;
        if (AK) { mm = idx / BK; kk = idx % BK; } else { mm = idx % BM; kk = idx / BM; }
        As_[stage][q][mm * PK + kk] = ra[q][i];
      }
#pragma unroll
      for (int i = 0; i < LB; ++i) {
        int idx = tid + i * NT;
        int nn, kk;
        if (BK_) { nn = idx / BK; kk = idx % BK; } else { nn = idx % BN; kk = idx / BN; }
        Bs_[stage][q][nn * PK + kk] = rb[q][i];
      }
    }
  };

  auto compute = [&](int st_) {
#pragma unroll
    for (int kk = 0; kk < BK / 32; ++kk) {
      const int kq = kk * 32 + 8 * (lane >> 4);
      bf8v ah[TM], bh[TN], al[TM], bl[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        ah[i] = *reinterpret_cast<const bf8v*>(&As_[st_][0][(wm0 + i * 16 + (lane & 15)) * PK + kq]);
        if (SPLIT) al[i] = *reinterpret_cast<const bf8v*>(&As_[st_][NS - 1][(wm0 + i * 16 + (lane & 15)) * PK + kq]);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        bh[j] = *reinterpret_cast<const bf8v*>(&Bs_[st_][0][(wn0 + j * 16 + (lane & 15)) * PK + kq]);
        if (SPLIT) bl[j] = *reinterpret_cast<const bf8v*>(&Bs_[st_][NS - 1][(wn0 + j * 16 + (lane & 15)) * PK + kq]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          if (SPLIT) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          }
        }
    }
  };

  if (kbeg < kend) {
    if constexpr (DB) {
      // stage s holds panel k0 while the registers carry panel k0 + BK; every wave has left
      // stage s ^ 1 (the barrier closing the previous step) before it is overwritten
      load(kbeg);
      store();
      __syncthreads();
      for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const bool more = k0 + BK < kend;
        if (more) load(k0 + BK);
        compute(stage);
        if (more) {
          stage ^= 1;
          store();
        }
        __syncthreads();
      }
    } else {
      load(kbeg);
      for (int k0 = kbeg; k0 < kend; k0 += BK) {
        __syncthreads();
        store();
        __syncthreads();
        if (k0 + BK < kend) load(k0 + BK);
        compute(0);
      }
    }
  }
  epilogue_store<TM, TN, 4>(
      d, epi, b, ks, acc, [&](int i, int r) { return m0 + wm0 + i * 16 + 4 * (lane >> 4) + r; },
      [&](int j) { return n0 + wn0 + j * 16 + (lane & 15); });
  __syncthreads();  // LDS is reused by the next tile of a persistent caller
}

template <int BM, int BN, bool AK, bool BK_, bool SPLIT, bool HAS_S, bool VEC, class Epi, int BKT = 32, int NT = 256,
          bool DB = false, int MINW = 1, int PAD = 8>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW))) void gemm_bf16_kernel(
    GemmDims d, Opnd<unsigned short> A, Opnd<unsigned short> B, const float* __restrict__ s, long s_bs, Epi epi) {
  int tm, tn;
  tile_of_block(d, tm, tn);
  gemm_bf16_tile<BM, BN, AK, BK_, SPLIT, HAS_S, VEC, Epi, BKT, NT, DB, PAD>(d, A, B, s, s_bs, epi, tm, tn, blockIdx.z,
                                                                       blockIdx.y);
}

// ------------------------------------------------------------------ host launch helpers
template <int BM, int BN>
inline GemmDims make_dims(int M, int N, int K, int batch, int ksplit, int bk, bool lower) {
  GemmDims d;
  d.M = M; d.N = N; d.K = K; d.batch = batch;
  d.tiles_m = (M + BM - 1) / BM;
  d.tiles_n = (N + BN - 1) / BN;
  if (ksplit < 1) ksplit = 1;
  int kc = (K + ksplit - 1) / ksplit;
  kc = ((kc + bk - 1) / bk) * bk;
  if (kc < bk) kc = bk;
  d.kchunk = kc;
  d.ksplit = (K + kc - 1) / kc;
  if (d.ksplit < 1) d.ksplit = 1;
  d.lower = lower ? 1 : 0;
  d.nfast = d.tiles_n <= d.tiles_m ? 1 : 0;
  d.group = 0;
  return d;
}

// 16-byte vector loads are legal when both operands are K-contiguous and every row start,
// the K extent and the chunk bounds are multiples of `elems` elements.
template <typename T>
inline bool vec_ok(const Opnd<T>& o, int K, int elems) {
  auto al = [&](const void* p) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) % 16) == 0; };
  return al(o.p) && al(o.lo) && o.ld % elems == 0 && o.bs % elems == 0 && K % elems == 0;
}

inline int gemm_tiles(const GemmDims& d) {
  return d.lower ? d.tiles_m * (d.tiles_m + 1) / 2 : d.tiles_m * d.tiles_n;
}

template <int BM, int BN, bool AK, bool BK_, class Epi, int MINW = 1>
inline hipError_t launch_gemm_f64(hipStream_t st, const GemmDims& d, Opnd<double> A, Opnd<double> B,
                                  const double* s, long s_bs, Epi epi) {
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  const bool vec = AK && BK_ && vec_ok(A, d.K, 2) && vec_ok(B, d.K, 2) && (s_bs % 2 == 0);
  if (vec) {
    if (s)
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, AK, BK_, true, true, Epi, MINW>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, AK, BK_, false, true, Epi, MINW>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  } else {
    if (s)
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, AK, BK_, true, false, Epi>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, AK, BK_, false, false, Epi>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  }
  return hipGetLastError();
}

// K-step of the vector-load path; d.kchunk must be a multiple of it (make_dims with this bk).
inline bool bf16_vec(const Opnd<unsigned short>& A, const Opnd<unsigned short>& B, int K, long s_bs) {
  return vec_ok(A, K, 8) && vec_ok(B, K, 8) && (s_bs % 8 == 0);
}

// 256 x 256 block tile (wave tile 128 x 128, one wave per SIMD), K step 32, vector loads:
// half the operand bytes per MFMA of the 128 x 128 tile. The caller checks bf16_vec().
template <bool AK, bool BK_, class Epi>
inline hipError_t launch_gemm_bf16_big(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                       Opnd<unsigned short> B, const float* s, long s_bs, bool split, Epi epi) {
  static_assert(AK && BK_, "K-contiguous operands only");
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  if (split) {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<256, 256, AK, BK_, true, true, true, Epi, 32>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<256, 256, AK, BK_, true, false, true, Epi, 32>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  } else {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<256, 256, AK, BK_, false, true, true, Epi, 32>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<256, 256, AK, BK_, false, false, true, Epi, 32>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  }
  return hipGetLastError();
}

// 256 x 128 block tile with 512 threads (8 waves, wave tile 64 x 64), K step 64, vector loads:
// a quarter fewer operand bytes per MFMA than 128 x 128 at the same per-wave register budget.
template <bool AK, bool BK_, class Epi>
inline hipError_t launch_gemm_bf16_tall(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                        Opnd<unsigned short> B, const float* s, long s_bs, bool split, Epi epi) {
  static_assert(AK && BK_, "K-contiguous operands only");
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  if (split) {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, true, true, true, Epi, 64, 512>), grid, dim3(512), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, true, false, true, Epi, 64, 512>), grid, dim3(512), 0, st, d, A, B, s, s_bs, epi);
  } else {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, false, true, true, Epi, 64, 512>), grid, dim3(512), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, false, false, true, Epi, 64, 512>), grid, dim3(512), 0, st, d, A, B, s, s_bs, epi);
  }
  return hipGetLastError();
}

// 256 x 128 block tile with 256 threads (wave tile 128 x 64: 32 accumulator tiles per wave),
// K step BKT, vector loads: 1.5x the MFMAs per LDS fragment read of the 64 x 64 wave tile.
template <int BKT, bool AK, bool BK_, class Epi>
inline hipError_t launch_gemm_bf16_wave128(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                           Opnd<unsigned short> B, const float* s, long s_bs, bool split, Epi epi) {
  static_assert(AK && BK_, "K-contiguous operands only");
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  if (split) {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, true, true, true, Epi, 64, 256>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, true, false, true, Epi, 64, 256>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  } else {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, false, true, true, Epi, BKT, 256>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<256, 128, AK, BK_, false, false, true, Epi, BKT, 256>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  }
  return hipGetLastError();
}

// 128 x 128 block tile, 256 threads, two LDS stages (one barrier per K step), K step BKT
// (64 for split operands), vector loads.
template <int BKT, bool AK, bool BK_, class Epi>
inline hipError_t launch_gemm_bf16_db(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                      Opnd<unsigned short> B, const float* s, long s_bs, bool split, Epi epi) {
  static_assert(AK && BK_, "K-contiguous operands only");
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  if (split) {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, true, true, true, Epi, 64, 256, true>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, true, false, true, Epi, 64, 256, true>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  } else {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, false, true, true, Epi, BKT, 256, true>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, false, false, true, Epi, BKT, 256, true>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  }
  return hipGetLastError();
}

// 128 x 128 block tile, 256 threads, one LDS stage, K step BKT for both forms, compiled for at
// least MINW waves per SIMD (registers capped accordingly).
template <int MINW, int BKT, bool AK, bool BK_, class Epi, int PAD = 8>
inline hipError_t launch_gemm_bf16_occ(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                       Opnd<unsigned short> B, const float* s, long s_bs, bool split, Epi epi) {
  static_assert(AK && BK_, "K-contiguous operands only");
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  if (split) {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, true, true, true, Epi, BKT, 256, false, MINW, PAD>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, true, false, true, Epi, BKT, 256, false, MINW, PAD>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  } else {
    if (s) hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, false, true, true, Epi, BKT, 256, false, MINW, PAD>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
    else hipLaunchKernelGGL((gemm_bf16_kernel<128, 128, AK, BK_, false, false, true, Epi, BKT, 256, false, MINW, PAD>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
  }
  return hipGetLastError();
}

// K step of the wide-K vector path: 64 for split operands (two planes each), 128 for plain BF16.
// d.kchunk of a launch with wide = true must be a multiple of it (make_dims with this bk).
inline int bf16_wide_kstep(bool split) { return split ? 64 : 128; }

template <int BM, int BN, bool AK, bool BK_, class Epi>
inline hipError_t launch_gemm_bf16(hipStream_t st, const GemmDims& d, Opnd<unsigned short> A,
                                   Opnd<unsigned short> B, const float* s, long s_bs, bool split,
                                   Epi epi, bool wide = false) {
  dim3 grid(gemm_tiles(d), d.ksplit, d.batch);
  const bool vec = AK && BK_ && bf16_vec(A, B, d.K, s_bs);
  if constexpr (AK && BK_) {
    if (vec && wide) {
      if (split) {
        if (s) hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, AK, BK_, true, true, true, Epi, 64>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
        else hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, AK, BK_, true, false, true, Epi, 64>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
      } else {
        if (s) hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, AK, BK_, false, true, true, Epi, 128>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
        else hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, AK, BK_, false, false, true, Epi, 128>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi);
      }
      return hipGetLastError();
    }
  }
#define XG_BF16_LAUNCH(SP, HS, VV) \
  hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, AK, BK_, SP, HS, VV, Epi>), grid, dim3(256), 0, st, d, A, B, s, s_bs, epi)
  if (vec) {
    if (split) { if (s) XG_BF16_LAUNCH(true, true, true); else XG_BF16_LAUNCH(true, false, true); }
    else { if (s) XG_BF16_LAUNCH(false, true, true); else XG_BF16_LAUNCH(false, false, true); }
  } else {
    if (split) { if (s) XG_BF16_LAUNCH(true, true, false); else XG_BF16_LAUNCH(true, false, false); }
    else { if (s) XG_BF16_LAUNCH(false, true, false); else XG_BF16_LAUNCH(false, false, false); }
  }
#undef XG_BF16_LAUNCH
  return hipGetLastError();
}

}  // namespace xg
