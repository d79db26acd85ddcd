// libxtpgpu: MI355X (gfx950) GW-BSE dense-contraction hot path behind a C-ABI.
//
// Reference surface replaced here (file:line in the reference tree):
//   TCMatrix_gwbse             xtp/include/votca/xtp/threecenter.h:70-142, xtp/src/libxtp/threecenter.cc:29-131
//   RPA::calculate_epsilon*    xtp/src/libxtp/gwbse/rpa.cc:75-202
//   BSE_OPERATOR::matmul/diag  xtp/src/libxtp/gwbse/bse_operator.cc:28-188
//   DavidsonSolver             xtp/include/votca/xtp/davidsonsolver.h:64-280, xtp/src/libxtp/davidsonsolver.cc
//   OpenMP_CUDA / CudaPipeline xtp/src/libxtp/openmp_cuda.cc, xtp/src/libxtp/cudapipeline.cc
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <tuple>
#include <type_traits>
#include <cmath>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "gemm_kernels.h"
#include "gemm_pipe.h"
#include "gemm_lds.h"
#include "host_eigen.h"
#include "kernels_misc.h"
#include "qp_solver.h"
#include "xtpgpu_internal.h"

using namespace xg;

static thread_local std::string g_last_error;

template <class F>
static int guarded(F&& f) {
  // a failure returned by an earlier call must not be reported again by this one
  (void)hipGetLastError();
  try {
    f();
    return XTPGPU_OK;
  } catch (const xg::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return XTPGPU_ERR_ARG;
  } catch (...) {
    g_last_error = "unknown error";
    return XTPGPU_ERR_ARG;
  }
}

namespace {

// ------------------------------------------------------------------ epilogues
struct EpiStore {  // C[b*bs + m*sm + n*sn] = alpha*v + beta*C
  double* C;
  long sm, sn, bs;
  double alpha, beta;
  static constexpr bool kSeparable = true;
  typedef long Row;
  __device__ long row(int b, int, int m) const { return (long)b * bs + (long)m * sm; }
  __device__ long col(int, int n) const { return (long)n * sn; }
  template <typename V>
  __device__ void put(long r, long c, V v) const {
    double* p = C + r + c;
    double x = alpha * (double)v;
    if (beta != 0.0) x += beta * (*p);
    *p = x;
  }
  template <typename V>
  __device__ void operator()(int b, int ks, int m, int n, V v) const { put(row(b, ks, m), col(b, n), v); }
};

struct EpiPartial {  // split-K partials: part[(ks*batch + b)*M*N + n*M + m]
  double* part;
  int M, N, batch;
  static constexpr bool kSeparable = true;
  typedef long Row;
  __device__ long row(int b, int ks, int m) const { return ((long)ks * batch + b) * (long)M * N + m; }
  __device__ long col(int, int n) const { return (long)n * M; }
  template <typename V>
  __device__ void put(long r, long c, V v) const { part[r + c] = (double)v; }
  template <typename V>
  __device__ void operator()(int b, int ks, int m, int n, V v) const { put(row(b, ks, m), col(b, n), v); }
};

struct EpiPart32 {  // FP32 split-K partials of a batched product: part[((ks*batch + b)*N + n)*M + m]
  float* part;
  long M, N, batch;
  static constexpr bool kSeparable = true;
  typedef long Row;
  __device__ long row(int b, int ks, int m) const { return ((long)ks * batch + b) * N * M + m; }
  __device__ long col(int, int n) const { return (long)n * M; }
  template <typename V>
  __device__ void put(long r, long c, V v) const { part[r + c] = (float)v; }
  template <typename V>
  __device__ void operator()(int b, int ks, int m, int n, V v) const { put(row(b, ks, m), col(b, n), v); }
};

struct EpiEpsSym {  // eps[b](m,n) = eps[b](n,m) = v + delta_mn, written once from m >= n
  double* out;
  long P;
  template <typename V>
  __device__ void operator()(int b, int, int m, int n, V v) const {
    if (m < n) return;
    double r = (double)v + (m == n ? 1.0 : 0.0);
    double* o = out + (long)b * P * P;
    o[(long)n * P + m] = r;
    o[(long)m * P + n] = r;
  }
};

// Hd / Hd2 first-stage product, stored K-contiguous for the second stage.
// Row m = p*rows + r, column n = kv*cols + j  ->  S[(r*kvec + kv)*(pc*colsp) + p*colsp + j]
// (by_row) or S[(kv*rows + r)*(pc*colsp) + p*colsp + j]; pads (j >= cols) stay zero.
struct EpiStage1 {
  double* S64;
  unsigned short* Sh;
  unsigned short* Sl;
  const double* scale;  // optional per-p scale (eps^-1), indexed p + p0
  long p0;
  int rows, cols, colsp, kvec, pc;  // colsp: padded row length of the K-contiguous output
  bool by_row;
  // idx = outer * pc * colsp + p * colsp + j with outer = kv * rows + r (or r * kvec + kv):
  // the row (p, r) and column (kv, j) terms separate.
  static constexpr bool kSeparable = true;
  struct Row {
    long off;
    double f;
  };
  __device__ Row row(int, int, int m) const {
    const int p = m / rows, r = m - p * rows;
    const long slab = (long)pc * colsp;
    Row o;
    o.off = (by_row ? (long)r * kvec * slab : (long)r * slab) + (long)p * colsp;
    o.f = scale ? scale[p0 + p] : 1.0;
    return o;
  }
  __device__ long col(int, int n) const {
    const int kv = n / cols, j = n - kv * cols;
    const long slab = (long)pc * colsp;
    return (by_row ? (long)kv * slab : (long)kv * rows * slab) + j;
  }
  template <typename V>
  __device__ void put(const Row& ro, long c, V v) const {
    const long idx = ro.off + c;
    if constexpr (std::is_same<V, float>::value) {
      // BF16 kernels: FP32 accumulator; the split residual x - hi is exact in FP32
      const float x = scale ? v * (float)ro.f : v;
      const unsigned short h = bf16_bits_rn(x);
      Sh[idx] = h;
      if (Sl) Sl[idx] = bf16_bits_rn(x - bf16_bits_to_f(h));
    } else {
      const double x = (double)v * ro.f;
      if (S64) {
        S64[idx] = x;
      } else {
        unsigned short h, l;
        split_bf16(x, h, l);
        Sh[idx] = h;
        if (Sl) Sl[idx] = l;
      }
    }
  }
  template <typename V>
  __device__ void operator()(int b, int ks, int m, int n, V v) const { put(row(b, ks, m), col(b, n), v); }
};

// ------------------------------------------------------------------ GEMM dispatch
// Row-block group for the grouped tile order (0 = plain order): used when both tile counts
// are large enough for panels to be evicted between reuses. XTPGPU_TILE_GROUP overrides.
inline int tile_group(const GemmDims& d) {
  const char* e = std::getenv("XTPGPU_TILE_GROUP");
  if (e) return std::max(0, std::atoi(e));
  return (d.tiles_m >= 64 && d.tiles_n >= 8) ? 8 : 0;
}

// Super-block size of the lower-triangle tile order (0 = row by row, the default): G x G tiles
// keep G row and G column panels in flight per XCD instead of one row panel and every column
// panel. XTPGPU_LOWER_GROUP=G selects it; on the eps builds it measured level with row by row
// (FP64 51.6 / 51.4, BF16x3 210.8 / 212.1 TFLOP/s at P = 2000).
inline int lower_tile_group(const GemmDims& d) {
  const char* e = std::getenv("XTPGPU_LOWER_GROUP");
  if (e) return std::max(0, std::atoi(e));
  (void)d;
  return 0;
}

template <bool AK, bool BK_, class Epi>
void gemm64(hipStream_t st, int M, int N, int K, int batch, int ksplit, Opnd<double> A, Opnd<double> B,
            const double* s, long s_bs, bool lower, Epi epi) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  // 128x128 tiles halve the operand traffic per FLOP; used when the grid still fills the chip
  const long big_tiles = (long)((M + 127) / 128) * ((N + 127) / 128) * batch * std::max(ksplit, 1);
  if (ksplit <= 1 && M >= 256 && N >= 256 && big_tiles >= 512) {
    GemmDims d = make_dims<128, 128>(M, N, std::max(K, 0), batch, ksplit, 16, lower);
    d.group = lower ? lower_tile_group(d) : tile_group(d);
    // registers capped for two waves per SIMD (210 VGPRs, accumulators in VGPRs): 61 against 44
    // TFLOP/s at one wave per SIMD on the Hd stage-1 product; XTPGPU_F64_KERNEL=occ1 selects the latter
    const char* e = std::getenv("XTPGPU_F64_KERNEL");
    if (e && std::strcmp(e, "occ1") == 0)
      XG_HIP((launch_gemm_f64<128, 128, AK, BK_>(st, d, A, B, s, s_bs, epi)));
    else
      XG_HIP((launch_gemm_f64<128, 128, AK, BK_, Epi, 2>(st, d, A, B, s, s_bs, epi)));
    return;
  }
  GemmDims d = make_dims<64, 64>(M, N, std::max(K, 0), batch, ksplit, 16, lower);
  XG_HIP((launch_gemm_f64<64, 64, AK, BK_>(st, d, A, B, s, s_bs, epi)));
}

// BF16 kernel choice (XTPGPU_BF16_KERNEL, for A/B timing; anything else is the default, 1):
// 1 = K-contiguous operands: split forms 128x128 with BK 64 (two workgroups per CU), plain BF16
// 128x128 with BK 64 capped for three waves per SIMD; staged32 = BK 32; wide = split BK 64 /
// plain BK 128 at two workgroups per CU; pipe = direct-to-LDS pipeline; big = 256x256; tall =
// 256x128 with 512 threads; wave128 / wave128k128 = 256x128 with 256 threads; db64 / db128 = two
// LDS stages; k64occ{2,3,4} / k32occ{3,4} = register caps for 2-4 waves per SIMD.
inline int bf16_kernel_choice() {
  const char* e = std::getenv("XTPGPU_BF16_KERNEL");
  if (e && std::strcmp(e, "staged32") == 0) return 0;
  if (e && std::strcmp(e, "pipe") == 0) return 2;
  if (e && std::strcmp(e, "big") == 0) return 3;
  if (e && std::strcmp(e, "tall") == 0) return 4;
  if (e && std::strcmp(e, "wave128") == 0) return 5;
  if (e && std::strcmp(e, "wave128k128") == 0) return 6;
  if (e && std::strcmp(e, "db64") == 0) return 7;
  if (e && std::strcmp(e, "db128") == 0) return 8;
  if (e && std::strcmp(e, "k64occ2") == 0) return 9;
  if (e && std::strcmp(e, "k64occ3") == 0) return 10;
  if (e && std::strcmp(e, "k64occ4") == 0) return 11;
  if (e && std::strcmp(e, "k32occ3") == 0) return 12;
  if (e && std::strcmp(e, "k32occ4") == 0) return 13;
  if (e && std::strcmp(e, "wide") == 0) return 14;
  if (e && std::strcmp(e, "lds256") == 0) return 15;
  if (e && std::strcmp(e, "lds128") == 0) return 16;
  return 1;
}

template <bool AK, bool BK_, class Epi>
void gemm16(hipStream_t st, int M, int N, int K, int batch, int ksplit, Opnd<unsigned short> A,
            Opnd<unsigned short> B, const float* s, long s_bs, bool split, bool lower, Epi epi) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  const int choice = bf16_kernel_choice();
  if constexpr (AK && BK_) {
    // direct-to-LDS pipeline with 128-row wave tiles (256 x 256 block, 2 stages; 256 x 128, 3 stages)
    if ((choice == 15 || choice == 16) && !s && lds_ok(A, B, K, split) && !(lower && choice == 16)) {
      if (choice == 15) {
        GemmDims d = make_dims<256, 256>(M, N, std::max(K, 0), batch, ksplit, split ? 32 : 64, lower);
        d.group = lower ? lower_tile_group(d) : tile_group(d);
        const char* sc = std::getenv("XTPGPU_LDS_SCHED");
        const int sched = sc ? std::atoi(sc) : 0;
        if constexpr (std::is_same<Epi, EpiStage1>::value) {
          if (sched == 1) { XG_HIP((launch_gemm_bf16_lds<256, 256, 2, Epi, 1>(st, d, A, B, split, epi))); return; }
          if (sched == 2) { XG_HIP((launch_gemm_bf16_lds<256, 256, 2, Epi, 2>(st, d, A, B, split, epi))); return; }
        }
        XG_HIP((launch_gemm_bf16_lds<256, 256, 2>(st, d, A, B, split, epi)));
      } else {
        GemmDims d = make_dims<256, 128>(M, N, std::max(K, 0), batch, ksplit, split ? 32 : 64, false);
        d.group = tile_group(d);
        XG_HIP((launch_gemm_bf16_lds<256, 128, 3>(st, d, A, B, split, epi)));
      }
      return;
    }
    if (choice == 3 && bf16_vec(A, B, K, s_bs) && (long)M * N >= 256L * 256 * 512) {
      GemmDims d = make_dims<256, 256>(M, N, std::max(K, 0), batch, ksplit, 32, lower);
      XG_HIP((launch_gemm_bf16_big<AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      return;
    }
    // the lower-triangle tile walk assumes square tiles: 256 x 128 only for full outputs
    if (choice == 4 && !lower && bf16_vec(A, B, K, s_bs) && (long)M * N >= 256L * 128 * 512) {
      GemmDims d = make_dims<256, 128>(M, N, std::max(K, 0), batch, ksplit, 64, false);
      d.group = tile_group(d);
      XG_HIP((launch_gemm_bf16_tall<AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      return;
    }
    if ((choice == 5 || choice == 6) && !lower && bf16_vec(A, B, K, s_bs) && (long)M * N >= 256L * 128 * 512) {
      const int bk = split ? 64 : (choice == 6 ? 128 : 64);
      GemmDims d = make_dims<256, 128>(M, N, std::max(K, 0), batch, ksplit, bk, false);
      d.group = tile_group(d);
      if (bk == 128)
        XG_HIP((launch_gemm_bf16_wave128<128, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      else
        XG_HIP((launch_gemm_bf16_wave128<64, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      return;
    }
    if ((choice == 7 || choice == 8) && bf16_vec(A, B, K, s_bs)) {
      const int bk = split ? 64 : (choice == 8 ? 128 : 64);
      GemmDims d = make_dims<128, 128>(M, N, std::max(K, 0), batch, ksplit, bk, lower);
      d.group = lower ? lower_tile_group(d) : tile_group(d);
      if (bk == 128)
        XG_HIP((launch_gemm_bf16_db<128, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      else
        XG_HIP((launch_gemm_bf16_db<64, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      return;
    }
    // default for plain BF16: 64-wide K steps with the registers capped for three waves per SIMD
    // (148 VGPRs, 37 KB LDS: three workgroups per CU) -- latency hiding by occupancy
    if (choice == 1 && !split && bf16_vec(A, B, K, s_bs)) {
      GemmDims d = make_dims<128, 128>(M, N, std::max(K, 0), batch, ksplit, 64, lower);
      d.group = lower ? lower_tile_group(d) : tile_group(d);
      XG_HIP((launch_gemm_bf16_occ<3, 64, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      return;
    }
    if (choice >= 9 && choice <= 13 && bf16_vec(A, B, K, s_bs)) {
      GemmDims d = make_dims<128, 128>(M, N, std::max(K, 0), batch, ksplit, choice >= 12 ? 32 : 64, lower);
      d.group = lower ? lower_tile_group(d) : tile_group(d);
      if (choice == 9) XG_HIP((launch_gemm_bf16_occ<2, 64, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      else if (choice == 10) XG_HIP((launch_gemm_bf16_occ<3, 64, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      else if (choice == 11) XG_HIP((launch_gemm_bf16_occ<4, 64, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      else if (choice == 12) XG_HIP((launch_gemm_bf16_occ<3, 32, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      else XG_HIP((launch_gemm_bf16_occ<4, 32, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
      return;
    }
    if ((choice == 1 || choice >= 3) && bf16_vec(A, B, K, s_bs)) {
      GemmDims d = make_dims<128, 128>(M, N, std::max(K, 0), batch, ksplit, bf16_wide_kstep(split), lower);
      d.group = lower ? lower_tile_group(d) : tile_group(d);
      XG_HIP((launch_gemm_bf16<128, 128, AK, BK_>(st, d, A, B, s, s_bs, split, epi, true)));
      return;
    }
    if (choice == 2 && split && !s && !lower && pipe_ok(A, B, K)) {
      if ((long)M * N >= 256L * 128 * 512) {
        GemmDims d = make_dims<256, 128>(M, N, std::max(K, 0), batch, ksplit, 32, false);
        XG_HIP((launch_gemm_bf16x3_pipe<256, 128, 3>(st, d, A, B, epi)));
      } else {
        GemmDims d = make_dims<128, 128>(M, N, std::max(K, 0), batch, ksplit, 32, false);
        XG_HIP((launch_gemm_bf16x3_pipe<128, 128, 4>(st, d, A, B, epi)));
      }
      return;
    }
  }
  GemmDims d = make_dims<128, 128>(M, N, std::max(K, 0), batch, ksplit, 32, lower);
  XG_HIP((launch_gemm_bf16<128, 128, AK, BK_>(st, d, A, B, s, s_bs, split, epi)));
}

// Number of K splits so that tiles*splits gives about four workgroups per CU, with K chunks
// of at least 512 so each workgroup amortises its launch and epilogue.
int pick_ksplit(long tiles, long K, int bk) {
  long want = std::max<long>(1, 1024 / std::max<long>(1, tiles));
  long maxs = std::max<long>(1, K / std::max(512, 4 * bk));
  return (int)std::max<long>(1, std::min(want, maxs));
}

void fill_zero(hipStream_t st, double* p, size_t n) {
  launch_map(st, n, [=] __device__(size_t i) { p[i] = 0.0; });
}

void split_to_bf16(hipStream_t st, const double* src, unsigned short* hi, unsigned short* lo, size_t n) {
  launch_map(st, n, [=] __device__(size_t i) {
    unsigned short h, l;
    split_bf16(src[i], h, l);
    hi[i] = h;
    if (lo) lo[i] = l;
  });
}

}  // namespace

// =================================================================== context / memory
extern "C" const char* xtpgpu_last_error(void) { return g_last_error.c_str(); }

extern "C" int xtpgpu_device_count(int* count) {
  return guarded([&] {
    XG_REQUIRE(count, XTPGPU_ERR_ARG, "null count");
    XG_HIP(hipGetDeviceCount(count));
  });
}

// ------------------------------------------------------------------ no-alloc scratch
// Device scratch of xtpgpu_bse_matmul_device_noalloc: one copy per GPU, allocated when the
// library's code object is loaded on that GPU and held until the process ends. In host code the
// symbol is only a stand-in; kernels get the per-device address that hipGetSymbolAddress returns.
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_dev;  // per device id; null until resolved

// Resolve the scratch address on `device` (which must be current) once per device.
void resolve_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize(device + 1, nullptr);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (rc != hipSuccess || p == nullptr)
    throw Error(XTPGPU_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
  g_scratch_dev[device] = static_cast<unsigned char*>(p);
}

unsigned char* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  XG_REQUIRE(device >= 0 && device < (int)g_scratch_dev.size() && g_scratch_dev[device], XTPGPU_ERR_ARG,
             "no-alloc scratch not resolved on this device (create a context on it first)");
  return g_scratch_dev[device];
}
}  // namespace

extern "C" int xtpgpu_ctx_create(int device, unsigned flags, xtpgpu_ctx* out) {
  return guarded([&] {
    XG_REQUIRE(out && flags == 0, XTPGPU_ERR_ARG, "xtpgpu_ctx_create: bad arguments");
    int n = 0;
    XG_HIP(hipGetDeviceCount(&n));
    XG_REQUIRE(device >= 0 && device < n, XTPGPU_ERR_ARG, "xtpgpu_ctx_create: device out of range");
    XG_HIP(hipSetDevice(device));
    resolve_scratch(device);
    auto c = std::make_unique<xtpgpu_ctx_s>();
    c->device = device;
    XG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c.release();
  });
}

extern "C" int xtpgpu_ctx_destroy(xtpgpu_ctx ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    ctx->collect();
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->blas) (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    delete ctx;
  });
}

extern "C" int xtpgpu_ctx_synchronize(xtpgpu_ctx ctx) {
  return guarded([&] {
    XG_REQUIRE(ctx, XTPGPU_ERR_ARG, "null ctx");
    XG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

// The context's grow-only scratch (the ε builds' scaled transition copies and K-split partials, up
// to the scale budget plus the partials) is kept between calls; this returns it to the device.
extern "C" int xtpgpu_ctx_release_scratch(xtpgpu_ctx ctx) {
  return guarded([&] {
    XG_REQUIRE(ctx, XTPGPU_ERR_ARG, "null ctx");
    XG_HIP(hipSetDevice(ctx->device));
    XG_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& b : ctx->ws) b.release();
  });
}

extern "C" int xtpgpu_ctx_set_profiling(xtpgpu_ctx ctx, int enable) {
  return guarded([&] {
    XG_REQUIRE(ctx, XTPGPU_ERR_ARG, "null ctx");
    ctx->profiling = enable != 0;
  });
}

extern "C" int xtpgpu_ctx_timer(xtpgpu_ctx ctx, const char* name, double* total_ms, int64_t* launches, double* flops,
                                double* bytes) {
  return guarded([&] {
    XG_REQUIRE(ctx && name, XTPGPU_ERR_ARG, "null argument");
    XG_HIP(hipSetDevice(ctx->device));
    XG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->collect();
    KTimer& t = ctx->timer(name);
    if (total_ms) *total_ms = t.ms;
    if (launches) *launches = t.launches;
    if (flops) *flops = t.flops;
    if (bytes) *bytes = t.bytes;
  });
}

extern "C" int xtpgpu_ctx_reset_timers(xtpgpu_ctx ctx) {
  return guarded([&] {
    XG_REQUIRE(ctx, XTPGPU_ERR_ARG, "null ctx");
    XG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->collect();
    ctx->timers.clear();
  });
}

extern "C" int xtpgpu_malloc(xtpgpu_ctx ctx, int64_t bytes, void** dptr) {
  return guarded([&] {
    XG_REQUIRE(ctx && dptr && bytes >= 0, XTPGPU_ERR_ARG, "xtpgpu_malloc: bad arguments");
    XG_HIP(hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    XG_HIP(hipMemGetInfo(&fr, &tot));
    XG_REQUIRE((size_t)bytes <= fr, XTPGPU_ERR_OOM, "xtpgpu_malloc: request exceeds free device memory");
    XG_HIP(hipMalloc(dptr, (size_t)std::max<int64_t>(bytes, 1)));
  });
}

extern "C" int xtpgpu_free(xtpgpu_ctx ctx, void* dptr) {
  return guarded([&] {
    XG_REQUIRE(ctx, XTPGPU_ERR_ARG, "null ctx");
    XG_HIP(hipStreamSynchronize(ctx->stream));
    if (dptr) XG_HIP(hipFree(dptr));
  });
}

extern "C" int xtpgpu_memcpy_h2d(xtpgpu_ctx ctx, void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    XG_REQUIRE(ctx && bytes >= 0, XTPGPU_ERR_ARG, "bad arguments");
    XG_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    XG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int xtpgpu_memcpy_d2h(xtpgpu_ctx ctx, void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    XG_REQUIRE(ctx && bytes >= 0, XTPGPU_ERR_ARG, "bad arguments");
    XG_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    XG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

namespace {
__global__ __launch_bounds__(256) void copy16_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                     size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}
}  // namespace

extern "C" int xtpgpu_copy_device(xtpgpu_ctx ctx, void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    XG_REQUIRE(ctx && dst && src && bytes >= 0 && bytes % 16 == 0 && (uintptr_t)dst % 16 == 0 &&
                   (uintptr_t)src % 16 == 0,
               XTPGPU_ERR_ARG, "xtpgpu_copy_device: pointers and size must be 16-byte multiples");
    XG_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)bytes / 16;
    if (n == 0) return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(copy16_kernel, dim3(grid), dim3(256), 0, ctx->stream, (uint4*)dst, (const uint4*)src, n);
    XG_HIP(hipGetLastError());
    XG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

// =================================================================== TCMatrix_gwbse
extern "C" int xtpgpu_mmn_create(xtpgpu_ctx ctx, int64_t mmin, int64_t mmax, int64_t nmin,
                                 int64_t nmax, int64_t auxsize, xtpgpu_mmn* out) {
  return guarded([&] {
    XG_REQUIRE(ctx && out, XTPGPU_ERR_ARG, "xtpgpu_mmn_create: null argument");
    XG_REQUIRE(mmax >= mmin && nmax >= nmin && auxsize > 0 && mmin >= 0 && nmin >= 0, XTPGPU_ERR_SHAPE,
               "xtpgpu_mmn_create: invalid index ranges");
    XG_HIP(hipSetDevice(ctx->device));
    auto t = std::make_unique<xtpgpu_mmn_s>();
    t->ctx = ctx;
    t->mmin = mmin; t->mmax = mmax; t->nmin = nmin; t->nmax = nmax;
    t->mtotal = mmax - mmin + 1;
    t->ntotal = nmax - nmin + 1;
    t->aux = auxsize;
    size_t n = (size_t)t->mtotal * t->ntotal * t->aux;
    t->data.alloc(n * sizeof(double));
    fill_zero(ctx->stream, t->d(), n);
    XG_HIP(hipStreamSynchronize(ctx->stream));
    *out = t.release();
  });
}

extern "C" int xtpgpu_mmn_destroy(xtpgpu_mmn mmn) {
  return guarded([&] {
    if (!mmn) return;
    (void)hipSetDevice(mmn->ctx->device);
    (void)hipStreamSynchronize(mmn->ctx->stream);
    delete mmn;
  });
}

extern "C" int xtpgpu_mmn_sizes(xtpgpu_mmn mmn, int64_t* mtotal, int64_t* ntotal, int64_t* auxsize) {
  return guarded([&] {
    XG_REQUIRE(mmn, XTPGPU_ERR_ARG, "null mmn");
    if (mtotal) *mtotal = mmn->mtotal;
    if (ntotal) *ntotal = mmn->ntotal;
    if (auxsize) *auxsize = mmn->aux;
  });
}

extern "C" int xtpgpu_mmn_set_slice(xtpgpu_mmn mmn, int64_t m, const double* slice, int64_t ld) {
  return guarded([&] {
    XG_REQUIRE(mmn && slice, XTPGPU_ERR_ARG, "null argument");
    XG_REQUIRE(m >= 0 && m < mmn->mtotal && ld >= mmn->ntotal, XTPGPU_ERR_SHAPE, "xtpgpu_mmn_set_slice: bad slice index / ld");
    xtpgpu_ctx ctx = mmn->ctx;
    XG_HIP(hipSetDevice(ctx->device));
    size_t n = (size_t)ld * mmn->aux;
    ctx->ws[0].ensure(n * sizeof(double));
    double* st = ctx->ws[0].as<double>();
    XG_HIP(hipMemcpyAsync(st, slice, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    double* D = mmn->d();
    long nt = mmn->ntotal, ps = (long)mmn->pstride();
    long ldl = ld;
    launch_map(ctx->stream, (size_t)nt * mmn->aux, [=] __device__(size_t i) {
      long p = (long)(i / nt), r = (long)(i % nt);
      D[p * ps + m * nt + r] = st[p * ldl + r];
    });
    XG_HIP(hipStreamSynchronize(ctx->stream));
    mmn->version++;
  });
}

extern "C" int xtpgpu_mmn_get_slice(xtpgpu_mmn mmn, int64_t m, double* slice, int64_t ld) {
  return guarded([&] {
    XG_REQUIRE(mmn && slice, XTPGPU_ERR_ARG, "null argument");
    XG_REQUIRE(m >= 0 && m < mmn->mtotal && ld >= mmn->ntotal, XTPGPU_ERR_SHAPE, "xtpgpu_mmn_get_slice: bad slice index / ld");
    xtpgpu_ctx ctx = mmn->ctx;
    XG_HIP(hipSetDevice(ctx->device));
    size_t n = (size_t)ld * mmn->aux;
    ctx->ws[0].ensure(n * sizeof(double));
    double* st = ctx->ws[0].as<double>();
    const double* D = mmn->d();
    long nt = mmn->ntotal, ps = (long)mmn->pstride();
    long ldl = ld;
    launch_map(ctx->stream, n, [=] __device__(size_t i) {
      long p = (long)(i / ldl), r = (long)(i % ldl);
      st[i] = r < nt ? D[p * ps + m * nt + r] : 0.0;
    });
    XG_HIP(hipMemcpyAsync(slice, st, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    XG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int xtpgpu_mmn_multiply_right(xtpgpu_mmn mmn, const double* A, int64_t lda) {
  return guarded([&] {
    XG_REQUIRE(mmn && A, XTPGPU_ERR_ARG, "null argument");
    XG_REQUIRE(lda >= mmn->aux, XTPGPU_ERR_SHAPE, "xtpgpu_mmn_multiply_right: lda < auxsize");
    xtpgpu_ctx ctx = mmn->ctx;
    XG_HIP(hipSetDevice(ctx->device));
    const long P = mmn->aux;
    const long MN = (long)mmn->pstride();
    DevBuf dA((size_t)lda * P * sizeof(double));
    XG_HIP(hipMemcpyAsync(dA.p, A, (size_t)lda * P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    DevBuf out((size_t)P * MN * sizeof(double));
    // new[Q][mn] = sum_P A(P,Q) old[P][mn]  ->  C(m=Q, n=mn), A_op(Q,P) = A[P + Q*lda] (K-contiguous),
    // B_op(P, mn) = old[P*MN + mn] (N-contiguous).
    for (long n0 = 0; n0 < MN; n0 += (1L << 30)) {
      int nn = (int)std::min<long>(MN - n0, 1L << 30);
      gemm64<true, false>(ctx->stream, (int)P, nn, (int)P, 1, 1, Opnd<double>{dA.as<double>(), nullptr, lda, 0},
                          Opnd<double>{mmn->d() + n0, nullptr, MN, 0}, nullptr, 0, false,
                          EpiStore{out.as<double>() + n0, MN, 1, 0, 1.0, 0.0});
    }
    XG_HIP(hipStreamSynchronize(ctx->stream));
    mmn->data = std::move(out);
    mmn->version++;
  });
}

extern "C" int xtpgpu_mmn_fill_ao2mo(xtpgpu_mmn mmn, const double* ao3c, int64_t nbf, const double* C,
                                     int64_t ldc, int64_t nmo) {
  return guarded([&] {
    XG_REQUIRE(mmn && ao3c && C, XTPGPU_ERR_ARG, "null argument");
    XG_REQUIRE(nbf > 0 && ldc >= nbf && nmo > mmn->mmax && nmo > mmn->nmax, XTPGPU_ERR_SHAPE,
               "xtpgpu_mmn_fill_ao2mo: MO window outside the coefficient matrix");
    xtpgpu_ctx ctx = mmn->ctx;
    XG_HIP(hipSetDevice(ctx->device));
    const long P = mmn->aux, nb = nbf, mt = mmn->mtotal, nt = mmn->ntotal;
    DevBuf dT((size_t)P * nb * nb * sizeof(double));
    DevBuf dC((size_t)ldc * nmo * sizeof(double));
    XG_HIP(hipMemcpyAsync(dT.p, ao3c, (size_t)P * nb * nb * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    XG_HIP(hipMemcpyAsync(dC.p, C, (size_t)ldc * nmo * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    // step 1: H[Q](m, nu) = sum_mu Cm(mu, m) (Q|mu nu)   (batched over Q)
    DevBuf dH((size_t)P * mt * nb * sizeof(double));
    const double* Cm = dC.as<double>() + (long)mmn->mmin * ldc;
    const double* Cn = dC.as<double>() + (long)mmn->nmin * ldc;
    // A(m, mu) = Cm[mu + m*ldc] (K-contiguous); B(mu, nu) = T[Q][mu + nu*nb] (K-contiguous)
    gemm64<true, true>(ctx->stream, (int)mt, (int)nb, (int)nb, (int)P, 1, Opnd<double>{Cm, nullptr, ldc, 0},
                       Opnd<double>{dT.as<double>(), nullptr, nb, nb * nb}, nullptr, 0, false,
                       EpiStore{dH.as<double>(), nb, 1, mt * nb, 1.0, 0.0});
    // step 2: M[Q][m][n] = sum_nu H[Q](m, nu) Cn(nu, n)  ->  data[Q*mt*nt + m*nt + n]
    gemm64<true, true>(ctx->stream, (int)mt, (int)nt, (int)nb, (int)P, 1,
                       Opnd<double>{dH.as<double>(), nullptr, nb, mt * nb}, Opnd<double>{Cn, nullptr, ldc, 0},
                       nullptr, 0, false, EpiStore{mmn->d(), nt, 1, mt * nt, 1.0, 0.0});
    XG_HIP(hipStreamSynchronize(ctx->stream));
    mmn->version++;
  });
}

extern "C" int xtpgpu_mmn_fill_synthetic(xtpgpu_mmn mmn, uint64_t seed, double sigma0, double decay) {
  return guarded([&] {
    XG_REQUIRE(mmn, XTPGPU_ERR_ARG, "null mmn");
    xtpgpu_ctx ctx = mmn->ctx;
    XG_HIP(hipSetDevice(ctx->device));
    const long P = mmn->aux;
    std::vector<double> sig(P);
    for (long p = 0; p < P; ++p) sig[p] = sigma0 * std::exp(-decay * (double)p / (double)P);
    DevBuf ds(P * sizeof(double));
    XG_HIP(hipMemcpyAsync(ds.p, sig.data(), P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const double* s = ds.as<double>();
    double* D = mmn->d();
    const long mt = mmn->mtotal, nt = mmn->ntotal, mmin = mmn->mmin, nmin = mmn->nmin;
    const size_t n = (size_t)P * mt * nt;
    launch_map(ctx->stream, n, [=] __device__(size_t i) {
      long p = (long)(i / (size_t)(mt * nt));
      long r = (long)(i % (size_t)(mt * nt));
      long m = r / nt + mmin, q = r % nt + nmin;
      long a = m < q ? m : q, b = m < q ? q : m;
      D[i] = synth_normal(seed, (uint64_t)p, (uint64_t)a, (uint64_t)b) * s[p];
    });
    XG_HIP(hipStreamSynchronize(ctx->stream));
    mmn->version++;
  });
}

extern "C" int xtpgpu_mmn_copy(xtpgpu_mmn dst, xtpgpu_mmn src) {
  return guarded([&] {
    XG_REQUIRE(dst && src, XTPGPU_ERR_ARG, "xtpgpu_mmn_copy: null argument");
    XG_REQUIRE(dst->mtotal == src->mtotal && dst->ntotal == src->ntotal && dst->aux == src->aux, XTPGPU_ERR_SHAPE,
               "xtpgpu_mmn_copy: tensors of different shape");
    XG_REQUIRE(dst->ctx->device == src->ctx->device, XTPGPU_ERR_ARG, "xtpgpu_mmn_copy: tensors on different devices");
    XG_HIP(hipSetDevice(dst->ctx->device));
    // order after pending work on both tensors' streams
    XG_HIP(hipStreamSynchronize(src->ctx->stream));
    XG_HIP(hipMemcpyAsync(dst->d(), src->d(), src->pstride() * (size_t)src->aux * 8, hipMemcpyDeviceToDevice,
                          dst->ctx->stream));
    XG_HIP(hipStreamSynchronize(dst->ctx->stream));
    dst->version++;
  });
}

extern "C" int xtpgpu_mmn_set_block(xtpgpu_mmn mmn, int64_t vmin, int64_t vtotal, int64_t cmin, int64_t ctotal,
                                    const double* m, int symmetric) {
  return guarded([&] {
    XG_REQUIRE(mmn && m, XTPGPU_ERR_ARG, "xtpgpu_mmn_set_block: null argument");
    XG_REQUIRE(vmin >= 0 && cmin >= 0 && vtotal > 0 && ctotal > 0 && vmin + vtotal <= mmn->mtotal &&
                   cmin + ctotal <= mmn->ntotal &&
                   (!symmetric || (cmin + ctotal <= mmn->mtotal && vmin + vtotal <= mmn->ntotal)),
               XTPGPU_ERR_SHAPE, "xtpgpu_mmn_set_block: block outside the tensor");
    xtpgpu_ctx ctx = mmn->ctx;
    XG_HIP(hipSetDevice(ctx->device));
    const long P = mmn->aux, n = (long)P * vtotal * ctotal;
    DevBuf b(n * 8);
    XG_HIP(hipMemcpyAsync(b.p, m, n * 8, hipMemcpyHostToDevice, ctx->stream));
    const double* s = b.as<double>();
    double* D = mmn->d();
    const long nt = mmn->ntotal, ps = (long)mmn->pstride(), vt = vtotal, ct = ctotal, v0 = vmin, c0 = cmin;
    const bool sym = symmetric != 0;
    launch_map(ctx->stream, (size_t)n, [=] __device__(size_t i) {
      long p = (long)(i / (vt * ct)), r = (long)(i % (vt * ct));
      long v = r / ct, c = r % ct;
      D[p * ps + (v0 + v) * nt + c0 + c] = s[i];
      if (sym) D[p * ps + (c0 + c) * nt + v0 + v] = s[i];
    });
    XG_HIP(hipStreamSynchronize(ctx->stream));
    mmn->version++;
  });
}

// =================================================================== RPA
namespace {

struct RpaPacked {
  DevBuf A64;          // [P][Kp] FP64
  DevBuf Ah, Al;       // [P][Kp] bf16 hi / lo
  long K = 0;          // transitions n_occ * n_unocc
  long Kp = 0;         // row length: K rounded up to 64 (8 below 64) for line-aligned rows, zero pads
};

void rpa_pack(xtpgpu_mmn mmn, long n_occ, long n_unocc, bool need64, bool need16, RpaPacked& pk) {
  hipStream_t st = mmn->ctx->stream;
  const long P = mmn->aux, K = n_occ * n_unocc, nt = mmn->ntotal, ps = (long)mmn->pstride();
  const long Kp = K >= 64 ? (K + 63) / 64 * 64 : (K + 7) / 8 * 8;  // 128-byte aligned rows
  const long r0 = nt - n_unocc;
  pk.K = K;
  pk.Kp = Kp;
  const double* D = mmn->d();
  if (need64) {
    pk.A64.alloc((size_t)P * Kp * sizeof(double));
    double* A = pk.A64.as<double>();
    launch_map(st, (size_t)P * Kp, [=] __device__(size_t i) {
      long p = (long)(i / Kp), k = (long)(i % Kp);
      long v = k / n_unocc, c = k % n_unocc;
      A[i] = k < K ? D[p * ps + v * nt + r0 + c] : 0.0;
    });
  }
  if (need16) {
    pk.Ah.alloc((size_t)P * Kp * 2);
    pk.Al.alloc((size_t)P * Kp * 2);
    unsigned short* H = pk.Ah.as<unsigned short>();
    unsigned short* L = pk.Al.as<unsigned short>();
    launch_map(st, (size_t)P * Kp, [=] __device__(size_t i) {
      long p = (long)(i / Kp), k = (long)(i % Kp);
      long v = k / n_unocc, c = k % n_unocc;
      unsigned short h = 0, l = 0;
      if (k < K) split_bf16(D[p * ps + v * nt + r0 + c], h, l);
      H[i] = h;
      L[i] = l;
    });
  }
}

// Bytes of scaled transition copies one mixed-precision eps build may hold at once: 12 GiB, or
// XTPGPU_EPS_SCALE_BYTES when set (a small value forces one frequency per group).
size_t eps_scale_budget() {
  const char* e = std::getenv("XTPGPU_EPS_SCALE_BYTES");
  if (e && *e) {
    const long long v = std::atoll(e);
    if (v > 0) return (size_t)v;
  }
  return (size_t)12 << 30;
}

// FP64 eps with the frequency weights in scaled copies (default: 43.7 against 37.3 TFLOP/s for the
// weights applied on the B-operand load at the single-segment shapes; XTPGPU_EPS64_SCALED=0 selects
// the latter).
bool eps64_scaled() {
  const char* e = std::getenv("XTPGPU_EPS64_SCALED");
  return !(e && std::strcmp(e, "0") == 0);
}

// K splits of a lower-triangle eps GEMM so that its workgroups fill whole rounds of the resident
// slots (`resident` = workgroups the kernel keeps on the chip at once): 136 tiles x 12 frequencies at
// P = 2000 are 3.2 rounds of 512 slots, the last one a fifth full; three splits make 9.6 of 10.
int eps_ksplit(long units, long resident, long K, int bk) {
  if (units <= 0 || units >= 8 * resident) return 1;
  int best = 1;
  double best_eff = 0.0;
  for (int ks = 1; ks <= 4; ++ks) {
    if (ks > 1 && K / ks < 16L * bk) break;
    const double rounds = (double)(units * ks) / (double)resident;
    const double eff = rounds / std::ceil(rounds);
    if (eff > best_eff + 0.02) {
      best_eff = eff;
      best = ks;
    }
  }
  return best;
}

// eps[b] = sum over the K splits of the partials + 1, written to both triangles (the partials hold
// the lower-triangle tiles: element (m, n), m >= n, of split q at part[((q*nb + b)*P + n)*P + m])
template <typename T>
void eps_reduce_splits(hipStream_t st, const T* part, int ks, long nb, long P, double* out) {
  launch_map(st, (size_t)nb * P * P, [=] __device__(size_t i) {
    const long b = (long)(i / (size_t)(P * P)), r = (long)(i % (size_t)(P * P));
    const long n = r / P, m = r % P;
    if (m < n) return;
    double acc = 0.0;
    for (int q = 0; q < ks; ++q) acc += (double)part[(((long)q * nb + b) * P + n) * P + m];
    const double v = acc + (m == n ? 1.0 : 0.0);
    double* o = out + b * P * P;
    o[n * P + m] = v;
    o[m * P + n] = v;
  });
}

void rpa_epsilon_impl(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo, int64_t rpamin,
                      int64_t rpamax, int kind, const double* omega_re, const double* omega_im, int n_omega,
                      double eta, int precision, double* out_dev) {
  XG_REQUIRE(ctx && mmn && energies && omega_re && out_dev, XTPGPU_ERR_ARG, "xtpgpu_rpa_epsilon: null argument");
  XG_REQUIRE(n_omega > 0, XTPGPU_ERR_ARG, "xtpgpu_rpa_epsilon: n_omega must be positive");
  XG_REQUIRE(kind >= 0 && kind <= 2, XTPGPU_ERR_ARG, "xtpgpu_rpa_epsilon: unknown kind");
  XG_REQUIRE(precision == XTPGPU_FP64 || precision == XTPGPU_BF16 || precision == XTPGPU_BF16X3,
             XTPGPU_ERR_ARG, "xtpgpu_rpa_epsilon: unknown precision");
  const long lumo = homo + 1;
  const long n_occ = lumo - rpamin;
  const long n_unocc = rpamax - lumo + 1;
  XG_REQUIRE(n_occ > 0 && n_unocc > 0, XTPGPU_ERR_SHAPE, "xtpgpu_rpa_epsilon: empty occupied or virtual range");
  XG_REQUIRE(n_occ <= mmn->mtotal && n_unocc <= mmn->ntotal, XTPGPU_ERR_SHAPE,
             "xtpgpu_rpa_epsilon: RPA window exceeds the Mmn tensor");
  XG_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const long P = mmn->aux;
  const long rpatotal = rpamax - rpamin + 1;
  const long K = n_occ * n_unocc;
  RpaPacked pk;
  const bool use16 = precision != XTPGPU_FP64;
  rpa_pack(mmn, n_occ, n_unocc, !use16, use16, pk);

  // weights d[w][K] (rpa.cc:110-125, 178-190; the complex kind carries the -2 of rpa.cc:197)
  DevBuf de(rpatotal * sizeof(double)), dw(2 * n_omega * sizeof(double));
  XG_HIP(hipMemcpyAsync(de.p, energies, rpatotal * sizeof(double), hipMemcpyHostToDevice, st));
  std::vector<double> w(2 * n_omega, 0.0);
  for (int i = 0; i < n_omega; ++i) {
    w[i] = omega_re[i];
    w[n_omega + i] = omega_im ? omega_im[i] : 0.0;
  }
  XG_HIP(hipMemcpyAsync(dw.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, st));
  const long Kp = pk.Kp;
  DevBuf dd((size_t)n_omega * Kp * (use16 ? sizeof(float) : sizeof(double)));
  const double* E = de.as<double>();
  const double* W = dw.as<double>();
  double* d64 = use16 ? nullptr : dd.as<double>();
  float* d32 = use16 ? dd.as<float>() : nullptr;
  const int nw = n_omega;
  launch_map(st, (size_t)n_omega * Kp, [=] __device__(size_t i) {
    long f = (long)(i / Kp), k = (long)(i % Kp);
    if (k >= K) {  // pad transitions carry zero weight
      if (d64) d64[i] = 0.0; else d32[i] = 0.0f;
      return;
    }
    long v = k / n_unocc, c = k % n_unocc;
    double D = E[n_occ + c] - E[v];
    double wr = W[f], wi = W[nw + f];
    double r;
    if (kind == XTPGPU_EPS_IMAG) {
      r = 4.0 * D / (D * D + wr * wr);
    } else if (kind == XTPGPU_EPS_REAL) {
      double a = D - wr, b = D + wr;
      r = 2.0 * (a / (a * a + eta * eta) + b / (b * b + eta * eta));
    } else {
      double dm = wr - D, dp = wr + D;
      double s1 = (wi + eta) * (wi + eta), s2 = (wi - eta) * (wi - eta);
      r = -2.0 * (dm / (dm * dm + s1) - dp / (dp * dp + s2));
    }
    if (d64) d64[i] = r; else d32[i] = (float)r;
  });

  EpiEpsSym epi{out_dev, P};
  int cus = 256;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) cus = prop.multiProcessorCount;
  }
  // algorithmic FLOPs: n_omega * K * P * (P+1) (symmetric rank-K update); bytes: M read once + eps written
  ProfScope prof(ctx, "rpa_eps_gemm", (double)n_omega * K * P * (P + 1),
                 (double)P * K * (use16 ? (precision == XTPGPU_BF16X3 ? 4 : 2) : 8) + (double)n_omega * P * P * 8);
  if (!use16 && !eps64_scaled()) {
    Opnd<double> A{pk.A64.as<double>(), nullptr, Kp, 0};
    gemm64<true, true>(st, (int)P, (int)P, (int)Kp, n_omega, 1, A, A, d64, Kp, true, epi);
  } else if (!use16) {
    // FP64 with the weights in scaled copies B_w = M diag(d_w), as the mixed forms do
    const size_t plane = (size_t)P * Kp;
    const size_t per_w = plane * 8;
    const long gmax = std::max<long>(1, std::min<long>(n_omega, (long)(eps_scale_budget() / per_w)));
    const long ngroups = (n_omega + gmax - 1) / gmax;
    const long gsz = (n_omega + ngroups - 1) / ngroups;
    DevBuf& sc = ctx->ws[1];  // kept on the context: a multi-GB allocation and free per build costs more than the build
    sc.ensure(per_w * gsz);
    double* Bs = sc.as<double>();
    const double* A64 = pk.A64.as<double>();
    const Opnd<double> A{A64, nullptr, Kp, 0};
    for (long w0 = 0; w0 < n_omega; w0 += gsz) {
      const long nw_g = std::min<long>(gsz, n_omega - w0);
      const double* dg = d64 + w0 * Kp;
      const long Kpl = Kp;
      launch_map(st, plane, [=] __device__(size_t i) {
        const double a = A64[i];
        const long k = (long)(i % (size_t)Kpl);
        for (long f = 0; f < nw_g; ++f) Bs[f * plane + i] = a * dg[f * Kpl + k];
      });
      EpiEpsSym eg{out_dev + w0 * P * P, P};
      const Opnd<double> Bo{Bs, nullptr, Kp, (long)plane};
      const long tl = (P + 127) / 128;
      const int ks = P >= 256 ? eps_ksplit(tl * (tl + 1) / 2 * nw_g, 2L * cus, Kp, 16) : 1;
      if (ks > 1) {  // 128 x 128 two-wave tiles over K splits, FP64 partials, one reduction
        GemmDims d = make_dims<128, 128>((int)P, (int)P, (int)Kp, (int)nw_g, ks, 16, true);
        d.group = lower_tile_group(d);
        DevBuf& pb = ctx->ws[2];
        pb.ensure((size_t)d.ksplit * nw_g * P * P * 8);
        XG_HIP((launch_gemm_f64<128, 128, true, true, EpiPartial, 2>(
            st, d, A, Bo, nullptr, 0, EpiPartial{pb.as<double>(), (int)P, (int)P, (int)nw_g})));
        eps_reduce_splits(st, pb.as<double>(), d.ksplit, nw_g, P, out_dev + w0 * P * P);
      } else {
        gemm64<true, true>(st, (int)P, (int)P, (int)Kp, (int)nw_g, 1, A, Bo, nullptr, 0, true, eg);
      }
    }
  } else {
    // the weights go into a scaled copy of the transitions, B_w = M diag(d_w) as BF16 hi (+ lo),
    // built once per frequency, so the GEMM's inner loop is plain MFMA work; frequencies are
    // taken in equal groups whose scaled copies fit eps_scale_budget()
    const bool split = precision == XTPGPU_BF16X3;
    const size_t plane = (size_t)P * Kp;
    const size_t per_w = plane * 2 * (split ? 2 : 1);
    const long gmax = std::max<long>(1, std::min<long>(n_omega, (long)(eps_scale_budget() / per_w)));
    const long ngroups = (n_omega + gmax - 1) / gmax;
    const long gsz = (n_omega + ngroups - 1) / ngroups;
    DevBuf& sc = ctx->ws[1];
    sc.ensure(per_w * gsz);
    unsigned short* Bh = sc.as<unsigned short>();
    unsigned short* Bl = split ? Bh + plane * gsz : nullptr;
    const unsigned short* Ah = pk.Ah.as<unsigned short>();
    const unsigned short* Al = pk.Al.as<unsigned short>();
    const Opnd<unsigned short> A{Ah, Al, Kp, 0};
    for (long w0 = 0; w0 < n_omega; w0 += gsz) {
      const long nw_g = std::min<long>(gsz, n_omega - w0);
      const float* dg = d32 + w0 * Kp;
      const long Kpl = Kp;
      const size_t plane8 = plane / 8;  // Kp is a multiple of 8: 8-element vectors stay in one row
      // one 16-byte vector of transitions per thread, read once and scaled for every frequency
      // of the group
      launch_map(st, plane8, [=] __device__(size_t i) {
        const us8v h = reinterpret_cast<const us8v*>(Ah)[i], l = reinterpret_cast<const us8v*>(Al)[i];
        float a[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = bf16_bits_to_f(h[e]) + bf16_bits_to_f(l[e]);
        const long k0 = (long)((i * 8) % (size_t)Kpl);
        for (long f = 0; f < nw_g; ++f) {
          const float4 d0 = *reinterpret_cast<const float4*>(dg + f * Kpl + k0);
          const float4 d1 = *reinterpret_cast<const float4*>(dg + f * Kpl + k0 + 4);
          const float dv[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
          us8v oh, ol;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const float x = a[e] * dv[e];
            oh[e] = bf16_bits_rn(x);
            ol[e] = bf16_bits_rn(x - bf16_bits_to_f(oh[e]));
          }
          reinterpret_cast<us8v*>(Bh)[f * plane8 + i] = oh;
          if (Bl) reinterpret_cast<us8v*>(Bl)[f * plane8 + i] = ol;
        }
      });
      EpiEpsSym eg{out_dev + w0 * P * P, P};
      const Opnd<unsigned short> Bo{Bh, Bl, Kp, (long)plane};
      const long tl = (P + 127) / 128;
      // the default kernels keep 2 (split) or 3 (plain) 128 x 128 workgroups per CU resident
      const int ks = (P >= 256 && bf16_kernel_choice() == 1 && bf16_vec(A, Bo, (int)Kp, 0))
                         ? eps_ksplit(tl * (tl + 1) / 2 * nw_g, (split ? 2L : 3L) * cus, Kp, 64) : 1;
      // split form from P = 1536: 256 x 256 tiles whose operands go straight to LDS (one
      // workgroup per CU), K split to fill its rounds -- half the panel traffic per FLOP of
      // 128 x 128 tiles, which is what bounds this GEMM (220 against 206 TFLOP/s at P = 2000, 280
      // against 229 at P = 6000); below, the 128 x 128 tiles fill the chip better (a loop of the
      // 256 uneven segments' builds, P 248-2100, took 532 ms with the large tiles from P = 512
      // against 480 ms). XTPGPU_EPS_BF16_KERNEL=wide keeps the 128 x 128 tiles throughout.
      const char* ek = std::getenv("XTPGPU_EPS_BF16_KERNEL");
      const bool wide = ek && std::strcmp(ek, "wide") == 0;
      if (!wide && bf16_kernel_choice() == 1 && split && P >= 1536 && lds_ok(A, Bo, (int)Kp, true)) {
        const long t2 = (P + 255) / 256;
        const int k2 = eps_ksplit(t2 * (t2 + 1) / 2 * nw_g, (long)cus, Kp, 32);
        GemmDims d = make_dims<256, 256>((int)P, (int)P, (int)Kp, (int)nw_g, k2, 32, true);
        DevBuf& pb = ctx->ws[2];
        pb.ensure((size_t)d.ksplit * nw_g * P * P * 4);
        XG_HIP((launch_gemm_bf16_lds<256, 256, 2>(st, d, A, Bo, true, EpiPart32{pb.as<float>(), P, P, nw_g})));
        eps_reduce_splits(st, pb.as<float>(), d.ksplit, nw_g, P, out_dev + w0 * P * P);
      } else if (ks > 1) {  // the default tiles over K splits, FP32 partials, one reduction
        GemmDims d = make_dims<128, 128>((int)P, (int)P, (int)Kp, (int)nw_g, ks, 64, true);
        d.group = lower_tile_group(d);
        DevBuf& pb = ctx->ws[2];
        pb.ensure((size_t)d.ksplit * nw_g * P * P * 4);
        const EpiPart32 ep{pb.as<float>(), P, P, nw_g};
        if (split)
          XG_HIP((launch_gemm_bf16<128, 128, true, true>(st, d, A, Bo, nullptr, 0, true, ep, true)));
        else
          XG_HIP((launch_gemm_bf16_occ<3, 64, true, true>(st, d, A, Bo, nullptr, 0, false, ep)));
        eps_reduce_splits(st, pb.as<float>(), d.ksplit, nw_g, P, out_dev + w0 * P * P);
      } else {
        gemm16<true, true>(st, (int)P, (int)P, (int)Kp, (int)nw_g, 1, A, Bo, nullptr, 0, split, true, eg);
      }
    }
  }
  XG_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" int xtpgpu_rpa_epsilon_device(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo,
                                         int64_t rpamin, int64_t rpamax, int kind, const double* omega_re,
                                         const double* omega_im, int n_omega, double eta, int precision,
                                         double* out_dev) {
  return guarded([&] {
    rpa_epsilon_impl(ctx, mmn, energies, homo, rpamin, rpamax, kind, omega_re, omega_im, n_omega, eta, precision,
                     out_dev);
  });
}

extern "C" int xtpgpu_rpa_epsilon(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo,
                                  int64_t rpamin, int64_t rpamax, int kind, const double* omega_re,
                                  const double* omega_im, int n_omega, double eta, int precision, double* out) {
  return guarded([&] {
    XG_REQUIRE(ctx && mmn && out, XTPGPU_ERR_ARG, "xtpgpu_rpa_epsilon: null argument");
    const size_t n = (size_t)mmn->aux * mmn->aux * std::max(n_omega, 0);
    DevBuf o(n * sizeof(double));
    rpa_epsilon_impl(ctx, mmn, energies, homo, rpamin, rpamax, kind, omega_re, omega_im, n_omega, eta, precision,
                     o.as<double>());
    XG_HIP(hipMemcpyAsync(out, o.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    XG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

// =================================================================== BSE operator
// Device layouts (vtp / Pp = vtotal / auxsize rounded up to 8; ctp = ctotal rounded up to 64 when
// ctotal >= 64, else to 8: rows of ctp elements start on 128-byte lines in BF16 and FP64, so a
// 64-wide K step reads whole cache lines; pads zero; contractions over c' run to round8(ct)):
//   X   [(kv, v')][ctp]        the k input vectors, BSE index (v', c')
//   Mcc [P][c1][ctp]           M^P_{c1 c'}            Hd stage 1 (A, K = c')
//   Bvv [v1][P][vtp]           eps_P M^P_{v1 v'}      Hd stage 2 (B, K = (P, v'))
//   Mvc [P][v][ctp]            M^P_{v c}              Hx V^T X (A, K = (v,c)); Hd2 stage 1 (A, K = c')
//   MvcT[(v,c)][Pp]            M^P_{v c}              Hx V W (A, K = P)
//   McvT[c1][P][vtp]           M^P_{c1 v'}            Hd2 stage 2 (A, K = (P, v'))
// so that every GEMM reads both operands K-contiguous with 16-byte vectors.
struct xtpgpu_bseop_s {
  xtpgpu_ctx ctx = nullptr;
  int cqp = 0, cx = 0, cd = 0, cd2 = 0;
  long vt = 0, ct = 0, N = 0, P = 0;
  long ctp = 0, vtp = 0, Pp = 0;
  bool hqp_diag = true;
  unsigned mask = 1;
  DevBuf hc, hv;        // Hqp blocks (column-major ct x ct / vt x vt)
  DevBuf hcd, hvd;      // diagonals
  DevBuf eps;           // eps0^-1 (P)
  DevBuf diag;          // operator diagonal (N)
  DevBuf Mcc, Bvv, Mvc, MvcT, McvT;                                  // FP64
  DevBuf Mcc_h, Mcc_l, Bvv_h, Bvv_l, Mvc_h, Mvc_l, McvT_h, McvT_l;  // BF16 hi / lo
  // scratch
  DevBuf xp, xh, xl, s1, s1h, s1l, part, wbuf;
  size_t s1_bytes = 0;
  // streamed FP64 M_cc' (xtpgpu_bse_create_streamed with XTPGPU_BSE_STREAM_FP64): the source,
  // the tensor index of the first conduction level, and the chunk scratch (source window, packed)
  xtpgpu_mmn_chunk_fn src_fn = nullptr;
  void* src_user = nullptr;
  bool stream64 = false;
  bool screen_full = false;  // Bvv carries the full eps^-1(0) matrix (unrotated aux basis)
  long cmin_r = 0, vmin_r = 0;  // tensor indices of the first conduction / valence level
  DevBuf diag0;                 // the diagonal without its Hqp part (xtpgpu_bse_set_hqp)
  DevBuf srcbuf, mcc64;
};

namespace {

inline long round8(long x) { return (x + 7) & ~7L; }

// grow-only scratch whose unwritten pad positions must read as zero
void ensure_zeroed(DevBuf& b, size_t bytes, hipStream_t st) {
  if (bytes <= b.bytes) return;
  b.alloc(bytes);
  unsigned char* p = b.as<unsigned char>();
  launch_map(st, bytes, [=] __device__(size_t i) { p[i] = 0; });
}

// First-stage scratch budget (bytes) of the aux-chunked Hd / Hd2 contraction: 3 GiB, or
// XTPGPU_HD_CHUNK_BYTES when set (a small value forces several uneven aux chunks).
size_t hd_chunk_budget() {
  const char* e = std::getenv("XTPGPU_HD_CHUNK_BYTES");
  if (e && *e) {
    const long long v = std::atoll(e);
    if (v > 0) return (size_t)v;
  }
  return (size_t)3 << 30;
}

// Consecutive 256-byte aligned slices of a fixed device region: the scratch of one H X call
// when it must not allocate (the no-alloc entry point hands in the library's static region).
struct ScratchArena {
  unsigned char* base = nullptr;
  size_t cap = 0, used = 0;
  static size_t up(size_t x) { return (x + 255) & ~(size_t)255; }
  size_t left() const { return up(used) < cap ? cap - up(used) : 0; }
  void* take(size_t bytes) {
    const size_t off = up(used);
    if (off + bytes > cap)
      throw Error(XTPGPU_ERR_SHAPE, "H X needs " + std::to_string(off + bytes) + " bytes of scratch, more than the " +
                                        std::to_string(cap) + "-byte no-alloc region");
    used = off + bytes;
    return base + off;
  }
};

void zero_bytes(hipStream_t st, void* p, size_t bytes) {
  unsigned char* q = reinterpret_cast<unsigned char*>(p);
  launch_map(st, bytes, [=] __device__(size_t i) { q[i] = 0; });
}

// Scratch bytes of one H X call other than the stage-1 intermediate, and the stage-1 bytes per
// aux function (both planes), for the given precision and block width.
struct MatmulPlan {
  size_t xp = 0, xh = 0, xl = 0, part = 0, wbuf = 0, per_p = 0;
};

MatmulPlan plan_matmul(const xtpgpu_bseop_s* op, int precision, long k) {
  MatmulPlan pl;
  const long vt = op->vt, ct = op->ct, P = op->P, Np = vt * op->ctp;
  const bool m16 = precision != XTPGPU_FP64, split = precision == XTPGPU_BF16X3;
  const bool need16 = m16 && (op->cd || op->cd2);
  pl.xp = (size_t)Np * k * 8;
  if (need16) {
    pl.xh = (size_t)Np * k * 2;
    if (split) pl.xl = (size_t)Np * k * 2;
  }
  if (op->cd || op->cd2) pl.per_p = (size_t)std::max(ct, vt) * k * op->vtp * (m16 ? 2 * (split ? 2 : 1) : 8);
  // Hd stage-2 partials, at most 4 K splits (FP32; FP64 in the grouped products of the batched solve)
  if (op->cd) pl.part = (size_t)4 * k * ct * vt * (m16 ? 4 : 8);
  if (op->cx) {
    const int tiles = (int)(((P + 63) / 64) * ((k + 63) / 64));
    const GemmDims dd = make_dims<64, 64>((int)P, (int)k, (int)Np, 1, pick_ksplit(tiles, Np, 16), 16, false);
    pl.part = std::max(pl.part, (size_t)dd.ksplit * P * k * 8);
    pl.wbuf = (size_t)op->Pp * k * 8;
  }
  return pl;
}

// BSE_OPERATOR::matmul (bse_operator.cc:39-119) on device buffers. Scratch comes from the
// operator's grow-only buffers, or, with `arena`, from that fixed region only (no allocation,
// no profiling events; the stage-1 aux chunk shrinks to what the region has left).
void bse_matmul_impl(xtpgpu_bseop op, int precision, const double* X, long ldx, long k, double* Y, long ldy,
                     size_t budget = hd_chunk_budget(), ScratchArena* arena = nullptr, hipStream_t on = nullptr) {
  xtpgpu_ctx ctx = op->ctx;
  // ProfScope records on the context's stream: inert without a context, so off for an arena call
  // and for work issued on another stream
  xtpgpu_ctx pctx = (arena || on) ? nullptr : ctx;
  hipStream_t st = on ? on : ctx->stream;
  const long N = op->N, vt = op->vt, ct = op->ct, P = op->P;
  const long ctp = op->ctp, vtp = op->vtp, Pp = op->Pp;
  const long Np = vt * ctp;  // padded length of one input vector
  XG_REQUIRE(k >= 0 && ldx >= N && ldy >= N, XTPGPU_ERR_SHAPE, "bse matmul: leading dimension smaller than operator size");
  XG_REQUIRE(precision == XTPGPU_FP64 || ((op->mask >> precision) & 1u), XTPGPU_ERR_ARG,
             "bse matmul: operator was created without this precision");
  if (k == 0) return;
  const bool m16 = precision != XTPGPU_FP64;
  const bool split = precision == XTPGPU_BF16X3;
  const bool need16 = m16 && (op->cd || op->cd2);
  const MatmulPlan pl = plan_matmul(op, precision, k);
  auto buf = [&](DevBuf& b, size_t bytes) -> void* {
    if (bytes == 0) return nullptr;
    if (arena) return arena->take(bytes);
    b.ensure(bytes);
    return b.p;
  };
  // grow-only scratch whose never-written pad positions must read as zero: the operator's buffers
  // are cleared when they grow, an arena slice on every call (it holds the last call's data)
  auto zbuf = [&](DevBuf& b, size_t bytes) -> void* {
    if (bytes == 0) return nullptr;
    if (arena) {
      void* p = arena->take(bytes);
      zero_bytes(st, p, bytes);
      return p;
    }
    ensure_zeroed(b, bytes, st);
    return b.p;
  };

  // padded copies of X (FP64 always; BF16 hi/lo for the mixed contractions)
  double* Xp = static_cast<double*>(buf(op->xp, pl.xp));
  unsigned short* Xh = need16 ? static_cast<unsigned short*>(buf(op->xh, pl.xh)) : nullptr;
  unsigned short* Xl = need16 && split ? static_cast<unsigned short*>(buf(op->xl, pl.xl)) : nullptr;
  void* part_p = buf(op->part, pl.part);
  double* Wb = op->cx ? static_cast<double*>(zbuf(op->wbuf, pl.wbuf)) : nullptr;
  launch_map(st, (size_t)Np * k, [=] __device__(size_t i) {
    long col = (long)(i / (size_t)ctp), c = (long)(i % (size_t)ctp);  // col = kv*vt + v
    long kv = col / vt, v = col % vt;
    double x = c < ct ? X[kv * ldx + v * ct + c] : 0.0;
    Xp[i] = x;
    if (Xh) {
      unsigned short h, l;
      split_bf16(x, h, l);
      Xh[i] = h;
      if (Xl) Xl[i] = l;
    }
  });
  // Y = 0
  launch_map(st, (size_t)N * k, [=] __device__(size_t i) { Y[(i / N) * ldy + i % N] = 0.0; });

  // ---------------- Hqp (bse_operator.cc:121-132): cqp (H_c^T X - X H_v) per vector
  if (op->cqp) {
    ProfScope prof(pctx, "hqp", 2.0 * N * k * (op->hqp_diag ? 1.0 : (double)(ct + vt)), 8.0 * 3.0 * N * k);
    const double cqp = op->cqp;
    if (op->hqp_diag) {
      const double* hc = op->hcd.as<double>();
      const double* hv = op->hvd.as<double>();
      launch_map(st, (size_t)N * k, [=] __device__(size_t i) {
        long kv = (long)(i / N), r = (long)(i % N);
        long v = r / ct, c = r % ct;
        Y[kv * ldy + r] += cqp * (hc[c] - hv[v]) * Xp[(kv * vt + v) * ctp + c];
      });
    } else {
      // out_kv(c1, v1) += cqp * sum_c' Hc(c', c1) X_kv(c', v1)   (batched over kv)
      gemm64<true, true>(st, (int)ct, (int)vt, (int)ct, (int)k, 1, Opnd<double>{op->hc.as<double>(), nullptr, ct, 0},
                         Opnd<double>{Xp, nullptr, ctp, vt * ctp}, nullptr, 0, false,
                         EpiStore{Y, 1, ct, ldy, cqp, 1.0});
      // out_kv(c1, v1) -= cqp * sum_v' X_kv(c1, v') Hv(v', v1)
      gemm64<false, true>(st, (int)ct, (int)vt, (int)vt, (int)k, 1, Opnd<double>{Xp, nullptr, ctp, vt * ctp},
                          Opnd<double>{op->hv.as<double>(), nullptr, vt, 0}, nullptr, 0, false,
                          EpiStore{Y, 1, ct, ldy, -cqp, 1.0});
    }
  }

  // aux chunk for the two-stage Hd / Hd2 contraction (first-stage scratch <= budget, or what the
  // arena has left); the stage-1 buffers are sized once for the largest chunk
  long pc = P;
  unsigned char* s1a = nullptr;  // FP64 intermediate, or the BF16 hi plane
  unsigned char* s1b = nullptr;  // BF16 lo plane of the split form
  if (op->cd || op->cd2) {
    const size_t avail = arena ? (arena->left() > 1024 ? arena->left() - 1024 : 0) : budget;
    XG_REQUIRE(pl.per_p <= avail, XTPGPU_ERR_SHAPE,
               "bse matmul: one auxiliary function's stage-1 block (" + std::to_string(pl.per_p) +
                   " bytes) exceeds the scratch budget");
    pc = std::max<long>(1, std::min<long>(P, (long)(avail / std::max<size_t>(pl.per_p, 1))));
    const size_t s1n = (size_t)std::max(ct, vt) * k * pc * vtp;
    if (!m16) {
      s1a = static_cast<unsigned char*>(zbuf(op->s1, s1n * 8));
    } else {
      s1a = static_cast<unsigned char*>(zbuf(op->s1h, s1n * 2));
      if (split) s1b = static_cast<unsigned char*>(zbuf(op->s1l, s1n * 2));
    }
  }

  // ---------------- Hd (bse_operator.cc:70-89, PrepareMatrix1/2): -cd sum_P eps_P M^P_cc' X M^P_vv'^T
  if (op->cd) {
    const double alpha = -(double)op->cd;
    for (long p0 = 0; p0 < P; p0 += pc) {
      const long np = std::min(pc, P - p0);
      const long M1 = np * ct, N1 = k * vt;
      const double eb = m16 ? (split ? 4.0 : 2.0) : 8.0;
      const double f1 = 2.0 * M1 * N1 * ct, b1 = eb * ((double)M1 * ct + (double)N * k + (double)M1 * N1);
      const double f2 = 2.0 * ct * vt * np * vt * k, b2 = eb * ((double)M1 * N1 + (double)vt * np * vt) + 16.0 * N * k;
      EpiStage1 e1{nullptr, nullptr, nullptr, nullptr, 0, (int)ct, (int)vt, (int)vtp, (int)k, (int)np, true};
      if (!m16) {
        double* S = reinterpret_cast<double*>(s1a);
        e1.S64 = S;
        const double* mcc = op->stream64 ? nullptr : op->Mcc.as<double>() + p0 * ct * ctp;
        if (op->stream64) {
          // this chunk's M^P_{c c'} from the source, packed to [p][c1][ctp]
          ProfScope prof(pctx, "hd_stream_fp64", 0.0, 8.0 * 2.0 * np * ct * (double)(ct + ctp));
          op->srcbuf.ensure((size_t)np * ct * ct * 8);
          op->mcc64.ensure((size_t)np * ct * ctp * 8);
          double* W = op->srcbuf.as<double>();
          double* A = op->mcc64.as<double>();
          const int rc = op->src_fn(op->src_user, p0, np, op->cmin_r, op->cmin_r + ct - 1, W, (void*)st);
          XG_REQUIRE(rc == 0, rc < 0 ? rc : XTPGPU_ERR_ARG, "bse matmul: the Mmn chunk source failed");
          XG_HIP(hipGetLastError());
          launch_map(st, (size_t)np * ct * ctp, [=] __device__(size_t i) {
            const long q = (long)(i / (size_t)(ct * ctp)), r = (long)(i % (size_t)(ct * ctp));
            const long c1 = r / ctp, c2 = r % ctp;
            A[i] = c2 < ct ? W[(q * ct + c1) * ct + c2] : 0.0;
          });
          mcc = A;
        }
        {
          ProfScope prof(pctx, "hd_stage1_fp64", f1, b1);
          gemm64<true, true>(st, (int)M1, (int)N1, (int)round8(ct), 1, 1, Opnd<double>{mcc, nullptr, ctp, 0},
                             Opnd<double>{Xp, nullptr, ctp, 0}, nullptr, 0, false, e1);
        }
        ProfScope prof(pctx, "hd_stage2_fp64", f2, b2);
        gemm64<true, true>(st, (int)ct, (int)vt, (int)(np * vtp), (int)k, 1,
                           Opnd<double>{S, nullptr, k * np * vtp, np * vtp},
                           Opnd<double>{op->Bvv.as<double>() + p0 * vtp, nullptr, P * vtp, 0}, nullptr, 0, false,
                           EpiStore{Y, 1, ct, ldy, alpha, 1.0});
      } else {
        unsigned short* Sh = reinterpret_cast<unsigned short*>(s1a);
        unsigned short* Sl = split ? reinterpret_cast<unsigned short*>(s1b) : nullptr;
        e1.Sh = Sh;
        e1.Sl = Sl;
        {
          ProfScope prof(pctx, split ? "hd_stage1" : "hd_stage1_bf16", f1, b1);
          gemm16<true, true>(st, (int)M1, (int)N1, (int)round8(ct), 1, 1,
                             Opnd<unsigned short>{op->Mcc_h.as<unsigned short>() + p0 * ct * ctp,
                                                  split ? op->Mcc_l.as<unsigned short>() + p0 * ct * ctp : nullptr,
                                                  ctp, 0},
                             Opnd<unsigned short>{Xh, Xl, ctp, 0}, nullptr, 0, split, false, e1);
        }
        ProfScope prof(pctx, split ? "hd_stage2" : "hd_stage2_bf16", f2, b2);
        const Opnd<unsigned short> s2a{Sh, Sl, k * np * vtp, np * vtp};
        const Opnd<unsigned short> s2b{op->Bvv_h.as<unsigned short>() + p0 * vtp,
                                       split ? op->Bvv_l.as<unsigned short>() + p0 * vtp : nullptr, P * vtp, 0};
        // few output tiles (ct x vt per vector) over a long K = (P, v'): split K so the grid
        // covers the chip, FP32 partials, then a fixed-order reduction into Y
        const long tiles2 = ((ct + 127) / 128) * ((vt + 127) / 128) * k;
        const int ks2 = (int)std::max<long>(1, std::min<long>(4, std::min<long>(512 / std::max<long>(1, tiles2),
                                                                               (np * vtp) / 2048)));
        if (ks2 > 1 && bf16_vec(s2a, s2b, (int)(np * vtp), 0)) {
          // launched directly so the partial count is exactly dd.ksplit (<= 4, within pl.part)
          const GemmDims dd = make_dims<128, 128>((int)ct, (int)vt, (int)(np * vtp), (int)k, ks2,
                                                  bf16_wide_kstep(split), false);
          float* pt = static_cast<float*>(part_p);
          XG_HIP((launch_gemm_bf16<128, 128, true, true>(st, dd, s2a, s2b, nullptr, 0, split, EpiPart32{pt, ct, vt, k},
                                                         true)));
          const long nout = k * ct * vt, nsp = dd.ksplit, ctl = ct, vtl = vt;
          launch_map(st, (size_t)nout, [=] __device__(size_t i) {
            double acc = 0.0;
            for (long q = 0; q < nsp; ++q) acc += (double)pt[q * nout + (long)i];
            const long b = (long)i / (ctl * vtl), r = (long)i - b * ctl * vtl;
            const long n = r / ctl, m = r - n * ctl;
            Y[b * ldy + n * ctl + m] += alpha * acc;
          });
        } else {
          gemm16<true, true>(st, (int)ct, (int)vt, (int)(np * vtp), (int)k, 1, s2a, s2b, nullptr, 0, split, false,
                             EpiStore{Y, 1, ct, ldy, alpha, 1.0});
        }
      }
    }
  }

  // ---------------- Hd2 (bse_operator.cc:74-86, Hd2 branch): -cd2 sum_P eps_P M^P_{c1 v'} M^P_{v1 c'} X_{v'c'}
  if (op->cd2) {
    ProfScope prof(pctx, "hd2", 4.0 * P * vt * vt * ct * k, (m16 ? (split ? 4.0 : 2.0) : 8.0) * 2.0 * P * N + 16.0 * N * k);
    const double alpha = -(double)op->cd2;
    for (long p0 = 0; p0 < P; p0 += pc) {
      const long np = std::min(pc, P - p0);
      const long M1 = np * vt, N1 = k * vt;
      EpiStage1 e1{nullptr, nullptr, nullptr, op->eps.as<double>(), p0, (int)vt, (int)vt, (int)vtp, (int)k, (int)np,
                   false};
      if (!m16) {
        double* S = reinterpret_cast<double*>(s1a);
        e1.S64 = S;
        gemm64<true, true>(st, (int)M1, (int)N1, (int)round8(ct), 1, 1,
                           Opnd<double>{op->Mvc.as<double>() + p0 * Np, nullptr, ctp, 0},
                           Opnd<double>{Xp, nullptr, ctp, 0}, nullptr, 0, false, e1);
        gemm64<true, true>(st, (int)ct, (int)vt, (int)(np * vtp), (int)k, 1,
                           Opnd<double>{op->McvT.as<double>() + p0 * vtp, nullptr, P * vtp, 0},
                           Opnd<double>{S, nullptr, np * vtp, vt * np * vtp}, nullptr, 0, false,
                           EpiStore{Y, 1, ct, ldy, alpha, 1.0});
      } else {
        unsigned short* Sh = reinterpret_cast<unsigned short*>(s1a);
        unsigned short* Sl = split ? reinterpret_cast<unsigned short*>(s1b) : nullptr;
        e1.Sh = Sh;
        e1.Sl = Sl;
        gemm16<true, true>(st, (int)M1, (int)N1, (int)round8(ct), 1, 1,
                           Opnd<unsigned short>{op->Mvc_h.as<unsigned short>() + p0 * Np,
                                                split ? op->Mvc_l.as<unsigned short>() + p0 * Np : nullptr, ctp, 0},
                           Opnd<unsigned short>{Xh, Xl, ctp, 0}, nullptr, 0, split, false, e1);
        gemm16<true, true>(st, (int)ct, (int)vt, (int)(np * vtp), (int)k, 1,
                           Opnd<unsigned short>{op->McvT_h.as<unsigned short>() + p0 * vtp,
                                                split ? op->McvT_l.as<unsigned short>() + p0 * vtp : nullptr, P * vtp, 0},
                           Opnd<unsigned short>{Sh, Sl, np * vtp, vt * np * vtp}, nullptr, 0, split, false,
                           EpiStore{Y, 1, ct, ldy, alpha, 1.0});
      }
    }
  }

  // ---------------- Hx (bse_operator.cc:98-116, MultiplyBlocks): cx V (V^T X), always FP64
  if (op->cx) {
    ProfScope prof(pctx, "hx", 4.0 * N * P * k, 8.0 * (2.0 * N * P + 2.0 * N * k));
    const int tiles = (int)(((P + 63) / 64) * ((k + 63) / 64));
    const int ks = pick_ksplit(tiles, Np, 16);
    GemmDims dd = make_dims<64, 64>((int)P, (int)k, (int)Np, 1, ks, 16, false);
    double* part = static_cast<double*>(part_p);
    gemm64<true, true>(st, (int)P, (int)k, (int)Np, 1, ks, Opnd<double>{op->Mvc.as<double>(), nullptr, Np, 0},
                       Opnd<double>{Xp, nullptr, Np, 0}, nullptr, 0, false, EpiPartial{part, (int)P, (int)k, 1});
    // W[kv][Pp] (pads stay zero)
    const int nsp = dd.ksplit;
    launch_map(st, (size_t)P * k, [=] __device__(size_t i) {
      long kv = (long)(i / P), p = (long)(i % P);
      double acc = 0.0;
      for (int q = 0; q < nsp; ++q) acc += part[(long)q * P * k + i];
      Wb[kv * Pp + p] = acc;
    });
    gemm64<true, true>(st, (int)N, (int)k, (int)Pp, 1, 1, Opnd<double>{op->MvcT.as<double>(), nullptr, Pp, 0},
                       Opnd<double>{Wb, nullptr, Pp, 0}, nullptr, 0, false,
                       EpiStore{Y, 1, ldy, 0, (double)op->cx, 1.0});
  }
}

}  // namespace

#include "grouped.inc"

// BSE::AdjustHqpSize (bse.cc:147-186): host-side, no device work. The BSE window [vmin, cmax]
// (vtotal = homo - vmin + 1 valence, ctotal = cmax - homo conduction levels) takes the GW-window
// block of Hqp where the windows overlap; a BSE level outside [qpmin, qpmax] has no QP correction
// and keeps its RPA input energy (indexed from vmin - rpamin) on the diagonal, zero elsewhere.
extern "C" int xtpgpu_bse_adjust_hqp(const double* Hqp, int64_t ldh, const double* rpa_energies, int64_t n_rpa,
                                     int64_t homo, int64_t rpamin, int64_t qpmin, int64_t qpmax, int64_t vmin,
                                     int64_t cmax, double* out, int64_t ldo) {
  return guarded([&] {
    XG_REQUIRE(Hqp && rpa_energies && out, XTPGPU_ERR_ARG, "xtpgpu_bse_adjust_hqp: null argument");
    const long hs = (homo - vmin + 1) + (cmax - homo);
    const long gw = qpmax - qpmin + 1;
    XG_REQUIRE(homo >= vmin && cmax > homo && gw > 0 && qpmin <= homo && qpmax > homo, XTPGPU_ERR_SHAPE,
               "xtpgpu_bse_adjust_hqp: empty BSE or GW window");
    XG_REQUIRE(ldh >= gw && ldo >= hs, XTPGPU_ERR_SHAPE, "xtpgpu_bse_adjust_hqp: leading dimension too small");
    XG_REQUIRE(vmin >= rpamin && cmax - rpamin < n_rpa, XTPGPU_ERR_SHAPE,
               "xtpgpu_bse_adjust_hqp: BSE window outside the RPA energies");
    for (long j = 0; j < hs; ++j)
      for (long i = 0; i < hs; ++i) out[j * ldo + i] = 0.0;
    // BSE row b is level vmin + b; it lies in the GW window when qpmin <= vmin + b <= qpmax
    const long shift = vmin - qpmin;  // GW index of BSE row b = b + shift
    for (long b = 0; b < hs; ++b) {
      const long gb = b + shift;
      if (gb < 0 || gb >= gw) {
        out[b * ldo + b] = rpa_energies[vmin - rpamin + b];
        continue;
      }
      for (long a = 0; a < hs; ++a) {
        const long ga = a + shift;
        if (ga >= 0 && ga < gw) out[b * ldo + a] = Hqp[gb * ldh + ga];
      }
    }
  });
}

namespace {

// The BSE-sized Hqp blocks (bse.cc:147-186) and the screening weights eps0^-1 of an operator whose
// vt, ct, P and cqp are set.
void setup_hqp(xtpgpu_bseop op, const double* Hqp, int64_t ldhqp) {
  hipStream_t st = op->ctx->stream;
  const long vt = op->vt, ct = op->ct, hs = vt + ct;
  if (op->cqp) {
    XG_REQUIRE(ldhqp >= hs, XTPGPU_ERR_SHAPE, "xtpgpu_bse_create: ldhqp < vtotal+ctotal");
    std::vector<double> hc(ct * ct), hv(vt * vt), hcd(ct), hvd(vt);
    bool diag = true;
    for (long j = 0; j < hs; ++j)
      for (long i = 0; i < hs; ++i)
        if (i != j && Hqp[j * ldhqp + i] != 0.0) diag = false;
    for (long j = 0; j < ct; ++j)
      for (long i = 0; i < ct; ++i) hc[j * ct + i] = Hqp[(vt + j) * ldhqp + vt + i];
    for (long j = 0; j < vt; ++j)
      for (long i = 0; i < vt; ++i) hv[j * vt + i] = Hqp[j * ldhqp + i];
    for (long i = 0; i < ct; ++i) hcd[i] = Hqp[(vt + i) * ldhqp + vt + i];
    for (long i = 0; i < vt; ++i) hvd[i] = Hqp[i * ldhqp + i];
    op->hqp_diag = diag;
    op->hc.alloc(hc.size() * 8); op->hv.alloc(hv.size() * 8);
    op->hcd.alloc(ct * 8); op->hvd.alloc(vt * 8);
    XG_HIP(hipMemcpyAsync(op->hc.p, hc.data(), hc.size() * 8, hipMemcpyHostToDevice, st));
    XG_HIP(hipMemcpyAsync(op->hv.p, hv.data(), hv.size() * 8, hipMemcpyHostToDevice, st));
    XG_HIP(hipMemcpyAsync(op->hcd.p, hcd.data(), ct * 8, hipMemcpyHostToDevice, st));
    XG_HIP(hipMemcpyAsync(op->hvd.p, hvd.data(), vt * 8, hipMemcpyHostToDevice, st));
    XG_HIP(hipStreamSynchronize(st));
  }
}

// diag = diag0 + cqp (Hqp_cc - Hqp_vv)
void finish_diag(xtpgpu_bseop op) {
  const long N = op->N, ct = op->ct;
  double* dg = op->diag.as<double>();
  const double* d0 = op->diag0.as<double>();
  const double* hcd = op->cqp ? op->hcd.as<double>() : nullptr;
  const double* hvd = op->cqp ? op->hvd.as<double>() : nullptr;
  const double fqp = op->cqp;
  launch_map(op->ctx->stream, (size_t)N, [=] __device__(size_t i) {
    const long v = (long)i / ct, c = (long)i % ct;
    dg[i] = d0[i] + (fqp != 0.0 ? fqp * (hcd[c] - hvd[v]) : 0.0);
  });
}

void setup_hqp_eps(xtpgpu_bseop op, const double* Hqp, int64_t ldhqp, const double* eps0inv) {
  hipStream_t st = op->ctx->stream;
  const long P = op->P;
  setup_hqp(op, Hqp, ldhqp);
  op->eps.alloc(P * 8);
  if (eps0inv) {
    XG_HIP(hipMemcpyAsync(op->eps.p, eps0inv, P * 8, hipMemcpyHostToDevice, st));
  } else {
    fill_zero(st, op->eps.as<double>(), P);
  }
}

}  // namespace

namespace {

// The layouts that hold the transition block M^P_{vc} (Mvc, MvcT, McvT and their BF16 shadows),
// packed from a resident tensor; xtpgpu_bse_refresh repacks them after the tensor's transition
// block changed.
void pack_transitions(xtpgpu_bseop op, xtpgpu_mmn mmn) {
  hipStream_t st = op->ctx->stream;
  const long vt = op->vt, ct = op->ct, P = op->P, N = op->N, ctp = op->ctp, vtp = op->vtp, Pp = op->Pp;
  const long Np = vt * ctp, nt = mmn->ntotal, ps = (long)mmn->pstride();
  const long vmin_r = op->vmin_r, cmin_r = op->cmin_r;
  const int cx = op->cx, cd2 = op->cd2;
  const bool need16 = ((op->mask >> XTPGPU_BF16) & 1u) || ((op->mask >> XTPGPU_BF16X3) & 1u);
  const double* D = mmn->d();
  auto el = [=] __device__(long p, long slice, long row) { return D[p * ps + slice * nt + row]; };
  if (cx || cd2) {
    op->Mvc.ensure((size_t)P * Np * 8);
    double* A = op->Mvc.as<double>();
    launch_map(st, (size_t)P * Np, [=] __device__(size_t i) {
      long p = (long)(i / (size_t)Np), r = (long)(i % (size_t)Np);
      long v = r / ctp, c = r % ctp;
      A[i] = c < ct ? el(p, vmin_r + v, cmin_r + c) : 0.0;
    });
    if (need16 && cd2) {
      op->Mvc_h.ensure((size_t)P * Np * 2); op->Mvc_l.ensure((size_t)P * Np * 2);
      split_to_bf16(st, A, op->Mvc_h.as<unsigned short>(), op->Mvc_l.as<unsigned short>(), (size_t)P * Np);
    }
  }
  if (cx) {
    op->MvcT.ensure((size_t)N * Pp * 8);
    double* A = op->MvcT.as<double>();
    launch_map(st, (size_t)N * Pp, [=] __device__(size_t i) {
      long r = (long)(i / (size_t)Pp), p = (long)(i % (size_t)Pp);
      long v = r / ct, c = r % ct;
      A[i] = p < P ? el(p, vmin_r + v, cmin_r + c) : 0.0;
    });
  }
  if (cd2) {
    op->McvT.ensure((size_t)ct * P * vtp * 8);
    double* A = op->McvT.as<double>();
    launch_map(st, (size_t)ct * P * vtp, [=] __device__(size_t i) {
      long c1 = (long)(i / (size_t)(P * vtp)), r = (long)(i % (size_t)(P * vtp));
      long p = r / vtp, v2 = r % vtp;
      A[i] = v2 < vt ? el(p, cmin_r + c1, vmin_r + v2) : 0.0;
    });
    if (need16) {
      op->McvT_h.ensure((size_t)ct * P * vtp * 2); op->McvT_l.ensure((size_t)ct * P * vtp * 2);
      split_to_bf16(st, A, op->McvT_h.as<unsigned short>(), op->McvT_l.as<unsigned short>(), (size_t)ct * P * vtp);
    }
  }
}

// BSE_OPERATOR::diagonal (bse_operator.cc:134-175) from a resident tensor: diag0 (exchange and
// direct terms), then the Hqp part.
void build_diag(xtpgpu_bseop op, xtpgpu_mmn mmn) {
  hipStream_t st = op->ctx->stream;
  const long ct = op->ct, P = op->P, N = op->N, nt = mmn->ntotal, ps = (long)mmn->pstride();
  const long vmin_r = op->vmin_r, cmin_r = op->cmin_r;
  const double* D = mmn->d();
  const double* E = op->eps.as<double>();
  auto el = [=] __device__(long p, long slice, long row) { return D[p * ps + slice * nt + row]; };
  op->diag.ensure((size_t)N * 8);
  op->diag0.ensure((size_t)N * 8);
  double* d0 = op->diag0.as<double>();
  const double fx = op->cx, fd = op->cd, fd2 = op->cd2;
  launch_map(st, (size_t)N, [=] __device__(size_t i) {
    long v = (long)i / ct, c = (long)i % ct;
    double e = 0.0;
    if (fx != 0.0) {
      double s = 0.0;
      for (long p = 0; p < P; ++p) { double m = el(p, vmin_r + v, cmin_r + c); s += m * m; }
      e += fx * s;
    }
    if (fd != 0.0) {
      double s = 0.0;
      for (long p = 0; p < P; ++p) s += el(p, cmin_r + c, cmin_r + c) * E[p] * el(p, vmin_r + v, vmin_r + v);
      e -= fd * s;
    }
    if (fd2 != 0.0) {
      double s = 0.0;
      for (long p = 0; p < P; ++p) s += el(p, cmin_r + c, vmin_r + v) * E[p] * el(p, vmin_r + v, cmin_r + c);
      e -= fd2 * s;
    }
    d0[i] = e;
  });
  finish_diag(op);
}

}  // namespace

extern "C" int xtpgpu_bse_create(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* eps0inv, const double* Hqp,
                                 int64_t ldhqp, int64_t homo, int64_t rpamin, int64_t qpmin, int64_t vmin,
                                 int64_t cmax, int cqp, int cx, int cd, int cd2, unsigned precision_mask,
                                 xtpgpu_bseop* out) {
  return guarded([&] {
    (void)qpmin;
    XG_REQUIRE(ctx && mmn && out, XTPGPU_ERR_ARG, "xtpgpu_bse_create: null argument");
    XG_REQUIRE(!(cd != 0 && cd2 != 0), XTPGPU_ERR_ARG, "Hamiltonian cannot contain Hd and Hd2 at the same time");
    XG_REQUIRE(cqp == 0 || Hqp, XTPGPU_ERR_ARG, "xtpgpu_bse_create: Hqp required when cqp != 0");
    XG_REQUIRE((cd == 0 && cd2 == 0) || eps0inv, XTPGPU_ERR_ARG, "xtpgpu_bse_create: eps0inv required for Hd/Hd2");
    auto op = std::make_unique<xtpgpu_bseop_s>();
    op->ctx = ctx;
    op->cqp = cqp; op->cx = cx; op->cd = cd; op->cd2 = cd2;
    op->mask = (precision_mask | 1u) & ((1u << XTPGPU_FP64) | (1u << XTPGPU_BF16) | (1u << XTPGPU_BF16X3));
    // BSE_OPERATOR::configure (bse_operator.cc:28-37)
    const long bse_vmax = homo, bse_cmin = homo + 1;
    const long vt = bse_vmax - vmin + 1, ct = cmax - bse_cmin + 1;
    XG_REQUIRE(vt > 0 && ct > 0, XTPGPU_ERR_SHAPE, "xtpgpu_bse_create: empty BSE window");
    // slice index / row index = level - rpamin (TCMatrix_gwbse for the BSE starts at rpamin)
    const long vmin_r = vmin - rpamin, cmin_r = bse_cmin - rpamin;
    XG_REQUIRE(vmin_r >= 0 && cmin_r + ct <= mmn->mtotal && cmin_r + ct <= mmn->ntotal, XTPGPU_ERR_SHAPE,
               "xtpgpu_bse_create: BSE window exceeds the Mmn tensor");
    op->vt = vt; op->ct = ct; op->N = vt * ct; op->P = mmn->aux;
    op->ctp = ct >= 64 ? (ct + 63) / 64 * 64 : round8(ct); op->vtp = round8(vt); op->Pp = round8(op->P);
    const long P = op->P, N = op->N, nt = mmn->ntotal, ps = (long)mmn->pstride();
    const long ctp = op->ctp, vtp = op->vtp, Pp = op->Pp, Np = vt * ctp;
    XG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const double* D = mmn->d();

    setup_hqp_eps(op.get(), Hqp, ldhqp, eps0inv);
    const double* E = op->eps.as<double>();
    const bool need16 = ((op->mask >> XTPGPU_BF16) & 1u) || ((op->mask >> XTPGPU_BF16X3) & 1u);

    auto el = [=] __device__(long p, long slice, long row) { return D[p * ps + slice * nt + row]; };
    if (cd) {
      op->Mcc.alloc((size_t)P * ct * ctp * 8);
      double* A = op->Mcc.as<double>();
      launch_map(st, (size_t)P * ct * ctp, [=] __device__(size_t i) {
        long p = (long)(i / (size_t)(ct * ctp)), r = (long)(i % (size_t)(ct * ctp));
        long c1 = r / ctp, c2 = r % ctp;
        A[i] = c2 < ct ? el(p, cmin_r + c1, cmin_r + c2) : 0.0;
      });
      op->Bvv.alloc((size_t)vt * P * vtp * 8);
      double* B = op->Bvv.as<double>();
      launch_map(st, (size_t)vt * P * vtp, [=] __device__(size_t i) {
        long v1 = (long)(i / (size_t)(P * vtp)), r = (long)(i % (size_t)(P * vtp));
        long p = r / vtp, v2 = r % vtp;
        B[i] = v2 < vt ? E[p] * el(p, vmin_r + v1, vmin_r + v2) : 0.0;
      });
      if (need16) {
        op->Mcc_h.alloc((size_t)P * ct * ctp * 2); op->Mcc_l.alloc((size_t)P * ct * ctp * 2);
        split_to_bf16(st, A, op->Mcc_h.as<unsigned short>(), op->Mcc_l.as<unsigned short>(), (size_t)P * ct * ctp);
        op->Bvv_h.alloc((size_t)vt * P * vtp * 2); op->Bvv_l.alloc((size_t)vt * P * vtp * 2);
        split_to_bf16(st, B, op->Bvv_h.as<unsigned short>(), op->Bvv_l.as<unsigned short>(), (size_t)vt * P * vtp);
      }
    }
    op->vmin_r = vmin_r;
    op->cmin_r = cmin_r;
    pack_transitions(op.get(), mmn);
    build_diag(op.get(), mmn);
    XG_HIP(hipStreamSynchronize(st));
    *out = op.release();
  });
}

extern "C" int xtpgpu_synthetic_chunk(void* user, int64_t p0, int64_t np, int64_t lmin, int64_t lmax, double* dst,
                                      void* stream) {
  return guarded([&] {
    XG_REQUIRE(user && dst && np >= 0 && lmax >= lmin && p0 >= 0, XTPGPU_ERR_ARG, "xtpgpu_synthetic_chunk: bad arguments");
    const xtpgpu_synthetic_source src = *static_cast<const xtpgpu_synthetic_source*>(user);
    XG_REQUIRE(p0 + np <= src.aux, XTPGPU_ERR_SHAPE, "xtpgpu_synthetic_chunk: aux range beyond the source");
    if (np == 0) return;
    // the same sigma table as xtpgpu_mmn_fill_synthetic (host exp, so the values match bit for bit)
    std::vector<double> sig(np);
    for (long q = 0; q < np; ++q)
      sig[q] = src.sigma0 * std::exp(-src.decay * (double)(p0 + q) / (double)src.aux);
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevBuf ds(np * sizeof(double));
    XG_HIP(hipMemcpyAsync(ds.p, sig.data(), np * sizeof(double), hipMemcpyHostToDevice, st));
    const double* sg = ds.as<double>();
    const long W = lmax - lmin + 1, l0 = lmin, pp0 = p0;
    const uint64_t seed = src.seed;
    launch_map(st, (size_t)np * W * W, [=] __device__(size_t i) {
      const long q = (long)(i / (size_t)(W * W)), r = (long)(i % (size_t)(W * W));
      const long m = r / W + l0, n = r % W + l0;
      const long a = m < n ? m : n, b = m < n ? n : m;
      dst[i] = synth_normal(seed, (uint64_t)(pp0 + q), (uint64_t)a, (uint64_t)b) * sg[q];
    });
    XG_HIP(hipStreamSynchronize(st));  // ds is freed on return
  });
}

extern "C" int xtpgpu_bse_create_streamed(xtpgpu_ctx ctx, xtpgpu_mmn_chunk_fn source, void* user, int64_t aux,
                                          const double* eps0inv, const double* epsinv_matrix, const double* Hqp,
                                          int64_t ldhqp, int64_t homo,
                                          int64_t rpamin, int64_t qpmin, int64_t vmin, int64_t cmax, int cqp,
                                          int cx, int cd, int cd2, unsigned precision_mask, int flags,
                                          int64_t chunk_bytes, xtpgpu_bseop* out) {
  return guarded([&] {
    (void)qpmin;
    XG_REQUIRE(ctx && source && out, XTPGPU_ERR_ARG, "xtpgpu_bse_create_streamed: null argument");
    XG_REQUIRE(aux > 0, XTPGPU_ERR_SHAPE, "xtpgpu_bse_create_streamed: aux must be positive");
    XG_REQUIRE(!(cd != 0 && cd2 != 0), XTPGPU_ERR_ARG, "Hamiltonian cannot contain Hd and Hd2 at the same time");
    XG_REQUIRE(cqp == 0 || Hqp, XTPGPU_ERR_ARG, "xtpgpu_bse_create: Hqp required when cqp != 0");
    XG_REQUIRE((cd == 0 && cd2 == 0) || eps0inv || epsinv_matrix, XTPGPU_ERR_ARG,
               "xtpgpu_bse_create: eps0inv required for Hd/Hd2");
    XG_REQUIRE(!(epsinv_matrix && cd2), XTPGPU_ERR_ARG,
               "xtpgpu_bse_create_streamed: the full inverse dielectric matrix supports the direct term only");
    XG_REQUIRE(vmin >= rpamin && homo >= vmin && cmax > homo, XTPGPU_ERR_SHAPE, "xtpgpu_bse_create_streamed: empty BSE window");
    auto op = std::make_unique<xtpgpu_bseop_s>();
    op->ctx = ctx;
    op->cqp = cqp; op->cx = cx; op->cd = cd; op->cd2 = cd2;
    op->mask = (precision_mask | 1u) & ((1u << XTPGPU_FP64) | (1u << XTPGPU_BF16) | (1u << XTPGPU_BF16X3));
    const long vt = homo - vmin + 1, ct = cmax - homo;
    op->vt = vt; op->ct = ct; op->N = vt * ct; op->P = aux;
    op->ctp = ct >= 64 ? (ct + 63) / 64 * 64 : round8(ct); op->vtp = round8(vt); op->Pp = round8(op->P);
    const long P = op->P, N = op->N, ctp = op->ctp, vtp = op->vtp, Pp = op->Pp, Np = vt * ctp;
    XG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // with the full matrix the valence block is packed unscaled (weights 1) and multiplied by W later
    std::vector<double> ones;
    if (epsinv_matrix) ones.assign(P, 1.0);
    setup_hqp_eps(op.get(), Hqp, ldhqp, epsinv_matrix ? ones.data() : eps0inv);
    const double* E = op->eps.as<double>();
    const bool need16 = ((op->mask >> XTPGPU_BF16) & 1u) || ((op->mask >> XTPGPU_BF16X3) & 1u);
    const bool keep_mcc64 = !(flags & XTPGPU_BSE_STREAM_FP64);
    // the source window is the BSE window: tensor indices [w0, w0 + W); v = w0 + v1, c = w0 + vt + c1
    const long w0 = vmin - rpamin, W = vt + ct;
    op->cmin_r = w0 + vt;
    if (!keep_mcc64 && cd) {
      op->stream64 = true;
      op->src_fn = source;
      op->src_user = user;
    }
    // layouts (pads zero)
    if (cd) {
      if (keep_mcc64) op->Mcc.alloc((size_t)P * ct * ctp * 8);
      op->Bvv.alloc((size_t)vt * P * vtp * 8);
      zero_bytes(st, op->Bvv.p, op->Bvv.bytes);
      if (need16) {
        op->Mcc_h.alloc((size_t)P * ct * ctp * 2); op->Mcc_l.alloc((size_t)P * ct * ctp * 2);
        op->Bvv_h.alloc((size_t)vt * P * vtp * 2); op->Bvv_l.alloc((size_t)vt * P * vtp * 2);
      }
    }
    if (cx || cd2) {
      op->Mvc.alloc((size_t)P * Np * 8);
      if (need16 && cd2) { op->Mvc_h.alloc((size_t)P * Np * 2); op->Mvc_l.alloc((size_t)P * Np * 2); }
    }
    if (cx) {
      op->MvcT.alloc((size_t)N * Pp * 8);
      zero_bytes(st, op->MvcT.p, op->MvcT.bytes);
    }
    if (cd2) {
      op->McvT.alloc((size_t)ct * P * vtp * 8);
      zero_bytes(st, op->McvT.p, op->McvT.bytes);
      if (need16) { op->McvT_h.alloc((size_t)ct * P * vtp * 2); op->McvT_l.alloc((size_t)ct * P * vtp * 2); }
    }
    op->diag.alloc((size_t)N * 8);
    double* dg = op->diag.as<double>();
    fill_zero(st, dg, N);
    // M^P_{cc} (diagonal of the conduction block) for the direct term's share of the diagonal
    DevBuf mccd(cd ? (size_t)P * ct * 8 : 0);
    double* Mccd = cd ? mccd.as<double>() : nullptr;
    // aux chunks of the source window
    const size_t budget = chunk_bytes > 0 ? (size_t)chunk_bytes : ((size_t)4 << 30);
    const long pc = std::max<long>(1, std::min<long>(P, (long)(budget / ((size_t)W * W * 8))));
    DevBuf win((size_t)pc * W * W * 8);
    double* Dc = win.as<double>();
    double* Mcc = keep_mcc64 && cd ? op->Mcc.as<double>() : nullptr;
    unsigned short* Mch = need16 && cd ? op->Mcc_h.as<unsigned short>() : nullptr;
    unsigned short* Mcl = need16 && cd ? op->Mcc_l.as<unsigned short>() : nullptr;
    double* Bvv = cd ? op->Bvv.as<double>() : nullptr;
    double* Mvc = (cx || cd2) ? op->Mvc.as<double>() : nullptr;
    double* MvcT = cx ? op->MvcT.as<double>() : nullptr;
    double* McvT = cd2 ? op->McvT.as<double>() : nullptr;
    const double fx = cx, fd = cd, fd2 = cd2;
    for (long p0 = 0; p0 < P; p0 += pc) {
      const long np = std::min(pc, P - p0);
      const int rc = source(user, p0, np, w0, w0 + W - 1, Dc, (void*)st);
      XG_REQUIRE(rc == 0, rc < 0 ? rc : XTPGPU_ERR_ARG, "xtpgpu_bse_create_streamed: the Mmn chunk source failed");
      XG_HIP(hipGetLastError());
      // element (p, a, b) of the chunk, a and b window indices (valence 0..vt-1, conduction vt..W-1)
      auto el = [=] __device__(long p, long a, long b) { return Dc[((p - p0) * W + a) * W + b]; };
      if (cd) {
        launch_map(st, (size_t)np * ct * ctp, [=] __device__(size_t i) {
          const long p = p0 + (long)(i / (size_t)(ct * ctp)), r = (long)(i % (size_t)(ct * ctp));
          const long c1 = r / ctp, c2 = r % ctp;
          const double x = c2 < ct ? el(p, vt + c1, vt + c2) : 0.0;
          const size_t o = (size_t)p * ct * ctp + r;
          if (Mcc) Mcc[o] = x;
          if (Mch) {
            unsigned short h, l;
            split_bf16(x, h, l);
            Mch[o] = h;
            Mcl[o] = l;
          }
        });
        launch_map(st, (size_t)vt * np * vt, [=] __device__(size_t i) {
          const long v1 = (long)(i / (size_t)(np * vt)), r = (long)(i % (size_t)(np * vt));
          const long p = p0 + r / vt, v2 = r % vt;
          Bvv[(v1 * P + p) * vtp + v2] = E[p] * el(p, v1, v2);
        });
        launch_map(st, (size_t)np * ct, [=] __device__(size_t i) {
          const long p = p0 + (long)(i / (size_t)ct), c = (long)(i % (size_t)ct);
          Mccd[p * ct + c] = el(p, vt + c, vt + c);
        });
      }
      if (Mvc) {
        launch_map(st, (size_t)np * Np, [=] __device__(size_t i) {
          const long p = p0 + (long)(i / (size_t)Np), r = (long)(i % (size_t)Np);
          const long v = r / ctp, c = r % ctp;
          Mvc[(size_t)p * Np + r] = c < ct ? el(p, v, vt + c) : 0.0;
        });
      }
      if (MvcT) {
        launch_map(st, (size_t)N * np, [=] __device__(size_t i) {
          const long r = (long)(i / (size_t)np), p = p0 + (long)(i % (size_t)np);
          const long v = r / ct, c = r % ct;
          MvcT[r * Pp + p] = el(p, v, vt + c);
        });
      }
      if (McvT) {
        launch_map(st, (size_t)ct * np * vt, [=] __device__(size_t i) {
          const long c1 = (long)(i / (size_t)(np * vt)), r = (long)(i % (size_t)(np * vt));
          const long p = p0 + r / vt, v2 = r % vt;
          McvT[(c1 * P + p) * vtp + v2] = el(p, vt + c1, v2);
        });
      }
      // BSE_OPERATOR::diagonal (bse_operator.cc:134-175), this chunk's share of the aux sums
      launch_map(st, (size_t)N, [=] __device__(size_t i) {
        const long v = (long)i / ct, c = (long)i % ct;
        double e = 0.0;
        if (fx != 0.0) {
          double sx = 0.0;
          for (long p = p0; p < p0 + np; ++p) { const double m = el(p, v, vt + c); sx += m * m; }
          e += fx * sx;
        }
        if (fd2 != 0.0) {
          double sd = 0.0;
          for (long p = p0; p < p0 + np; ++p) sd += el(p, vt + c, v) * E[p] * el(p, v, vt + c);
          e -= fd2 * sd;
        }
        dg[i] += e;
      });
    }
    op->screen_full = epsinv_matrix != nullptr;
    if (cd && epsinv_matrix) {
      // Bvv_v1 (P x vtp, rows P) <- W Bvv_v1 for every v1: the screened valence block in the
      // unrotated aux basis
      DevBuf wm((size_t)P * P * 8), braw((size_t)vt * P * vtp * 8);
      XG_HIP(hipMemcpyAsync(wm.p, epsinv_matrix, (size_t)P * P * 8, hipMemcpyHostToDevice, st));
      XG_HIP(hipMemcpyAsync(braw.p, Bvv, (size_t)vt * P * vtp * 8, hipMemcpyDeviceToDevice, st));
      // per v1: C(j, q) = sum_p Braw(p, j) W(q, p), with Braw read M-contiguous (element (p, j) at
      // p * vtp + j) and W's column q read as row q (W = eps^-1(0) is symmetric)
      gemm64<false, true>(st, (int)vtp, (int)P, (int)P, (int)vt, 1,
                          Opnd<double>{braw.as<double>(), nullptr, vtp, P * vtp},
                          Opnd<double>{wm.as<double>(), nullptr, P, 0}, nullptr, 0, false,
                          EpiStore{Bvv, 1, vtp, P * vtp, 1.0, 0.0});
      XG_HIP(hipStreamSynchronize(st));
    }
    if (cd) {
      // direct term of the diagonal: sum_P M^P_{cc} Bvv[v][P][v]
      const double fdd = cd;
      launch_map(st, (size_t)N, [=] __device__(size_t i) {
        const long v = (long)i / ct, c = (long)i % ct;
        double sd = 0.0;
        for (long p = 0; p < P; ++p) sd += Mccd[p * ct + c] * Bvv[(v * P + p) * vtp + v];
        dg[i] -= fdd * sd;
      });
    }
    if (need16 && cd) {
      split_to_bf16(st, Bvv, op->Bvv_h.as<unsigned short>(), op->Bvv_l.as<unsigned short>(), (size_t)vt * P * vtp);
    }
    if (need16 && cd2) {
      split_to_bf16(st, Mvc, op->Mvc_h.as<unsigned short>(), op->Mvc_l.as<unsigned short>(), (size_t)P * Np);
      split_to_bf16(st, McvT, op->McvT_h.as<unsigned short>(), op->McvT_l.as<unsigned short>(), (size_t)ct * P * vtp);
    }
    op->diag0.alloc((size_t)N * 8);
    XG_HIP(hipMemcpyAsync(op->diag0.p, dg, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
    finish_diag(op.get());
    XG_HIP(hipStreamSynchronize(st));
    *out = op.release();
  });
}

extern "C" int xtpgpu_bse_set_hqp(xtpgpu_bseop op, const double* Hqp, int64_t ldhqp) {
  return guarded([&] {
    XG_REQUIRE(op && (Hqp || op->cqp == 0), XTPGPU_ERR_ARG, "xtpgpu_bse_set_hqp: null argument");
    XG_HIP(hipSetDevice(op->ctx->device));
    hipStream_t st = op->ctx->stream;
    XG_HIP(hipStreamSynchronize(st));  // no product may still read the old blocks
    setup_hqp(op, Hqp, ldhqp);
    finish_diag(op);
    XG_HIP(hipStreamSynchronize(st));
  });
}

extern "C" int xtpgpu_bse_refresh(xtpgpu_bseop op, xtpgpu_mmn mmn) {
  return guarded([&] {
    XG_REQUIRE(op && mmn, XTPGPU_ERR_ARG, "xtpgpu_bse_refresh: null argument");
    XG_REQUIRE(!op->stream64 && !op->screen_full, XTPGPU_ERR_ARG,
               "xtpgpu_bse_refresh: operator was packed from a chunk source");
    XG_REQUIRE(mmn->aux == op->P && op->cmin_r + op->ct <= mmn->mtotal && op->cmin_r + op->ct <= mmn->ntotal,
               XTPGPU_ERR_SHAPE, "xtpgpu_bse_refresh: tensor does not match the operator");
    XG_HIP(hipSetDevice(op->ctx->device));
    hipStream_t st = op->ctx->stream;
    XG_HIP(hipStreamSynchronize(st));
    pack_transitions(op, mmn);
    build_diag(op, mmn);
    XG_HIP(hipStreamSynchronize(st));
  });
}

extern "C" int xtpgpu_bse_memory(xtpgpu_bseop op, int64_t* bytes) {
  return guarded([&] {
    XG_REQUIRE(op && bytes, XTPGPU_ERR_ARG, "null argument");
    const DevBuf* all[] = {&op->hc, &op->hv, &op->hcd, &op->hvd, &op->eps, &op->diag, &op->Mcc, &op->Bvv,
                           &op->Mvc, &op->MvcT, &op->McvT, &op->Mcc_h, &op->Mcc_l, &op->Bvv_h, &op->Bvv_l,
                           &op->Mvc_h, &op->Mvc_l, &op->McvT_h, &op->McvT_l, &op->xp, &op->xh, &op->xl,
                           &op->s1, &op->s1h, &op->s1l, &op->part, &op->wbuf, &op->srcbuf, &op->mcc64};
    size_t t = 0;
    for (const DevBuf* b : all) t += b->bytes;
    *bytes = (int64_t)t;
  });
}

extern "C" int xtpgpu_bse_destroy(xtpgpu_bseop op) {
  return guarded([&] {
    if (!op) return;
    (void)hipSetDevice(op->ctx->device);
    (void)hipStreamSynchronize(op->ctx->stream);
    delete op;
  });
}

extern "C" int xtpgpu_bse_size(xtpgpu_bseop op, int64_t* size) {
  return guarded([&] {
    XG_REQUIRE(op && size, XTPGPU_ERR_ARG, "null argument");
    *size = op->N;
  });
}

extern "C" int xtpgpu_bse_diagonal(xtpgpu_bseop op, double* out) {
  return guarded([&] {
    XG_REQUIRE(op && out, XTPGPU_ERR_ARG, "null argument");
    XG_HIP(hipSetDevice(op->ctx->device));
    XG_HIP(hipMemcpyAsync(out, op->diag.p, op->N * 8, hipMemcpyDeviceToHost, op->ctx->stream));
    XG_HIP(hipStreamSynchronize(op->ctx->stream));
  });
}

extern "C" int xtpgpu_bse_matmul_device(xtpgpu_bseop op, int precision, const double* X, int64_t ldx, int64_t k,
                                        double* Y, int64_t ldy) {
  return guarded([&] {
    XG_REQUIRE(op && X && Y, XTPGPU_ERR_ARG, "null argument");
    XG_HIP(hipSetDevice(op->ctx->device));
    bse_matmul_impl(op, precision, X, ldx, k, Y, ldy);
    XG_HIP(hipStreamSynchronize(op->ctx->stream));
  });
}

extern "C" int xtpgpu_bse_matmul_device_noalloc(xtpgpu_bseop op, int precision, const double* X, int64_t ldx,
                                                int64_t k, double* Y, int64_t ldy) {
  return guarded([&] {
    XG_REQUIRE(op && X && Y, XTPGPU_ERR_ARG, "null argument");
    XG_REQUIRE(k >= 0 && ldx >= op->N && ldy >= op->N, XTPGPU_ERR_SHAPE,
               "xtpgpu_bse_matmul_device_noalloc: bad leading dimension");
    XG_REQUIRE(precision == XTPGPU_FP64 || ((op->mask >> precision) & 1u), XTPGPU_ERR_ARG,
               "bse matmul: operator was created without this precision");
    XG_REQUIRE(!(op->stream64 && precision == XTPGPU_FP64), XTPGPU_ERR_ARG,
               "xtpgpu_bse_matmul_device_noalloc: an operator that streams its FP64 M_cc' from a source "
               "needs chunk buffers; use xtpgpu_bse_matmul_device");
    const MatmulPlan pl = plan_matmul(op, precision, k);
    const size_t fixed = ScratchArena::up(pl.xp) + ScratchArena::up(pl.xh) + ScratchArena::up(pl.xl) +
                         ScratchArena::up(pl.part) + ScratchArena::up(pl.wbuf) + ScratchArena::up(pl.per_p) + 1024;
    XG_REQUIRE(fixed <= kNoAllocScratchBytes, XTPGPU_ERR_SHAPE,
               "xtpgpu_bse_matmul_device_noalloc: operator size x " + std::to_string(k) + " columns needs " +
                   std::to_string(fixed) + " bytes of scratch, more than its " +
                   std::to_string(kNoAllocScratchBytes) + "-byte region");
    XG_HIP(hipSetDevice(op->ctx->device));
    ScratchArena arena{scratch_on(op->ctx->device), kNoAllocScratchBytes, 0};
    bse_matmul_impl(op, precision, X, ldx, k, Y, ldy, hd_chunk_budget(), &arena);
    XG_HIP(hipStreamSynchronize(op->ctx->stream));
  });
}

extern "C" int xtpgpu_bse_matmul(xtpgpu_bseop op, int precision, const double* X, int64_t ldx, int64_t k, double* Y,
                                 int64_t ldy) {
  return guarded([&] {
    XG_REQUIRE(op && X && Y, XTPGPU_ERR_ARG, "null argument");
    XG_REQUIRE(ldx >= op->N && ldy >= op->N && k >= 0, XTPGPU_ERR_SHAPE, "xtpgpu_bse_matmul: bad leading dimension");
    XG_REQUIRE(precision == XTPGPU_FP64 || ((op->mask >> precision) & 1u), XTPGPU_ERR_ARG,
               "bse matmul: operator was created without this precision");
    if (k == 0) return;
    XG_HIP(hipSetDevice(op->ctx->device));
    hipStream_t st = op->ctx->stream;
    const size_t n = (size_t)op->N * k;
    DevBuf dx(n * 8), dy(n * 8);
    XG_HIP(hipMemcpy2DAsync(dx.p, op->N * 8, X, ldx * 8, op->N * 8, k, hipMemcpyHostToDevice, st));
    if (grouped_matmul_env())
      grouped_matmul_one(op, precision, dx.as<double>(), k, dy.as<double>(), st);
    else
      bse_matmul_impl(op, precision, dx.as<double>(), op->N, k, dy.as<double>(), op->N);
    XG_HIP(hipMemcpy2DAsync(Y, ldy * 8, dy.p, op->N * 8, op->N * 8, k, hipMemcpyDeviceToHost, st));
    XG_HIP(hipStreamSynchronize(st));
  });
}

#include "mfop.inc"
#include "davidson.inc"
#include "davidson_ham.inc"
#include "adjoint.inc"
#include "gw.inc"
#include "batched.inc"
