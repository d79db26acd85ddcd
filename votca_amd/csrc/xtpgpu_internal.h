// Internal runtime pieces of libxtpgpu: error state, device buffers, context.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xtpgpu.h"

namespace xg {

struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define XG_HIP(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      throw ::xg::Error(e_ == hipErrorOutOfMemory ? XTPGPU_ERR_OOM : XTPGPU_ERR_HIP,     \
                        std::string("HIP error '") + hipGetErrorString(e_) + "' at " +   \
                            __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")"); \
    }                                                                                     \
  } while (0)

#define XG_REQUIRE(cond, code, msg)                        \
  do {                                                     \
    if (!(cond)) throw ::xg::Error((code), std::string(msg)); \
  } while (0)

// Owning device allocation. The allocation checks free memory first, mirroring the
// reference's guard in CudaMatrix (xtp/src/libxtp/cudamatrix.cc:97-113).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n == 0) return;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && n > fr)
      throw Error(XTPGPU_ERR_OOM, "device allocation of " + std::to_string(n) +
                                      " bytes exceeds free device memory (" + std::to_string(fr) + ")");
    XG_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  void ensure(size_t n) {
    if (n > bytes) alloc(n);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace xg

namespace xg {
// Named kernel timers: HIP events recorded on the context's stream around tagged
// launches, with the algorithmic FLOPs / bytes of each launch.
struct KTimer {
  struct Rec { hipEvent_t a, b; double flops, bytes; };
  std::vector<Rec> pending;
  double ms = 0.0, flops = 0.0, bytes = 0.0;
  long launches = 0;
};
}  // namespace xg

struct xtpgpu_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // grow-only scratch buffers, reused across calls on this context (ws[1]: the eps builds' scaled
  // transition copies, at most the scale budget)
  xg::DevBuf ws[6];
  bool profiling = false;
  void* blas = nullptr;  // rocblas_handle for the rocSOLVER calls, bound to `stream`
  std::vector<std::pair<std::string, xg::KTimer>> timers;
  std::vector<hipEvent_t> event_pool;
  xg::KTimer& timer(const std::string& n) {
    for (auto& t : timers)
      if (t.first == n) return t.second;
    timers.emplace_back(n, xg::KTimer());
    return timers.back().second;
  }
  hipEvent_t event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    XG_HIP(hipEventCreate(&e));
    return e;
  }
  // fold finished event pairs into the totals (call after the stream is synchronized)
  void collect() {
    for (auto& t : timers) {
      for (auto& r : t.second.pending) {
        float ms = 0.f;
        XG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        t.second.ms += ms;
        t.second.flops += r.flops;
        t.second.bytes += r.bytes;
        t.second.launches += 1;
        event_pool.push_back(r.a);
        event_pool.push_back(r.b);
      }
      t.second.pending.clear();
    }
  }
};

namespace xg {
// RAII scope that brackets the launches issued inside it with an event pair.
struct ProfScope {
  xtpgpu_ctx_s* c = nullptr;
  KTimer* t = nullptr;
  hipEvent_t a = nullptr;
  double flops, bytes;
  ProfScope(xtpgpu_ctx_s* ctx, const char* name, double f, double by) : flops(f), bytes(by) {
    if (ctx && ctx->profiling && name) {
      c = ctx;
      t = &ctx->timer(name);
      a = ctx->event();
      XG_HIP(hipEventRecord(a, ctx->stream));
    }
  }
  ~ProfScope() {
    if (c) {
      hipEvent_t b = c->event();
      if (hipEventRecord(b, c->stream) == hipSuccess) t->pending.push_back({a, b, flops, bytes});
    }
  }
};
}  // namespace xg

struct xtpgpu_mmn_s {
  xtpgpu_ctx ctx = nullptr;
  int64_t mmin = 0, mmax = 0, nmin = 0, nmax = 0, mtotal = 0, ntotal = 0, aux = 0;
  // Device layout, aux-major: element (m, n, P) at data[P*mtotal*ntotal + m*ntotal + n].
  xg::DevBuf data;
  uint64_t version = 0;  // bumped whenever the tensor changes
  double* d() const { return data.as<double>(); }
  size_t pstride() const { return (size_t)mtotal * (size_t)ntotal; }
};
