// C++ host surface over the xtpgpu C-ABI, shaped like the reference's operator classes so that
// VOTCA-XTP's GW loop and DavidsonSolver can use the device path unchanged:
//   xtpgpu::TCMatrix     ~ TCMatrix_gwbse        (xtp/include/votca/xtp/threecenter.h:70-142)
//   xtpgpu::BSEOperator  ~ BSE_OPERATOR<...>     (xtp/include/votca/xtp/bse_operator.h:40-87),
//                          a MatrixFreeOperator (rows(), diagonal(), matmul()) as the solver needs
//                          (xtp/include/votca/xtp/matrixfreeoperator.h:45-70)
// Errors come back as std::runtime_error with the library's message, as the reference throws.
// Matrices are plain column-major double buffers (Eigen::MatrixXd::data() fits directly).
#ifndef XTPGPU_HPP
#define XTPGPU_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "xtpgpu.h"

namespace xtpgpu {

inline void check(int rc) {
  if (rc != XTPGPU_OK) throw std::runtime_error(std::string("xtpgpu: ") + xtpgpu_last_error());
}

class Context {
 public:
  explicit Context(int device = 0) { check(xtpgpu_ctx_create(device, 0, &h_)); }
  ~Context() { xtpgpu_ctx_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  xtpgpu_ctx get() const { return h_; }

 private:
  xtpgpu_ctx h_ = nullptr;
};

class TCMatrix {
 public:
  TCMatrix(Context& ctx, int64_t basissize, int64_t mmin, int64_t mmax, int64_t nmin, int64_t nmax)
      : ntotal_(nmax - nmin + 1), mtotal_(mmax - mmin + 1), aux_(basissize) {
    check(xtpgpu_mmn_create(ctx.get(), mmin, mmax, nmin, nmax, basissize, &h_));
  }
  ~TCMatrix() { xtpgpu_mmn_destroy(h_); }
  TCMatrix(const TCMatrix&) = delete;
  TCMatrix& operator=(const TCMatrix&) = delete;
  int64_t auxsize() const { return aux_; }
  int64_t msize() const { return mtotal_; }
  int64_t nsize() const { return ntotal_; }
  void set_slice(int64_t m, const double* slice, int64_t ld) { check(xtpgpu_mmn_set_slice(h_, m, slice, ld)); }
  void get_slice(int64_t m, double* slice, int64_t ld) const { check(xtpgpu_mmn_get_slice(h_, m, slice, ld)); }
  void MultiplyRightWithAuxMatrix(const double* A, int64_t lda) { check(xtpgpu_mmn_multiply_right(h_, A, lda)); }
  xtpgpu_mmn get() const { return h_; }

 private:
  xtpgpu_mmn h_ = nullptr;
  int64_t ntotal_, mtotal_, aux_;
};

struct BSEOperatorOptions {  // BSEOperator_Options (bse_operator.h:32-38)
  int64_t homo, rpamin, qpmin, vmin, cmax;
};

template <int cqp, int cx, int cd, int cd2>
class BSEOperator {
 public:
  BSEOperator(Context& ctx, TCMatrix& Mmn, const double* eps0inv, const double* Hqp, int64_t ldhqp,
              const BSEOperatorOptions& o, unsigned precision_mask = 1u << XTPGPU_FP64) {
    static_assert(!(cd != 0 && cd2 != 0), "Hamiltonian cannot contain Hd and Hd2 at the same time");
    check(xtpgpu_bse_create(ctx.get(), Mmn.get(), eps0inv, Hqp, ldhqp, o.homo, o.rpamin, o.qpmin, o.vmin, o.cmax, cqp,
                            cx, cd, cd2, precision_mask, &h_));
    check(xtpgpu_bse_size(h_, &size_));
  }
  ~BSEOperator() { xtpgpu_bse_destroy(h_); }
  BSEOperator(const BSEOperator&) = delete;
  BSEOperator& operator=(const BSEOperator&) = delete;
  int64_t rows() const { return size_; }
  int64_t cols() const { return size_; }
  std::vector<double> diagonal() const {
    std::vector<double> d(size_);
    check(xtpgpu_bse_diagonal(h_, d.data()));
    return d;
  }
  // Y (size x k) = H X (size x k), column-major host buffers
  void matmul(const double* X, int64_t k, double* Y, int precision = XTPGPU_FP64) const {
    check(xtpgpu_bse_matmul(h_, precision, X, size_, k, Y, size_));
  }
  xtpgpu_bseop get() const { return h_; }

 private:
  xtpgpu_bseop h_ = nullptr;
  int64_t size_ = 0;
};

using SingletOperator_TDA = BSEOperator<1, 2, 1, 0>;
using TripletOperator_TDA = BSEOperator<1, 0, 1, 0>;
using SingletOperator_BTDA_B = BSEOperator<0, 2, 0, 1>;
using HqpOperator = BSEOperator<1, 0, 0, 0>;
using HxOperator = BSEOperator<0, 1, 0, 0>;
using HdOperator = BSEOperator<0, 0, 1, 0>;
using Hd2Operator = BSEOperator<0, 0, 0, 1>;

}  // namespace xtpgpu

#endif  // XTPGPU_HPP
