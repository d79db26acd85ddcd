// C++ host surface over the xtpgpu C-ABI, shaped like the reference's operator classes so that
// VOTCA-XTP's GW loop and DavidsonSolver can use the device path unchanged:
//   xtpgpu::TCMatrix     ~ TCMatrix_gwbse        (xtp/include/votca/xtp/threecenter.h:70-142)
//   xtpgpu::MatrixFreeOperator ~ MatrixFreeOperator (xtp/include/votca/xtp/matrixfreeoperator.h:45-98):
//                          size, diagonal() and a matmul on device buffers; the base the solver takes
//   xtpgpu::HostMatrixFreeOperator  a MatrixFreeOperator whose matmul runs on the host (the shape of
//                          the reference's own test operators, test_davidson.cc:127-165)
//   xtpgpu::BSEOperator  ~ BSE_OPERATOR<...>     (xtp/include/votca/xtp/bse_operator.h:40-87),
//                          derived from MatrixFreeOperator as BSE_OPERATOR is
//   xtpgpu::RPA          ~ RPA                   (xtp/include/votca/xtp/rpa.h:35-115)
//   xtpgpu::DavidsonSolver ~ DavidsonSolver      (xtp/include/votca/xtp/davidsonsolver.h:47-334)
//   xtpgpu::Sigma_PPM    ~ Sigma_PPM + GW::SolveQP (sigma_ppm.h, gw.cc:323-410)
// Errors come back as std::runtime_error with the library's message, as the reference throws.
// Matrices are plain column-major double buffers (Eigen::MatrixXd::data() fits directly).
#ifndef XTPGPU_HPP
#define XTPGPU_HPP

#include <algorithm>
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
  // return the grow-only scratch the eps builds keep between calls to the device
  void release_scratch() { check(xtpgpu_ctx_release_scratch(h_)); }

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

// MatrixFreeOperator (matrixfreeoperator.h:45-98). A subclass gives diagonal() and matmul_device():
// Y (size x k, leading dimension ldy) = A X (leading dimension ldx) on device buffers, issued on
// `stream` (the context's hipStream_t), for XTPGPU_FP64 or a mixed precision it lists in
// precision_mask(). The library sees it through mfop(): a callback view made on first use.
class MatrixFreeOperator {
 public:
  explicit MatrixFreeOperator(Context& ctx) : ctx_(&ctx) {}
  virtual ~MatrixFreeOperator() {
    if (mf_) xtpgpu_mfop_destroy(mf_);
  }
  MatrixFreeOperator(const MatrixFreeOperator&) = delete;
  MatrixFreeOperator& operator=(const MatrixFreeOperator&) = delete;
  int64_t size() const { return size_; }
  void set_size(int64_t n) { size_ = n; }
  int64_t rows() const { return size_; }
  int64_t cols() const { return size_; }
  virtual std::vector<double> diagonal() const = 0;
  virtual void matmul_device(int precision, const double* X, int64_t ldx, int64_t k, double* Y, int64_t ldy,
                             void* stream) const = 0;
  virtual unsigned precision_mask() const { return 1u << XTPGPU_FP64; }
  Context& context() const { return *ctx_; }
  // the library handle of this operator (valid while the operator lives)
  virtual xtpgpu_mfop mfop() const {
    if (!mf_) {
      const std::vector<double> d = diagonal();
      if ((int64_t)d.size() != size_) throw std::runtime_error("MatrixFreeOperator: diagonal() has the wrong length");
      check(xtpgpu_mfop_create(ctx_->get(), size_, d.data(), &MatrixFreeOperator::trampoline,
                               const_cast<MatrixFreeOperator*>(this), precision_mask(), &mf_));
    }
    return mf_;
  }

 protected:
  mutable xtpgpu_mfop mf_ = nullptr;
  int64_t size_ = 0;

 private:
  static int trampoline(void* user, int precision, const double* X, int64_t ldx, int64_t k, double* Y, int64_t ldy,
                        void* stream) {
    try {
      static_cast<const MatrixFreeOperator*>(user)->matmul_device(precision, X, ldx, k, Y, ldy, stream);
      return 0;
    } catch (...) {
      return XTPGPU_ERR_ARG;
    }
  }
  Context* ctx_;
};

// A MatrixFreeOperator with a host matmul (column-major size x k in, size x k out), as the
// reference's test operators are written: the columns travel to the host and back per call.
class HostMatrixFreeOperator : public MatrixFreeOperator {
 public:
  using MatrixFreeOperator::MatrixFreeOperator;
  virtual void matmul(const double* X, int64_t k, double* Y) const = 0;
  void matmul_device(int, const double* X, int64_t ldx, int64_t k, double* Y, int64_t ldy, void*) const override {
    std::vector<double> in((size_t)ldx * k), out((size_t)size_ * k), back((size_t)ldy * k, 0.0);
    check(xtpgpu_memcpy_d2h(context().get(), in.data(), X, (int64_t)in.size() * 8));
    std::vector<double> packed((size_t)size_ * k);
    for (int64_t j = 0; j < k; ++j)
      std::copy(in.begin() + j * ldx, in.begin() + j * ldx + size_, packed.begin() + j * size_);
    matmul(packed.data(), k, out.data());
    for (int64_t j = 0; j < k; ++j) std::copy(out.begin() + j * size_, out.begin() + (j + 1) * size_, back.begin() + j * ldy);
    check(xtpgpu_memcpy_h2d(context().get(), Y, back.data(), (int64_t)back.size() * 8));
  }
};

// AOCoulomb::Pseudo_InvSqrt_GWBSE (aomatrix.cc:53-86): V and S are n x n column-major (aux Coulomb,
// aux overlap); returns ((S^-1/2 V S^-1/2)^-1/2 S^-1/2)^T, the factor TCMatrix::MultiplyRightWithAuxMatrix
// applies in Fill; removedfunctions counts the eigenvalues below etol.
inline std::vector<double> Pseudo_InvSqrt_GWBSE(Context& ctx, const std::vector<double>& V, const std::vector<double>& S,
                                                int64_t n, double etol = 5e-7, int64_t* removedfunctions = nullptr) {
  if (n <= 0 || V.size() != (size_t)(n * n) || S.size() != (size_t)(n * n))
    throw std::runtime_error("Pseudo_InvSqrt_GWBSE: V and S must both hold n x n values");
  std::vector<double> out((size_t)(n * n));
  check(xtpgpu_aux_pseudo_invsqrt(ctx.get(), V.data(), S.data(), n, etol, out.data(), removedfunctions));
  return out;
}

struct BSEOperatorOptions {  // BSEOperator_Options (bse_operator.h:32-38)
  int64_t homo, rpamin, qpmin, vmin, cmax;
};

template <int cqp, int cx, int cd, int cd2>
class BSEOperator : public MatrixFreeOperator {
 public:
  BSEOperator(Context& ctx, TCMatrix& Mmn, const double* eps0inv, const double* Hqp, int64_t ldhqp,
              const BSEOperatorOptions& o, unsigned precision_mask = 1u << XTPGPU_FP64)
      : MatrixFreeOperator(ctx), mask_(precision_mask | (1u << XTPGPU_FP64)) {
    static_assert(!(cd != 0 && cd2 != 0), "Hamiltonian cannot contain Hd and Hd2 at the same time");
    check(xtpgpu_bse_create(ctx.get(), Mmn.get(), eps0inv, Hqp, ldhqp, o.homo, o.rpamin, o.qpmin, o.vmin, o.cmax, cqp,
                            cx, cd, cd2, precision_mask, &h_));
    check(xtpgpu_bse_size(h_, &size_));
  }
  ~BSEOperator() override {
    if (mf_) xtpgpu_mfop_destroy(mf_);
    mf_ = nullptr;
    xtpgpu_bse_destroy(h_);
  }
  std::vector<double> diagonal() const override {
    std::vector<double> d(size_);
    check(xtpgpu_bse_diagonal(h_, d.data()));
    return d;
  }
  // Y (size x k) = H X (size x k), column-major host buffers
  void matmul(const double* X, int64_t k, double* Y, int precision = XTPGPU_FP64) const {
    check(xtpgpu_bse_matmul(h_, precision, X, size_, k, Y, size_));
  }
  void matmul_device(int precision, const double* X, int64_t ldx, int64_t k, double* Y, int64_t ldy,
                     void*) const override {
    check(xtpgpu_bse_matmul_device(h_, precision, X, ldx, k, Y, ldy));
  }
  unsigned precision_mask() const override { return mask_; }
  // the fused device operator itself, not a callback
  xtpgpu_mfop mfop() const override {
    if (!mf_) check(xtpgpu_mfop_from_bse(h_, &mf_));
    return mf_;
  }
  xtpgpu_bseop get() const { return h_; }

 private:
  xtpgpu_bseop h_ = nullptr;
  unsigned mask_;
};

// RPA (xtp/include/votca/xtp/rpa.h:35-115): dielectric builds and the two-particle Hamiltonian.
class RPA {
 public:
  RPA(Context& ctx, TCMatrix& Mmn) : ctx_(ctx.get()), mmn_(Mmn.get()), aux_(Mmn.auxsize()) {}
  void configure(int64_t homo, int64_t rpamin, int64_t rpamax) {
    homo_ = homo;
    rpamin_ = rpamin;
    rpamax_ = rpamax;
  }
  void setRPAInputEnergies(const std::vector<double>& e) { energies_ = e; }
  // eps(i w) (aux x aux, column-major) at one imaginary frequency (rpa.cc:75-140)
  std::vector<double> calculate_epsilon_i(double frequency, int precision = XTPGPU_FP64) const {
    return epsilon(XTPGPU_EPS_IMAG, frequency, 0.0, precision);
  }
  // eps at a real frequency (rpa.cc:124-132) or a complex one (rpa.cc:145-202)
  std::vector<double> calculate_epsilon_r(double frequency, int precision = XTPGPU_FP64) const {
    return epsilon(XTPGPU_EPS_REAL, frequency, 0.0, precision);
  }
  std::vector<double> calculate_epsilon_r(double re, double im, int precision = XTPGPU_FP64) const {
    return epsilon(XTPGPU_EPS_COMPLEX, re, im, precision);
  }
  struct rpa_eigensolution {  // rpa.h:78-82
    std::vector<double> omega, XpY;
    double ERPA_correlation = 0.0;
  };
  // RPA::Diagonalize_H2p (rpa.cc:204-258)
  rpa_eigensolution Diagonalize_H2p() const {
    const int64_t n = (homo_ + 1 - rpamin_) * (rpamax_ - homo_);
    rpa_eigensolution sol;
    sol.omega.resize(n);
    sol.XpY.resize(n * n);
    check_energies();
    check(xtpgpu_rpa_h2p(ctx_, mmn_, energies_.data(), homo_, rpamin_, rpamax_, sol.omega.data(), sol.XpY.data(),
                         &sol.ERPA_correlation));
    return sol;
  }

  // the inputs of the eps builds, for EpsBatch
  xtpgpu_mmn mmn() const { return mmn_; }
  const std::vector<double>& energies() const { return energies_; }
  int64_t homo() const { return homo_; }
  int64_t rpamin() const { return rpamin_; }
  int64_t rpamax() const { return rpamax_; }
  int64_t auxsize() const { return aux_; }

 private:
  // the library reads rpamax - rpamin + 1 input energies
  void check_energies() const {
    if ((int64_t)energies_.size() != rpamax_ - rpamin_ + 1)
      throw std::runtime_error("RPA: RPA input energies do not match rpamax - rpamin + 1");
  }
  std::vector<double> epsilon(int kind, double re, double im, int precision) const {
    check_energies();
    std::vector<double> out(aux_ * aux_);
    check(xtpgpu_rpa_epsilon(ctx_, mmn_, energies_.data(), homo_, rpamin_, rpamax_, kind, &re, &im, 1, 1e-4,
                             precision, out.data()));
    return out;
  }
  xtpgpu_ctx ctx_;
  xtpgpu_mmn mmn_;
  int64_t aux_, homo_ = 0, rpamin_ = 0, rpamax_ = 0;
  std::vector<double> energies_;
};

// eps(i w) of many segments in one hipGraph-replayed persistent launch (xtpgpu_epsbatch_*): the
// per-segment RPA::calculate_epsilon_i of the jobs ParallelXJobCalc runs (parallelxjobcalc.cc:52-111).
// The RPA objects (configured, with their input energies) must outlive the batch.
class EpsBatch {
 public:
  EpsBatch(Context& ctx, const std::vector<const RPA*>& rpas, const std::vector<double>& omegas,
           int precision = XTPGPU_FP64, int grid = 0) {
    const size_t n = rpas.size();
    if (n == 0 || omegas.empty()) throw std::runtime_error("EpsBatch: no segments or no frequencies");
    std::vector<xtpgpu_mmn> t(n);
    std::vector<const double*> e(n);
    std::vector<int64_t> homo(n), rmin(n), rmax(n);
    for (size_t s = 0; s < n; ++s) {
      const RPA& r = *rpas[s];
      if ((int64_t)r.energies().size() != r.rpamax() - r.rpamin() + 1)
        throw std::runtime_error("EpsBatch: RPA input energies do not match rpamax - rpamin + 1");
      t[s] = r.mmn();
      e[s] = r.energies().data();
      homo[s] = r.homo();
      rmin[s] = r.rpamin();
      rmax[s] = r.rpamax();
      aux_.push_back(r.auxsize());
    }
    nw_ = (int)omegas.size();
    check(xtpgpu_epsbatch_create(ctx.get(), (int)n, t.data(), e.data(), homo.data(), rmin.data(), rmax.data(), nw_,
                                 omegas.data(), precision, grid, &h_));
  }
  ~EpsBatch() { xtpgpu_epsbatch_destroy(h_); }
  EpsBatch(const EpsBatch&) = delete;
  EpsBatch& operator=(const EpsBatch&) = delete;
  int autotune() {
    int g = 0;
    check(xtpgpu_epsbatch_autotune(h_, &g));
    return g;
  }
  void run(bool use_graph = true) { check(xtpgpu_epsbatch_run(h_, use_graph ? 1 : 0)); }
  // segment s: P x P x n_omega, column-major
  std::vector<double> result(int s) const {
    std::vector<double> out((size_t)aux_.at(s) * aux_.at(s) * nw_);
    check(xtpgpu_epsbatch_result(h_, s, out.data()));
    return out;
  }

 private:
  xtpgpu_epsbatch h_ = nullptr;
  int nw_ = 0;
  std::vector<int64_t> aux_;
};

// DavidsonSolver (xtp/include/votca/xtp/davidsonsolver.h:47-334) with the device-resident loop:
// the same options and string values, eigenvalues()/eigenvectors() after solve().
class DavidsonSolver {
 public:
  DavidsonSolver() {
    o_.tolerance = 1e-4;
    o_.correction = 0;
    o_.size_update = 1;
    o_.iter_max = 50;
    o_.max_search_space = 0;
    o_.precision = XTPGPU_FP64;
    o_.refine_tolerance = 1e-4;
    o_.refine_max = 0;
    o_.precision_start = -1;
    o_.switch_tolerance = 0.0;
  }
  void set_iter_max(int n) { o_.iter_max = n; }
  void set_max_search_space(int n) { o_.max_search_space = n; }
  void set_tolerance(double t) { o_.tolerance = t; }
  void set_tolerance(const std::string& t) {  // davidsonsolver.cc:122-134
    if (t == "loose") o_.tolerance = 1e-3;
    else if (t == "normal") o_.tolerance = 1e-4;
    else if (t == "strict") o_.tolerance = 1e-5;
    else if (t == "lapack") o_.tolerance = 1e-9;
    else throw std::runtime_error(t + " is not a valid Davidson tolerance");
  }
  void set_correction(const std::string& m) {  // davidsonsolver.cc:111-120
    if (m == "DPR") o_.correction = 0;
    else if (m == "OLSEN") o_.correction = 1;
    else throw std::runtime_error(m + " is not a valid Davidson correction method");
  }
  void set_size_update(const std::string& u) {  // davidsonsolver.cc:136-147
    if (u == "min") o_.size_update = 0;
    else if (u == "safe") o_.size_update = 1;
    else if (u == "max") o_.size_update = 2;
    else throw std::runtime_error(u + " is not a valid Davidson update");
  }
  // operator precision of the iterations (XTPGPU_BF16 / XTPGPU_BF16X3 add FP64 refinement rounds)
  void set_precision(int precision, double refine_tolerance, int refine_max = 4, int start = -1,
                     double switch_tolerance = 1e-5) {
    o_.precision = precision;
    o_.refine_tolerance = refine_tolerance;
    o_.refine_max = precision == XTPGPU_FP64 ? 0 : refine_max;
    o_.precision_start = start;
    o_.switch_tolerance = switch_tolerance;
  }
  // DavidsonSolver::solve(A, neigen) (SYMM) on any MatrixFreeOperator: a BSEOperator runs its fused
  // device contraction, other operators their matmul_device. Returns false when not converged
  // (unconverged roots zeroed, as storeNotConvergedData does).
  bool solve(const MatrixFreeOperator& A, int64_t neigen) {
    n_ = A.rows();
    eval_.assign(neigen, 0.0);
    evec_.assign(n_ * neigen, 0.0);
    return finish(xtpgpu_davidson_solve_mf(A.mfop(), neigen, &o_, eval_.data(), evec_.data(), n_, &info_));
  }
  // DavidsonSolver::solve on HamiltonianOperator<R, C> with the HAM matrix type (set_matrix_type("HAM")
  // in the reference): eigenvalues() and the unit 2*size-long eigenvectors()
  bool solve_ham(const MatrixFreeOperator& R, const MatrixFreeOperator& C, int64_t neigen) {
    n_ = 2 * R.rows();
    eval_.assign(neigen, 0.0);
    evec_.assign(n_ * neigen, 0.0);
    return finish(xtpgpu_davidson_solve_ham_mf(R.mfop(), C.mfop(), neigen, &o_, nullptr, 0, 0, eval_.data(),
                                               evec_.data(), n_, &info_));
  }
  // many segments' TDA problems (the jobs ParallelXJobCalc runs one per thread,
  // parallelxjobcalc.cc:52-111) on `workers` streams in `order`: ops are BSEOperator::get()
  // handles; evals gets one row of neigen per operator, seconds (optional) each segment's completion
  // time from the call's start. Returns false when a segment did not converge.
  // order: XTPGPU_BATCH_*, optionally | XTPGPU_BATCH_GROUPED (products issued together)
  bool solve_batched(const std::vector<xtpgpu_bseop>& ops, int64_t neigen, int workers,
                     std::vector<std::vector<double>>& evals, std::vector<double>* seconds = nullptr,
                     int order = XTPGPU_BATCH_LARGEST_FIRST) {
    const size_t n = ops.size();
    std::vector<double> ev(n * neigen, 0.0), sec(n, 0.0);
    const int rc = xtpgpu_davidson_solve_batched((int64_t)n, ops.data(), neigen, &o_, workers, order, 0, ev.data(),
                                                 nullptr, nullptr, sec.data(), nullptr);
    evals.assign(n, {});
    for (size_t s = 0; s < n; ++s) evals[s].assign(ev.begin() + s * neigen, ev.begin() + (s + 1) * neigen);
    if (seconds) *seconds = sec;
    return finish(rc);
  }
  // full BSE through the HamiltonianOperator [[A, B], [-B, -A]] (BSE::Solve_nonhermitian_Davidson)
  template <class OpA, class OpB>
  bool solve_full(const OpA& A, const OpB& B, int64_t neigen) {
    n_ = A.rows();
    eval_.assign(neigen, 0.0);
    evec_.assign(n_ * neigen, 0.0);
    evec2_.assign(n_ * neigen, 0.0);
    return finish(xtpgpu_davidson_solve_full(A.get(), B.get(), neigen, &o_, eval_.data(), evec_.data(),
                                             evec2_.data(), n_, &info_));
  }
  const std::vector<double>& eigenvalues() const { return eval_; }
  const std::vector<double>& eigenvectors() const { return evec_; }   // size x neigen
  const std::vector<double>& eigenvectors2() const { return evec2_; } // Y of the full BSE
  int num_iterations() const { return info_.iterations; }
  const xtpgpu_davidson_info& info() const { return info_; }

 private:
  bool finish(int rc) {
    if (rc == XTPGPU_ERR_NOT_CONVERGED) return false;
    check(rc);
    return true;
  }
  xtpgpu_davidson_opts o_{};
  xtpgpu_davidson_info info_{};
  int64_t n_ = 0;
  std::vector<double> eval_, evec_, evec2_;
};

// GW::options QP-solver fields with the reference's defaults (gw.h:55-100): grid solver, adaptive scan
// with dense fallback, bisection, restricted search, 0.75 Ha half window, 0.002 Ha dense spacing,
// 0.025 Ha shells.
inline xtpgpu_qp_opts qp_defaults() {
  xtpgpu_qp_opts o{};
  o.solver = 1;
  o.qp_solver_alpha = 0.75;
  o.g_sc_limit = 1e-5;
  o.g_sc_max_iterations = 50;
  o.full_window_half_width = 0.75;
  o.dense_spacing = 0.002;
  o.restrict_search = 1;
  o.zero_margin = 1e-6;
  o.virtual_min_energy = -0.1;
  o.min_accepted_Z = 0.05;
  o.max_accepted_Z = 1.5;
  o.brent_max_iter = 0;
  o.grid_mode = 0;
  o.root_finder = 0;
  o.adaptive_shell_width = 0.025;
  o.adaptive_shell_count = 0;
  o.gw_sc_iteration = 0;
  return o;
}
// the legacy qp_grid_steps / qp_grid_spacing pair, as qp_solver::NormalizeGridSearchOptions maps
// it (qp_solver_utils.h:115-170)
inline void qp_legacy_grid(xtpgpu_qp_opts& o, int steps, double spacing) {
  if (steps <= 1 || spacing <= 0.0) throw std::runtime_error("qp_legacy_grid: need steps > 1 and spacing > 0");
  const double full = spacing * (double)(steps - 1);
  const int coarse = std::max(21, steps / 4);
  o.full_window_half_width = 0.5 * full;
  o.dense_spacing = spacing;
  o.adaptive_shell_width = full / (double)(coarse - 1);
}

// Sigma_PPM (xtp/src/libxtp/self_energy_evaluators/sigma_ppm.h) with the exchange of Sigma_base
// (sigma_base.h:33-106) and GW::SolveQP (gw.cc:323-410) on the device. PrepareScreening rotates
// the tensor in place (Mmn * phi), as the reference does.
class Sigma_PPM {
 public:
  Sigma_PPM(Context& ctx, TCMatrix& Mmn) : ctx_(ctx.get()), mmn_(Mmn.get()), aux_(Mmn.auxsize()) {}
  ~Sigma_PPM() { xtpgpu_sigma_destroy(h_); }
  Sigma_PPM(const Sigma_PPM&) = delete;
  Sigma_PPM& operator=(const Sigma_PPM&) = delete;
  void configure(int64_t homo, int64_t qpmin, int64_t qpmax, int64_t rpamin, int64_t rpamax, double eta) {
    homo_ = homo;
    qpmin_ = qpmin;
    qpmax_ = qpmax;
    rpamin_ = rpamin;
    rpamax_ = rpamax;
    eta_ = eta;
  }
  void setRPAInputEnergies(const std::vector<double>& e) { energies_ = e; }
  // Sigma_base::CalcExchangeMatrix (sigma_base.cc:36-52): qptotal x qptotal, column-major
  std::vector<double> CalcExchangeMatrix() const {
    const int64_t q = qpmax_ - qpmin_ + 1;
    std::vector<double> out(q * q);
    check(xtpgpu_sigma_x(ctx_, mmn_, homo_, rpamin_, qpmin_, qpmax_, out.data()));
    return out;
  }
  // PPM::PPM_construct_parameters + Sigma_PPM::PrepareScreening (ppm.cc:31-59, sigma_ppm.cc:32-35)
  void PrepareScreening() {
    if ((int64_t)energies_.size() != rpamax_ - rpamin_ + 1)
      throw std::runtime_error("Sigma_PPM: RPA input energies do not match rpamax - rpamin + 1");
    weight_.resize(aux_);
    freq_.resize(aux_);
    check(xtpgpu_ppm_prepare(ctx_, mmn_, energies_.data(), homo_, rpamin_, rpamax_, weight_.data(), freq_.data()));
    xtpgpu_sigma_destroy(h_);
    h_ = nullptr;
    check(xtpgpu_sigma_ppm_create(ctx_, mmn_, energies_.data(), weight_.data(), freq_.data(), homo_, rpamin_, qpmin_,
                                  eta_, &h_));
  }
  const std::vector<double>& getPpm_weight() const { return weight_; }
  const std::vector<double>& getPpm_freq() const { return freq_; }
  // CalcCorrelationDiagElement (+Derivative) for every QP level at its own frequency
  std::vector<double> CalcCorrelationDiag(const std::vector<double>& frequencies,
                                          std::vector<double>* derivative = nullptr) const {
    const int64_t q = (int64_t)frequencies.size();
    if (q > qpmax_ - qpmin_ + 1) throw std::runtime_error("CalcCorrelationDiag: more frequencies than QP levels");
    std::vector<int64_t> lv(q);
    for (int64_t l = 0; l < q; ++l) lv[l] = l;
    std::vector<double> out(q);
    if (derivative) derivative->resize(q);
    check(xtpgpu_sigma_ppm_diag(h_, q, lv.data(), frequencies.data(), out.data(),
                                derivative ? derivative->data() : nullptr));
    return out;
  }
  // Sigma_base::CalcCorrelationOffDiag (sigma_base.cc:65-78); the diagonal is left zero
  std::vector<double> CalcCorrelationOffDiag(const std::vector<double>& frequencies) const {
    const int64_t q = (int64_t)frequencies.size();
    if (q != qpmax_ - qpmin_ + 1) throw std::runtime_error("CalcCorrelationOffDiag: frequencies must have qptotal entries");
    std::vector<double> out(q * q, 0.0);
    check(xtpgpu_sigma_ppm_offdiag(h_, q, frequencies.data(), out.data()));
    return out;
  }
  // GW::SolveQP: w_l = intercept_l + Sigma_c(l, w_l) for all QP levels
  std::vector<double> SolveQP(const std::vector<double>& intercepts, const std::vector<double>& frequencies,
                              const xtpgpu_qp_opts& opts, std::vector<int>* converged = nullptr) const {
    const int64_t q = (int64_t)intercepts.size();
    if ((int64_t)frequencies.size() != q || q != qpmax_ - qpmin_ + 1)
      throw std::runtime_error("SolveQP: intercepts and frequencies must both have qptotal entries");
    std::vector<double> out(q);
    std::vector<int> conv(q);
    check(xtpgpu_gw_solve_qp(h_, q, intercepts.data(), frequencies.data(), &opts, out.data(), conv.data()));
    if (converged) *converged = conv;
    return out;
  }

 private:
  xtpgpu_ctx ctx_;
  xtpgpu_mmn mmn_;
  int64_t aux_, homo_ = 0, qpmin_ = 0, qpmax_ = 0, rpamin_ = 0, rpamax_ = 0;
  double eta_ = 1e-3;
  std::vector<double> energies_, weight_, freq_;
  xtpgpu_sigma h_ = nullptr;
};

// BSE::AdjustHqpSize (bse.cc:147-186): the GW-window Hqp (gw x gw, column-major) cut or extended to
// the BSE window [vmin, cmax]; levels outside the GW window take their RPA input energy.
inline std::vector<double> AdjustHqpSize(const std::vector<double>& Hqp, const std::vector<double>& rpa_energies,
                                         int64_t homo, int64_t rpamin, int64_t qpmin, int64_t qpmax, int64_t vmin,
                                         int64_t cmax) {
  const int64_t gw = qpmax - qpmin + 1, hs = (homo - vmin + 1) + (cmax - homo);
  if (gw <= 0 || (int64_t)Hqp.size() != gw * gw) throw std::runtime_error("AdjustHqpSize: Hqp is not qptotal x qptotal");
  std::vector<double> out((size_t)hs * hs);
  check(xtpgpu_bse_adjust_hqp(Hqp.data(), gw, rpa_energies.data(), (int64_t)rpa_energies.size(), homo, rpamin, qpmin,
                              qpmax, vmin, cmax, out.data(), hs));
  return out;
}

using SingletOperator_TDA = BSEOperator<1, 2, 1, 0>;
using TripletOperator_TDA = BSEOperator<1, 0, 1, 0>;
using SingletOperator_BTDA_B = BSEOperator<0, 2, 0, 1>;
using HqpOperator = BSEOperator<1, 0, 0, 0>;
using HxOperator = BSEOperator<0, 1, 0, 0>;
using HdOperator = BSEOperator<0, 0, 1, 0>;
using Hd2Operator = BSEOperator<0, 0, 0, 1>;

}  // namespace xtpgpu

#endif  // XTPGPU_HPP
