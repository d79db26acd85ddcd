/*
 * xtpgpu -- C-ABI of the MI355X (gfx950) GW-BSE dense-contraction hot path.
 *
 * Every entry point returns an int status (XTPGPU_OK = 0, negative on error) and never
 * lets a C++ exception cross the boundary; xtpgpu_last_error() returns the message of
 * the last failure on the calling thread. Host pointers are borrowed for the duration of
 * a call only (copy-in / copy-out); device state lives in opaque handles that own their
 * memory until the matching *_destroy. Matrices at the boundary are column-major FP64
 * with explicit leading dimensions and 64-bit sizes, as Eigen::MatrixXd in the reference
 * (tools/include/votca/tools/types.h). Index conventions follow the reference:
 *   - an Mmn tensor is mtotal slices, each (ntotal x auxsize), slice m / row n / column P
 *     (xtp/include/votca/xtp/threecenter.h:70-142);
 *   - the BSE compound index is I(v,c) = ctotal*v + c (xtp/include/votca/xtp/vc2index.h:36-54).
 * Calls on one context are stream-ordered; use one context per calling thread.
 */
#ifndef XTPGPU_H
#define XTPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  XTPGPU_OK = 0,
  XTPGPU_ERR_HIP = -1,            /* HIP runtime / library failure (reference: std::runtime_error, cudamatrix.cc:25-37) */
  XTPGPU_ERR_OOM = -2,            /* device memory request exceeds free memory (cudamatrix.cc:97-113) */
  XTPGPU_ERR_SHAPE = -3,          /* inconsistent sizes / indices */
  XTPGPU_ERR_NOT_CONVERGED = -4,  /* iterative solver did not converge (davidsonsolver.cc:524-547) */
  XTPGPU_ERR_ARG = -5,            /* invalid argument */
  XTPGPU_ERR_NUMERIC = -6         /* numerical failure, e.g. Gram-Schmidt linear dependence (davidsonsolver.cc:490-494) */
};

/* Arithmetic of a contraction. FP64: plain FP64 MFMA (the drop-in form).
 * BF16: BF16 operands, FP32 accumulation. BF16X3: operands split a = a_hi + a_lo in
 * BF16, products hi*hi + hi*lo + lo*hi accumulated in FP32. */
enum { XTPGPU_FP64 = 0, XTPGPU_BF16 = 1, XTPGPU_BF16X3 = 3 };

/* RPA frequency kinds: RPA::calculate_epsilon_i (rpa.h:47-49), calculate_epsilon_r(double)
 * (rpa.h:51-53), calculate_epsilon_r(std::complex<double>) (rpa.cc:145-202). */
enum { XTPGPU_EPS_IMAG = 0, XTPGPU_EPS_REAL = 1, XTPGPU_EPS_COMPLEX = 2 };

typedef struct xtpgpu_ctx_s* xtpgpu_ctx;
typedef struct xtpgpu_mmn_s* xtpgpu_mmn;
typedef struct xtpgpu_bseop_s* xtpgpu_bseop;
typedef struct xtpgpu_mfop_s* xtpgpu_mfop;

/* ---------------------------------------------------------------- errors, devices */
const char* xtpgpu_last_error(void);
int xtpgpu_device_count(int* count);
/* One context per (device, stream); replaces OpenMP_CUDA::SetNoGPUs / GPU_data
 * (xtp/src/libxtp/openmp_cuda.cc:28-71). flags must be 0. */
int xtpgpu_ctx_create(int device, unsigned flags, xtpgpu_ctx* out);
int xtpgpu_ctx_destroy(xtpgpu_ctx ctx);
int xtpgpu_ctx_synchronize(xtpgpu_ctx ctx);
/* Free the context's grow-only scratch (kept between calls: the eps builds' scaled transition
 * copies, up to XTPGPU_EPS_SCALE_BYTES, and their K-split partials); the next call that needs it
 * allocates it again. No counterpart in the reference (its GPU buffers live per call). */
int xtpgpu_ctx_release_scratch(xtpgpu_ctx ctx);
/* Live kernel timing: when enabled, HIP events are recorded on the context's stream around
 * the dominant launches (names: "rpa_eps_gemm", "hd_stage1" / "hd_stage2" for the split BF16x3
 * form, "hd_stage1_bf16" / "hd_stage2_bf16" for plain BF16, "hd_stage1_fp64" / "hd_stage2_fp64"
 * for FP64, "hd2", "hx", "hqp"), together with each
 * launch's algorithmic FLOPs and bytes. xtpgpu_ctx_timer synchronizes the stream and returns
 * the totals. */
int xtpgpu_ctx_set_profiling(xtpgpu_ctx ctx, int enable);
int xtpgpu_ctx_timer(xtpgpu_ctx ctx, const char* name, double* total_ms, int64_t* launches,
                     double* flops, double* bytes);
int xtpgpu_ctx_reset_timers(xtpgpu_ctx ctx);
/* Stream-ordered device copies of host memory, for callers that keep data resident. */
int xtpgpu_malloc(xtpgpu_ctx ctx, int64_t bytes, void** dptr);
int xtpgpu_free(xtpgpu_ctx ctx, void* dptr);
int xtpgpu_memcpy_h2d(xtpgpu_ctx ctx, void* dst, const void* src, int64_t bytes);
int xtpgpu_memcpy_d2h(xtpgpu_ctx ctx, void* dst, const void* src, int64_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte vectors; bytes and both pointers
 * multiples of 16) on the context's stream. No counterpart in the reference: a utility for
 * resident data, and the known-size reference kernel for HBM traffic counters. */
int xtpgpu_copy_device(xtpgpu_ctx ctx, void* dst, const void* src, int64_t bytes);

/* ---------------------------------------------------------------- TCMatrix_gwbse */
/* TCMatrix_gwbse::Initialize (threecenter.cc:29-48): zero tensor, slices m in [mmin,mmax],
 * rows n in [nmin,nmax], auxsize columns. */
int xtpgpu_mmn_create(xtpgpu_ctx ctx, int64_t mmin, int64_t mmax, int64_t nmin, int64_t nmax,
                      int64_t auxsize, xtpgpu_mmn* out);
int xtpgpu_mmn_destroy(xtpgpu_mmn mmn);
/* Copy one slice Mmn_[m] (ntotal x auxsize, column-major, leading dim ld) in / out
 * (TCMatrix_gwbse::operator[], threecenter.h:101-104). m is relative to mmin. */
int xtpgpu_mmn_set_slice(xtpgpu_mmn mmn, int64_t m, const double* slice, int64_t ld);
int xtpgpu_mmn_get_slice(xtpgpu_mmn mmn, int64_t m, double* slice, int64_t ld);
/* TCMatrix_gwbse::MultiplyRightWithAuxMatrix (threecenter.cc:54-65): Mmn_[m] <- Mmn_[m] * A
 * for every slice; A is auxsize x auxsize column-major on the host. */
int xtpgpu_mmn_multiply_right(xtpgpu_mmn mmn, const double* A, int64_t lda);
/* TCMatrix_gwbse::Fill3cMO AO->MO step (libint2_calls.cc:595-651): from the AO three-centre
 * integrals (Q|mu nu) (auxsize x nbf x nbf, element (Q,mu,nu) at Q*nbf*nbf + nu*nbf + mu)
 * and MO coefficients C (nbf x nmo column-major), sets Mmn_[m](n,Q) = sum C(mu,n)(Q|mu nu)C(nu,m)
 * for the tensor's m/n windows. The aux metric is applied afterwards with
 * xtpgpu_mmn_multiply_right, as TCMatrix_gwbse::Fill does (threecenter.cc:72-90). */
int xtpgpu_mmn_fill_ao2mo(xtpgpu_mmn mmn, const double* ao3c, int64_t nbf, const double* C,
                          int64_t ldc, int64_t nmo);
/* Synthetic tensor for benchmarks: M^P_{mn} = sigma_P * g(seed, P, min(m,n), max(m,n)),
 * sigma_P = sigma0 * exp(-decay * P / auxsize), g a counter-based standard normal variate
 * (splitmix64 hash of (seed,P,i,j), Irwin-Hall sum of 12 uniforms: integer arithmetic, so a
 * host restatement reproduces it bit for bit); symmetric in (m,n) as a three-centre quantity is. */
int xtpgpu_mmn_fill_synthetic(xtpgpu_mmn mmn, uint64_t seed, double sigma0, double decay);
int xtpgpu_mmn_sizes(xtpgpu_mmn mmn, int64_t* mtotal, int64_t* ntotal, int64_t* auxsize);
/* dst <- src for two tensors of one shape on one device: restores a kept copy where the reference
 * calls TCMatrix_gwbse::Rebuild (threecenter.h:97-99, refilling from the DFT orbitals) between
 * evGW iterations. */
int xtpgpu_mmn_copy(xtpgpu_mmn dst, xtpgpu_mmn src);
/* Write a transition block: Mmn_[v](c, P) = m[(P*vtotal + v)*ctotal + c] for slices v in
 * [vmin, vmin+vtotal) and rows c in [cmin, cmin+ctotal) (indices relative to mmin / nmin), and
 * with symmetric != 0 also Mmn_[c](v, P), as a three-centre tensor is symmetric in (m, n).
 * m is a host buffer (auxsize x vtotal x ctotal). */
int xtpgpu_mmn_set_block(xtpgpu_mmn mmn, int64_t vmin, int64_t vtotal, int64_t cmin, int64_t ctotal,
                         const double* m, int symmetric);

/* ---------------------------------------------------------------- RPA */
/* RPA::calculate_epsilon<imag> and calculate_epsilon_r(complex) (rpa.cc:75-202):
 * eps(w) = 1 + sum_{v<n_occ} M_v^T diag(d_v(w)) M_v, M_v = Mmn_[v].bottomRows(n_unocc),
 * for n_omega frequencies in one pass. energies: rpatotal RPA input energies
 * (RPA::setRPAInputEnergies). out: auxsize x auxsize x n_omega column-major (host).
 * eta: RPA broadening (the reference fixes 1e-4, rpa.h:90). */
int xtpgpu_rpa_epsilon(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo,
                       int64_t rpamin, int64_t rpamax, int kind, const double* omega_re,
                       const double* omega_im, int n_omega, double eta, int precision,
                       double* out);
/* Same, result left on the device (out_dev: auxsize*auxsize*n_omega doubles). */
int xtpgpu_rpa_epsilon_device(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo,
                              int64_t rpamin, int64_t rpamax, int kind, const double* omega_re,
                              const double* omega_im, int n_omega, double eta, int precision,
                              double* out_dev);

/* ---------------------------------------------------------------- BSE operator */
/* BSE_OPERATOR<cqp,cx,cd,cd2>(eps0inv, Mmn, Hqp) + configure(BSEOperator_Options)
 * (bse_operator.h:32-48, bse_operator.cc:28-37). Hqp is the BSE-sized (vtotal+ctotal)^2 QP
 * Hamiltonian (BSE::AdjustHqpSize, bse.cc:147-186). The device operator packs the Mmn blocks
 * it contracts (c c', v v', v c) into K-contiguous layouts once; Mmn must not change while the
 * operator lives (the reference holds a const reference with the same rule, bse_operator.h:73-75).
 * precision_mask: bit (1<<XTPGPU_FP64) always; add (1<<XTPGPU_BF16)|(1<<XTPGPU_BF16X3) to
 * also keep BF16 shadows for the mixed-precision forms. */
int xtpgpu_bse_create(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* eps0inv, const double* Hqp,
                      int64_t ldhqp, int64_t homo, int64_t rpamin, int64_t qpmin, int64_t vmin,
                      int64_t cmax, int cqp, int cx, int cd, int cd2, unsigned precision_mask,
                      xtpgpu_bseop* out);
int xtpgpu_bse_destroy(xtpgpu_bseop op);
/* A source of three-centre slices by auxiliary chunk, for operators too large for a resident Mmn
 * window: fills M^P_{mn} for P in [p0, p0 + np), m, n in [lmin, lmax] (tensor indices, level -
 * rpamin) into dst (device memory, aux-major: dst[(P - p0) * W * W + (m - lmin) * W + (n - lmin)],
 * W = lmax - lmin + 1), issued on `stream` (the context's hipStream_t); returns 0 on success.
 * The reference fills the same slices from the AO three-centre integrals (TCMatrix_gwbse::Fill,
 * threecenter.cc:29-48); a source may do that per chunk, or regenerate synthetic data. */
typedef int (*xtpgpu_mmn_chunk_fn)(void* user, int64_t p0, int64_t np, int64_t lmin, int64_t lmax,
                                   double* dst, void* stream);
/* The library's synthetic generator (xtpgpu_mmn_fill_synthetic's law: element (m, n, P) of a
 * tensor with `aux` functions is sigma_P * N(seed, P, min(m, n), max(m, n)), sigma_P = sigma0 *
 * exp(-decay P / aux)) as a chunk source; user points to an xtpgpu_synthetic_source. */
typedef struct {
  uint64_t seed;
  double sigma0, decay;
  int64_t aux;
} xtpgpu_synthetic_source;
int xtpgpu_synthetic_chunk(void* user, int64_t p0, int64_t np, int64_t lmin, int64_t lmax, double* dst,
                           void* stream);
/* xtpgpu_bse_create with the operator packed from a chunk source instead of a resident tensor: the
 * source is asked for the BSE window [vmin, cmax] (as tensor indices vmin - rpamin ...) in aux
 * chunks of at most chunk_bytes (0: 4 GiB), so no full Mmn window is ever held. With
 * XTPGPU_BSE_STREAM_FP64 in flags the FP64 M^P_{cc'} block (the largest, aux * ctotal^2 doubles) is
 * not kept: FP64 products re-read it from the source chunk by chunk (the source must outlive the
 * operator); the mixed forms use the resident BF16 planes. No counterpart in the reference, whose
 * TCMatrix_gwbse always holds the window in host memory.
 * Screening: eps0inv (aux values) as for xtpgpu_bse_create, for slices already rotated to the
 * eigenbasis of eps(0) (BSE::SetupDirectInteractionOperator, bse.cc:188-204); or, with
 * epsinv_matrix (aux x aux, column-major; eps0inv ignored), the unrotated slices and the full
 * inverse dielectric matrix, W_PQ = (eps^-1(0))_PQ, applied to the valence block
 * (sum_PQ M^P_vv' W_PQ M^Q_cc' = the reference's sum over the rotated basis), so the source needs no
 * rotation across aux chunks. epsinv_matrix supports the direct term (cd) only, not cd2. */
enum { XTPGPU_BSE_STREAM_FP64 = 1 };
int xtpgpu_bse_create_streamed(xtpgpu_ctx ctx, xtpgpu_mmn_chunk_fn source, void* user, int64_t aux,
                               const double* eps0inv, const double* epsinv_matrix, const double* Hqp,
                               int64_t ldhqp, int64_t homo,
                               int64_t rpamin, int64_t qpmin, int64_t vmin, int64_t cmax, int cqp, int cx,
                               int cd, int cd2, unsigned precision_mask, int flags, int64_t chunk_bytes,
                               xtpgpu_bseop* out);
/* Replace the operator's QP Hamiltonian (BSE-sized, as at create) without repacking the tensor
 * blocks; the diagonal follows. */
int xtpgpu_bse_set_hqp(xtpgpu_bseop op, const double* Hqp, int64_t ldhqp);
/* Repack the layouts of the transition block M^P_{vc} (exchange and Hd2 terms) and the diagonal
 * from `mmn` after its transition block changed (xtpgpu_mmn_set_block); the (v v') and (c c')
 * blocks and the screening are kept. For operators created from a resident tensor. */
int xtpgpu_bse_refresh(xtpgpu_bseop op, xtpgpu_mmn mmn);
/* Bytes of device memory the operator holds (packed layouts and grow-only scratch). */
int xtpgpu_bse_memory(xtpgpu_bseop op, int64_t* bytes);
/* BSE::AdjustHqpSize (bse.cc:147-186), host only: the GW-window QP Hamiltonian Hqp (gw x gw,
 * gw = qpmax - qpmin + 1, leading dim ldh) cut or extended to the BSE window [vmin, cmax]
 * (hs = (homo - vmin + 1) + (cmax - homo) levels): overlapping levels take their Hqp block, a BSE
 * level outside the GW window gets its RPA input energy rpa_energies[level - rpamin] (n_rpa values)
 * on the diagonal. out: hs x hs, leading dim ldo. */
int xtpgpu_bse_adjust_hqp(const double* Hqp, int64_t ldh, const double* rpa_energies, int64_t n_rpa,
                          int64_t homo, int64_t rpamin, int64_t qpmin, int64_t qpmax, int64_t vmin,
                          int64_t cmax, double* out, int64_t ldo);
int xtpgpu_bse_size(xtpgpu_bseop op, int64_t* size);
/* BSE_OPERATOR::diagonal (bse_operator.cc:134-175). */
int xtpgpu_bse_diagonal(xtpgpu_bseop op, double* out);
/* BSE_OPERATOR::matmul (bse_operator.cc:39-119): Y = H X for k columns, host buffers. */
int xtpgpu_bse_matmul(xtpgpu_bseop op, int precision, const double* X, int64_t ldx, int64_t k,
                      double* Y, int64_t ldy);
/* Same on device buffers (X, Y: size x k, leading dims ldx/ldy). */
int xtpgpu_bse_matmul_device(xtpgpu_bseop op, int precision, const double* X, int64_t ldx,
                             int64_t k, double* Y, int64_t ldy);
/* No counterpart in the reference. xtpgpu_bse_matmul_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): all scratch (padded X copies, the aux-chunked Hd / Hd2 intermediate, split-K
 * partials) comes from a fixed 256 MiB region per GPU that the library holds from the first
 * xtpgpu_ctx_create on that GPU until the process ends; the aux chunk shrinks to fit it. Calls
 * whose fixed scratch (X copies, partials and one aux function of the intermediate) exceeds the
 * region fail with XTPGPU_ERR_SHAPE. Every call on one GPU shares the region and it keeps the
 * last call's contents: calls that overlap on one GPU (any context) race. */
int xtpgpu_bse_matmul_device_noalloc(xtpgpu_bseop op, int precision, const double* X, int64_t ldx,
                                     int64_t k, double* Y, int64_t ldy);

/* ---------------------------------------------------------------- screening, GW */
typedef struct xtpgpu_sigma_s* xtpgpu_sigma;

/* AOCoulomb::Pseudo_InvSqrt_GWBSE (aomatrix.cc:53-86): V (aux Coulomb) and S (aux overlap), n x n
 * column-major host matrices -> out = ((S^-1/2 V S^-1/2)^-1/2 S^-1/2)^T, the right factor
 * TCMatrix_gwbse::Fill applies (threecenter.cc:72-90); eigenvalues below etol (5e-7 in the
 * reference) of either eigendecomposition are dropped and counted in removed (nullable). */
int xtpgpu_aux_pseudo_invsqrt(xtpgpu_ctx ctx, const double* V, const double* S, int64_t n, double etol,
                              double* out, int64_t* removed);
/* BSE::SetupDirectInteractionOperator (bse.cc:188-204): eps_r(energy) on the device, its
 * eigendecomposition (rocSOLVER), Mmn <- Mmn * U in place, eps0inv[P] = 1/lambda_P (0 below 1e-8). */
int xtpgpu_bse_setup_screening(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo,
                               int64_t rpamin, int64_t rpamax, double energy, double* eps0inv);
/* Sigma_base::CalcExchangeMatrix (sigma_base.cc:36-52): out (qptotal x qptotal, column-major)
 * = -sum_{n<occ,P} M^P_{l1 n} M^P_{l2 n}. */
int xtpgpu_sigma_x(xtpgpu_ctx ctx, xtpgpu_mmn mmn, int64_t homo, int64_t rpamin, int64_t qpmin,
                   int64_t qpmax, double* out);
/* RPA::Diagonalize_H2p (rpa.cc:204-258, Calculate_H2p_AmB/ApB rpa.cc:260-331, Diagonalize_H2p_C
 * rpa.cc:333-345): C = (A-B)^1/2 (A+B) (A-B)^1/2 with (A-B)_{I(v,c)} = e_c - e_v and
 * A+B = diag(A-B) + 4 V V^T, V_{I(v,c),P} = M^P_{vc}; energies are the RPA input energies from
 * rpamin. omega (N = n_occ * n_unocc) ascending; xpy (N x N, column-major, may be NULL) the X+Y
 * vectors; erpa (may be NULL) the RPA correlation energy. XTPGPU_ERR_NUMERIC if an eigenvalue
 * of C is not positive. */
int xtpgpu_rpa_h2p(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo, int64_t rpamin,
                   int64_t rpamax, double* omega, double* xpy, double* erpa);
/* PPM::PPM_construct_parameters (ppm.cc:31-59) + Sigma_PPM::PrepareScreening (sigma_ppm.cc:32-35):
 * eps_r(0), eps_i(0.5) on the device, phi = eigenvectors of eps_r(0), weight = 1 - 1/lambda,
 * freq from diag((phi^T eps_i phi)^-1); then Mmn <- Mmn * phi in place. */
int xtpgpu_ppm_prepare(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo,
                       int64_t rpamin, int64_t rpamax, double* ppm_weight, double* ppm_freq);
/* Same with the two eps builds in `precision` (XTPGPU_FP64 / XTPGPU_BF16X3 / XTPGPU_BF16); the
 * eigendecomposition, the inversion and the rotation of the tensor stay FP64. No counterpart in
 * the reference (FP64 only): the mixed-precision GW form. */
int xtpgpu_ppm_prepare_prec(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* energies, int64_t homo,
                            int64_t rpamin, int64_t rpamax, int precision, double* ppm_weight,
                            double* ppm_freq);
/* Sigma_PPM evaluator bound to a (PPM-rotated) tensor (sigma_ppm.h; options sigma_base.h:47-57). */
int xtpgpu_sigma_ppm_create(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* rpa_energies,
                            const double* ppm_weight, const double* ppm_freq, int64_t homo,
                            int64_t rpamin, int64_t qpmin, double eta, xtpgpu_sigma* out);
int xtpgpu_sigma_destroy(xtpgpu_sigma s);
/* Sigma_PPM::CalcCorrelationDiagElement / ...Derivative (sigma_ppm.cc:37-89), batched over n
 * (gw_level, omega) pairs; sigma / dsigma may be null. */
int xtpgpu_sigma_ppm_diag(xtpgpu_sigma s, int64_t n, const int64_t* gw_levels, const double* omegas,
                          double* sigma, double* dsigma);
/* Sigma_base::CalcCorrelationOffDiag with Sigma_PPM::CalcCorrelationOffDiagElement
 * (sigma_base.cc:65-78, sigma_ppm.cc:91-126); the diagonal of out is left zero. */
int xtpgpu_sigma_ppm_offdiag(xtpgpu_sigma s, int64_t qptotal, const double* frequencies, double* out);

/* Sigma_CDA (sigma_cda.cc:30-141) with ImaginaryAxisIntegration (ImaginaryAxisIntegration.cc:90-176):
 * PrepareScreening at create time -- kappa0 = eps_r(0)^-1 - 1 and D_j = -(eps_i(x_j)^-1 - 1) +
 * kappa0 exp(-(alpha x_j)^2) on the device -- with the frequency-independent contractions
 * S_{j,l,n} = (M_l D_j M_l^T)_nn and T_{l,n} = (M_l kappa0 M_l^T)_nn done once. points/weights are
 * the scaled quadrature nodes and weights (QuadratureFactory scheme and order), symmetry the
 * scheme's UseSymmetry(). The tensor must span the RPA window in n. eta is the GW option's
 * broadening, kept for the option surface: Sigma_c^CDA itself uses the RPA's fixed eta (1e-4,
 * rpa.h:90) in the residues and the imaginary-axis integrand, as sigma_cda.cc:107,125 do. */
typedef struct xtpgpu_cda_s* xtpgpu_cda;
int xtpgpu_sigma_cda_create(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* rpa_energies, int64_t homo,
                            int64_t rpamin, int64_t rpamax, int64_t qpmin, int64_t qpmax, double eta,
                            double alpha, const double* points, const double* weights, int64_t order,
                            int symmetry, xtpgpu_cda* out);
/* Same with every eps build (the screening of PrepareScreening and the residues' eps_r(|delta| + i eta))
 * in `precision`; the inversions, LU solves and contractions stay FP64. */
int xtpgpu_sigma_cda_create_prec(xtpgpu_ctx ctx, xtpgpu_mmn mmn, const double* rpa_energies, int64_t homo,
                                 int64_t rpamin, int64_t rpamax, int64_t qpmin, int64_t qpmax, double eta,
                                 double alpha, const double* points, const double* weights, int64_t order,
                                 int symmetry, int precision, xtpgpu_cda* out);
int xtpgpu_sigma_cda_destroy(xtpgpu_cda c);
/* The reference's Sigma_CDA reads the RPA input energies by reference (ImaginaryAxisIntegration.h:69,
 * sigma_cda.cc:85); after RPA::UpdateRPAInputEnergies the caller passes the new ones here
 * (rpamax - rpamin + 1 values). The dielectric matrices of the last PrepareScreening are kept. */
int xtpgpu_sigma_cda_set_rpa_energies(xtpgpu_cda c, const double* rpa_energies, int64_t n);
/* CalcCorrelationDiagElement for n (gw_level, omega) pairs (residues + tail + imaginary-axis
 * integral); dsigma (nullable) by the reference's central difference, h = 1e-3 (sigma_cda.h:56-62). */
int xtpgpu_sigma_cda_diag(xtpgpu_cda c, int64_t n, const int64_t* gw_levels, const double* omegas,
                          double* sigma, double* dsigma);

/* GW::options QP-solver part (gw.h:55-100) after qp_solver::NormalizeGridSearchOptions
 * (qp_solver_utils.h:115-170). */
typedef struct {
  int solver;                    /* 0 = fixedpoint (then grid, then linearisation); 1 = grid (then linearisation) */
  double qp_solver_alpha;        /* Newton damping, gw.h:70 (0.75) */
  double g_sc_limit;             /* QP tolerance (1e-5) */
  int g_sc_max_iterations;       /* Newton iterations (50) */
  double full_window_half_width; /* qp_full_window_half_width, Ha (0.75; legacy 0.5 * spacing * (steps - 1)) */
  double dense_spacing;          /* qp_dense_spacing, Ha (0.002; legacy qp_grid_spacing) */
  int restrict_search;           /* qp_restrict_search: occupied roots < -zero_margin, virtual > virtual_min_energy */
  double zero_margin;            /* 1e-6 */
  double virtual_min_energy;     /* -0.1 */
  double min_accepted_Z;         /* 0.05 */
  double max_accepted_Z;         /* 1.5 */
  int brent_max_iter;            /* Brent iteration cap; <= 0: g_sc_max_iterations, as the reference passes (gw.cc:444) */
  int grid_mode;                 /* qp_grid_search_mode: 0 adaptive_with_dense_fallback (default), 1 adaptive, 2 dense */
  int root_finder;               /* qp_root_finder: 0 bisection (default), 1 brent */
  double adaptive_shell_width;   /* qp_adaptive_shell_width, Ha (0.025; legacy full width / (max(21, steps/4) - 1)) */
  int adaptive_shell_count;      /* qp_adaptive_shell_count (0: use the width) */
  int gw_sc_iteration;           /* evGW iteration (the adaptive scan centres on the linearised root only at 0) */
} xtpgpu_qp_opts;

/* GW::SolveQP (gw.cc:323-410) with the grid search of gw.cc:433-739 and qp_solver_utils.h:154-712:
 * solve w = intercept_l + Sigma_c(l, w) for every QP level; the levels run concurrently and every
 * Sigma_c request round is one device pass over all of them. out[qptotal] = QP energies;
 * converged[l] = 1 where the fixed point or grid search succeeded (else the linearised value). */
int xtpgpu_gw_solve_qp(xtpgpu_sigma s, int64_t qptotal, const double* intercepts,
                       const double* frequencies, const xtpgpu_qp_opts* opts, double* out,
                       int* converged);
/* The same on a Sigma_CDA evaluator (the reference's SolveQP works on any Sigma_base,
 * sigmafactory.cc:32-36); the derivative is Sigma_CDA's central difference, h = 1e-3. */
int xtpgpu_gw_solve_qp_cda(xtpgpu_cda c, int64_t qptotal, const double* intercepts,
                           const double* frequencies, const xtpgpu_qp_opts* opts, double* out,
                           int* converged);
/* The same on a caller's Sigma_c evaluator: fn(user, n, levels, omegas, sigma, dsigma) fills
 * sigma[k] = Sigma_c(levels[k], omegas[k]) and, when dsigma is not null, its derivative; returns
 * 0 on success. It is called from library threads, one call at a time. homo / qpmin place the
 * levels for the restricted search (level l is occupied when l + qpmin <= homo). */
typedef int (*xtpgpu_sigma_fn)(void* user, int64_t n, const int64_t* levels, const double* omegas,
                               double* sigma, double* dsigma);
int xtpgpu_gw_solve_qp_fn(xtpgpu_sigma_fn fn, void* user, int64_t qptotal, int64_t homo, int64_t qpmin,
                          const double* intercepts, const double* frequencies, const xtpgpu_qp_opts* opts,
                          double* out, int* converged);

/* ---------------------------------------------------------------- Davidson */
typedef struct {
  double tolerance;        /* residual-norm tolerance (DavidsonSolver::set_tolerance, davidsonsolver.cc:122-134) */
  int correction;          /* 0 = DPR, 1 = OLSEN (davidsonsolver.cc:105-115) */
  int size_update;         /* 0 = min, 1 = safe, 2 = max (davidsonsolver.cc:136-170) */
  int iter_max;            /* default 50 */
  int max_search_space;    /* reference BSE uses 10*nmax (bse.cc:280) */
  int precision;           /* operator arithmetic during the iterations */
  /* Mixed-precision refinement (precision != FP64): after the iterations converge to
   * `tolerance` with the mixed operator, up to refine_max rounds of FP64 Rayleigh-Ritz on
   * the Ritz vectors plus DPR corrections, until the FP64 residual norm < refine_tolerance. */
  double refine_tolerance;
  int refine_max;
  /* Precision ladder (SYMM solve, precision != FP64): iterate first with precision_start
   * (e.g. BF16) until every wanted residual norm is below switch_tolerance or has stopped
   * decreasing for 20 iterations, then restart on the leading Ritz vectors (A*V recomputed)
   * and continue with `precision`. precision_start < 0 (or equal to precision) disables it.
   * No counterpart in the reference (whose Davidson is FP64 only). */
  int precision_start;
  double switch_tolerance;
} xtpgpu_davidson_opts;

typedef struct {
  int converged;           /* 1 if all neigen roots met the tolerance */
  int iterations;          /* DavidsonSolver::num_iterations */
  int matvec_columns;      /* operator columns applied in `precision` */
  int matvec_columns_fp64; /* FP64 operator columns applied during refinement */
  int refine_rounds;
  double max_residual;     /* final max residual norm over the neigen roots (FP64 operator when refined) */
  double seconds;          /* wall time of the solve */
} xtpgpu_davidson_info;

/* DavidsonSolver::solve(A, neigen) for the SYMM matrix type (davidsonsolver.h:64-122) on a
 * device BSE operator (TDA). evals: neigen; evecs: size x neigen column-major (host, may be
 * null). Non-converged roots are zeroed like storeNotConvergedData (davidsonsolver.cc:530-547)
 * and XTPGPU_ERR_NOT_CONVERGED is returned. */
int xtpgpu_davidson_solve(xtpgpu_bseop op, int64_t neigen, const xtpgpu_davidson_opts* opts,
                          double* evals, double* evecs, int64_t ldv, xtpgpu_davidson_info* info);

/* Full BSE: DavidsonSolver::solve on HamiltonianOperator<A,B> = [[A, B], [-B, -A]]
 * (bseoperator_btda.h:44-150) with the HAM matrix type (harmonic Ritz, davidsonsolver.cc:241-332),
 * the X-ranked initial guess (bse_initialization.h:47-93) and the X/Y normalisation of
 * BSE::Solve_nonhermitian_Davidson (bse.cc:320-358). A = singlet/triplet TDA operator, B = the
 * matching coupling operator (cqp=0, cx=2/0, cd=0, cd2=1). X, Y: size x neigen (may be null). */
int xtpgpu_davidson_solve_full(xtpgpu_bseop A, xtpgpu_bseop B, int64_t neigen,
                               const xtpgpu_davidson_opts* opts, double* evals, double* X,
                               double* Y, int64_t ldv, xtpgpu_davidson_info* info);
/* The initial-guess overload of DavidsonSolver::solve (davidsonsolver.h:123-152) for the full BSE:
 * the nguess (>= neigen) columns of guess (2*size x nguess, column-major, leading dimension
 * ldg >= 2*size, [X; Y] stacked) are scaled to unit length, orthonormalized and start the search
 * space (the reference's Gram-Schmidt assumes a unit first column, which its ranked guess has;
 * a guess narrower than the size update simply yields fewer correction vectors); the restart size
 * is min(nguess, max search space). Used to continue from converged eigenvectors of a nearby
 * problem (e.g. a perturbed Hqp). */
int xtpgpu_davidson_solve_full_guess(xtpgpu_bseop A, xtpgpu_bseop B, int64_t neigen,
                                     const xtpgpu_davidson_opts* opts, const double* guess, int64_t ldg,
                                     int64_t nguess, double* evals, double* X, double* Y, int64_t ldv,
                                     xtpgpu_davidson_info* info);

/* ---------------------------------------------------------------- matrix-free operators */
/* MatrixFreeOperator (matrixfreeoperator.h:45-98): what DavidsonSolver consumes, a size, a
 * diagonal and matmul. The callback writes Y = A X (size x k, column-major, leading dimensions
 * ldx / ldy) into device memory on `stream` (the context's hipStream_t) for the requested
 * precision (XTPGPU_FP64, or a mixed form the operator declared in its precision mask) and
 * returns 0, or a non-zero code that aborts the solve with that code (XTPGPU_ERR_ARG if positive).
 * X and Y are device buffers owned by the solver, valid for the duration of the call. */
typedef int (*xtpgpu_matmul_fn)(void* user, int precision, const double* X, int64_t ldx, int64_t k,
                                double* Y, int64_t ldy, void* stream);
/* A caller's operator: `diagonal` (host, size values) is copied; precision_mask as for
 * xtpgpu_bse_create (bit 1<<XTPGPU_FP64 is implied). */
int xtpgpu_mfop_create(xtpgpu_ctx ctx, int64_t size, const double* diagonal, xtpgpu_matmul_fn fn,
                       void* user, uint32_t precision_mask, xtpgpu_mfop* out);
/* The library's BSE operator seen through the same interface (BSE_OPERATOR derives from
 * MatrixFreeOperator, bse_operator.h:41); the view is valid while `op` lives. */
int xtpgpu_mfop_from_bse(xtpgpu_bseop op, xtpgpu_mfop* out);
/* MatrixFreeOperator::matmul on device buffers. */
int xtpgpu_mfop_matmul_device(xtpgpu_mfop op, int precision, const double* X, int64_t ldx, int64_t k,
                              double* Y, int64_t ldy);
int xtpgpu_mfop_destroy(xtpgpu_mfop op);
/* DavidsonSolver::solve(A, neigen) (SYMM, davidsonsolver.h:64-122) on any matrix-free operator:
 * the same solver, options, ladder and refinement as xtpgpu_davidson_solve. */
int xtpgpu_davidson_solve_mf(xtpgpu_mfop op, int64_t neigen, const xtpgpu_davidson_opts* opts,
                             double* evals, double* evecs, int64_t ldv, xtpgpu_davidson_info* info);
/* DavidsonSolver::solve on HamiltonianOperator<R, C> = [[R, C], [-C, -R]] (bseoperator_btda.h:44-150)
 * with the HAM matrix type: guess == null starts from the unit vectors on the lowest positive
 * diagonal entries (setupInitialEigenvectors, davidsonsolver.cc:180-206; 2*neigen of them), else
 * from the nguess columns of guess (2*size x nguess, leading dimension ldg). evecs: the unit
 * 2*size-long eigenvectors, 2*size x neigen, leading dimension ldv >= 2*size (may be null). */
int xtpgpu_davidson_solve_ham_mf(xtpgpu_mfop R, xtpgpu_mfop C, int64_t neigen,
                                 const xtpgpu_davidson_opts* opts, const double* guess, int64_t ldg,
                                 int64_t nguess, double* evals, double* evecs, int64_t ldv,
                                 xtpgpu_davidson_info* info);
/* Many segments' BSE-TDA problems, each solved as xtpgpu_davidson_solve would with the same
 * options: the per-segment jobs that ParallelXJobCalc hands to its worker threads
 * (parallelxjobcalc.cc:52-111), here `workers` host threads with a HIP stream, a Davidson
 * workspace and scratch_bytes of operator-product scratch each (0: the largest segment's
 * single-chunk need, capped at 2 GiB), taking segments in `order`: largest operator first (the
 * shortest overall time), smallest first (the earliest completions) or as given. evals: nops x
 * neigen (segment s at evals + s*neigen); evecs (nullable): nops host pointers, each size_s x
 * neigen column-major (entries may be null); infos, seconds, status (nullable, nops each): the
 * segment's solve info, the time from the call's start to the segment's completion, and its
 * return code. Returns the first segment error, else XTPGPU_ERR_NOT_CONVERGED when a segment did
 * not converge (its roots zeroed as in xtpgpu_davidson_solve), else XTPGPU_OK. */
enum { XTPGPU_BATCH_LARGEST_FIRST = 0, XTPGPU_BATCH_SMALLEST_FIRST = 1, XTPGPU_BATCH_AS_GIVEN = 2 };
/* OR-ed into `order`: the workers' operator products meet at a rendezvous and are issued together,
 * one launch per contraction step for the segments posted (half the workers still solving at a
 * time; XTPGPU_BATCH_SUBGROUPS=1 waits for all of them). The Davidson iterations stay per segment;
 * products of streamed-FP64, Hd2 or non-diagonal-Hqp operators run one by one. */
enum { XTPGPU_BATCH_GROUPED = 0x100 };
int xtpgpu_davidson_solve_batched(int64_t nops, const xtpgpu_bseop* ops, int64_t neigen,
                                  const xtpgpu_davidson_opts* opts, int32_t workers, int32_t order,
                                  int64_t scratch_bytes, double* evals, double* const* evecs,
                                  xtpgpu_davidson_info* infos, double* seconds, int32_t* status);

/* ---------------------------------------------------------------- batched segments */
/* eps(i w) for many segments of uneven size (the per-segment calculate_epsilon_i of
 * rpa.cc:75-140 that the reference runs job by job, parallelxjobcalc.cc:52-111) as one
 * persistent launch over (segment, frequency, lower-triangle tile) work items. create packs the
 * segments' transition blocks and weights; run launches (optionally as a captured hipGraph replay);
 * result copies segment seg's eps (P x P x n_omega, column-major) to the host. */
typedef struct xtpgpu_epsbatch_s* xtpgpu_epsbatch;
int xtpgpu_epsbatch_create(xtpgpu_ctx ctx, int nseg, const xtpgpu_mmn* tensors,
                           const double* const* energies, const int64_t* homo, const int64_t* rpamin,
                           const int64_t* rpamax, int n_omega, const double* omegas, int precision,
                           int grid, xtpgpu_epsbatch* out);
int xtpgpu_epsbatch_run(xtpgpu_epsbatch b, int use_graph);
int xtpgpu_epsbatch_set_grid(xtpgpu_epsbatch b, int grid);
/* Time the batch at 1, 2, 4 and 8 workgroups per CU and keep the fastest grid (grid_out, nullable). */
int xtpgpu_epsbatch_autotune(xtpgpu_epsbatch b, int* grid_out);
int xtpgpu_epsbatch_result(xtpgpu_epsbatch b, int seg, double* out);
int xtpgpu_epsbatch_destroy(xtpgpu_epsbatch b);

/* ---------------------------------------------------------------- adjoint */
/* Hellmann-Feynman derivatives of converged BSE roots at fixed screening, the contraction
 * pattern of BSE::ExpectationValue_Operator (bse.cc:490-606): dOmega = X^T dA X + Y^T dA Y +
 * 2 X^T dB Y. B and Y null for TDA. X, Y: size x nroots (host, leading dim ldv), normalised as
 * the solvers return them. d_eps (nullable): (vtotal+ctotal) x nroots, dOmega/de for the BSE
 * levels (occupied first); d_m (nullable): auxsize x vtotal x ctotal x nroots, dOmega/dM^P_{vc}
 * (element (P,v,c) of root r at r*P*N + P*N + v*ctotal + c) for the symmetric transition block
 * M^P_{vc} = M^P_{cv}. No reference counterpart computes these; the reference obtains
 * excited-state derivatives by finite differences (xtp/src/pyxtp/pyxtp/xtp.py:592-682). */
int xtpgpu_bse_adjoint(xtpgpu_bseop A, xtpgpu_bseop B, int64_t nroots, const double* X,
                       const double* Y, int64_t ldv, double* d_eps, double* d_m);
/* Same with d_m a device buffer (stays resident for device-side consumers such as autograd). */
int xtpgpu_bse_adjoint_device(xtpgpu_bseop A, xtpgpu_bseop B, int64_t nroots, const double* X,
                              const double* Y, int64_t ldv, double* d_eps, double* d_m_dev);
/* The derivatives of sum_r weights[r] Omega_r (a backward pass's vector-Jacobian product): d_eps
 * (nullable, host) vtotal+ctotal values, d_m_dev (nullable, device) auxsize x vtotal x ctotal; one
 * root's derivative at a time, so the device holds two P x N arrays instead of nroots. */
int xtpgpu_bse_adjoint_weighted_device(xtpgpu_bseop A, xtpgpu_bseop B, int64_t nroots, const double* X,
                                       const double* Y, int64_t ldv, const double* weights,
                                       double* d_eps, double* d_m_dev);

#ifdef __cplusplus
}
#endif
#endif /* XTPGPU_H */
