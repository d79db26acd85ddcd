// Quasiparticle-equation solver of GW::SolveQP (xtp/src/libxtp/gwbse/gw.cc:323-410) for any
// batched Sigma_c evaluator. Host code.
//
// Each QP level runs the reference's sequential algorithm: Newton fixed point
// (gw.cc:741-757, newton_rapson.h:49-70) when asked, then the grid search (gw.cc:677-739: restricted
// window first; a full dense window when it finds no accepted root), then linearisation
// (gw.cc:412-431). Grid modes (gw.cc:433-675): adaptive outward shell scan with a 12-step local
// scan per bracket (qp_solver_utils.h:404-708), dense sign-change scan, or adaptive with a dense
// fallback; brackets refined by bisection or Brent (qp_solver_utils.h:199-318), roots accepted on
// residual and Z = 1/(1 - dSigma/dw) (qp_solver_utils.h:320-344) and ranked by
// Z - 0.1 |w - w0| (qp_solver_utils.h:346-348).
//
// Levels run on worker threads; every Sigma request a level makes blocks until all live workers
// have posted theirs, and the batch is then evaluated in one call (one device pass for all
// levels), so the per-level code stays the reference's sequential algorithm.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

namespace xg {
namespace qp {

struct Options {
  bool fixedpoint = false;      // qp_solver "fixedpoint" (else "grid")
  double alpha = 0.75;          // Newton damping (gw.h:70)
  double g_sc_limit = 1e-5;
  int g_sc_max_iterations = 50;
  double half_width = 0.75;     // qp_full_window_half_width
  double dense_spacing = 0.002; // qp_dense_spacing
  double shell_width = 0.025;   // qp_adaptive_shell_width
  int shell_count = 0;          // qp_adaptive_shell_count (> 0 overrides the width)
  bool restrict_search = true;
  double zero_margin = 1e-6;
  double virtual_min = -0.1;
  double min_Z = 0.05, max_Z = 1.5;
  int mode = 0;                 // 0 adaptive_with_dense_fallback, 1 adaptive, 2 dense
  bool brent = false;
  int brent_max_iter = 50;      // the reference passes g_sc_max_iterations (gw.cc:444)
  int gw_sc_iteration = 0;      // the adaptive scan centres on the linearised root only at 0
};

// eval(n, levels, omegas, sigma, dsigma): Sigma_c(level, omega) for n pairs; dsigma (nullable)
// receives dSigma_c/domega.
using BatchEval = std::function<void(long, const long*, const double*, double*, double*)>;

class Batcher {
 public:
  Batcher(BatchEval ev, int workers) : ev_(std::move(ev)), active_(workers) {}

  // Sigma (and, with ds, its derivative) at (level, w); blocks until the batch is served.
  void request(long level, double w, double* s, double* ds) {
    std::unique_lock<std::mutex> lk(mu_);
    if (err_) std::rethrow_exception(err_);
    Req r{level, w, s, ds};
    pend_.push_back(&r);
    const uint64_t g = gen_;
    if ((long)pend_.size() == active_)
      run();
    else
      cv_.wait(lk, [&] { return gen_ != g; });
    if (err_) std::rethrow_exception(err_);
  }

  // a worker has no more levels: serve the others if they are all waiting
  void leave() {
    std::unique_lock<std::mutex> lk(mu_);
    --active_;
    if (active_ > 0 && (long)pend_.size() == active_) run();
  }

 private:
  struct Req {
    long level;
    double w;
    double* s;
    double* ds;
  };
  // called with the lock held and every live worker waiting; with XTPGPU_QP_LOG set, one line per
  // served batch on stderr (a long QP solve then shows its progress)
  void run() {
    const char* lg = std::getenv("XTPGPU_QP_LOG");
    const bool log = lg && *lg && std::strcmp(lg, "0") != 0;
    const auto t0 = std::chrono::steady_clock::now();
    try {
      for (int pass = 0; pass < 2; ++pass) {  // value-only requests, then those with a derivative
        std::vector<long> lv;
        std::vector<double> om;
        std::vector<Req*> rs;
        for (Req* r : pend_)
          if ((r->ds != nullptr) == (pass == 1)) {
            lv.push_back(r->level);
            om.push_back(r->w);
            rs.push_back(r);
          }
        if (rs.empty()) continue;
        std::vector<double> s(rs.size()), ds(pass == 1 ? rs.size() : 0);
        ev_((long)rs.size(), lv.data(), om.data(), s.data(), pass == 1 ? ds.data() : nullptr);
        for (size_t i = 0; i < rs.size(); ++i) {
          *rs[i]->s = s[i];
          if (pass == 1) *rs[i]->ds = ds[i];
        }
      }
    } catch (...) {
      err_ = std::current_exception();
    }
    if (log) {
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::fprintf(stderr, "[qp] batch %llu: %zu Sigma requests, %.2f s\n", (unsigned long long)gen_, pend_.size(), dt);
      std::fflush(stderr);
    }
    pend_.clear();
    ++gen_;
    cv_.notify_all();
  }

  BatchEval ev_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<Req*> pend_;
  long active_;
  uint64_t gen_ = 0;
  std::exception_ptr err_;
};

// QPFunc (gw.h:221-293): value(w) = Sigma_c(w) + intercept - w, deriv(w) = dSigma_c/dw - 1.
struct Func {
  Batcher* b;
  long level;
  double offset;
  double sigma(double w) const {
    double s = 0.0;
    b->request(level, w, &s, nullptr);
    return s;
  }
  double value(double w) const { return sigma(w) + offset - w; }
  double deriv(double w) const {
    double s = 0.0, ds = 0.0;
    b->request(level, w, &s, &ds);
    return ds - 1.0;
  }
};

struct Root {
  double omega = 0.0, residual = 0.0, deriv = 0.0, Z = 0.0, dist = 0.0;
  bool accepted = false;
};

inline double shell_width(const Options& o) {
  return o.shell_count > 0 ? o.half_width / (double)o.shell_count : o.shell_width;
}

inline double bisection(double lo, double flo, double hi, double fhi, const Func& f, const Options& o) {
  if (flo * fhi > 0.0) throw std::runtime_error("Bisection needs a positive and negative function value");
  for (;;) {
    const double c = 0.5 * (lo + hi);
    if (std::fabs(hi - lo) < o.g_sc_limit) return c;
    const double yc = f.value(c);
    if (std::fabs(yc) < o.g_sc_limit) return c;
    if (yc * flo > 0.0) {
      lo = c;
      flo = yc;
    } else {
      hi = c;
      fhi = yc;
    }
  }
}

// Brent's method (bracketing inverse quadratic / secant steps with a bisection safeguard)
inline double brent(double lo, double flo, double hi, double fhi, const Func& f, const Options& o) {
  if (flo * fhi > 0.0) throw std::runtime_error("Brent needs a positive and negative function value");
  double a = lo, b = hi, fa = flo, fb = fhi, c = a, fc = fa, d = b - a, e = d;
  const double tol = o.g_sc_limit;
  for (int it = 0; it < o.brent_max_iter; ++it) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
      c = a;
      fc = fa;
      d = b - a;
      e = d;
    }
    if (std::fabs(fc) < std::fabs(fb)) {  // keep b the best estimate
      const double ob = b, ofb = fb;
      a = ob;
      fa = ofb;
      b = c;
      fb = fc;
      c = ob;
      fc = ofb;
    }
    const double m = 0.5 * (c - b);
    if (std::fabs(m) < tol || std::fabs(fb) < o.g_sc_limit) return b;
    if (std::fabs(e) >= tol && std::fabs(fa) > std::fabs(fb)) {
      const double s = fb / fa;
      double p, q;
      if (a == c) {
        p = 2.0 * m * s;
        q = 1.0 - s;
      } else {
        const double q1 = fa / fc, r = fb / fc;
        p = s * (2.0 * m * q1 * (q1 - r) - (b - a) * (r - 1.0));
        q = (q1 - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (p > 0.0) q = -q;
      p = std::fabs(p);
      if (q != 0.0 && 2.0 * p < std::min(3.0 * m * q - std::fabs(tol * q), std::fabs(e * q))) {
        e = d;
        d = p / q;
      } else {
        d = m;
        e = m;
      }
    } else {
      d = m;
      e = m;
    }
    a = b;
    fa = fb;
    b += std::fabs(d) > tol ? d : (m > 0.0 ? tol : -tol);
    fb = f.value(b);
  }
  throw std::runtime_error("Brent did not converge within qp_bisection_max_iter");
}

inline bool accept(const Root& r, const Options& o) {
  if (!std::isfinite(r.omega) || !std::isfinite(r.Z)) return false;
  if (std::fabs(r.residual) > o.g_sc_limit) return false;
  return r.Z > 0.0 && r.Z >= o.min_Z && r.Z <= o.max_Z;
}

inline double score(const Root& r) { return r.Z - 0.1 * r.dist; }

// the first root of maximal score (std::max_element order)
inline const Root* best_of(const std::vector<Root>& v) {
  const Root* b = nullptr;
  for (const Root& r : v)
    if (!b || score(*b) < score(r)) b = &r;
  return b;
}

inline bool refine(double lo, double flo, double hi, double fhi, const Func& f, double ref, const Options& o,
                   Root& r) {
  const bool l0 = std::fabs(flo) <= o.g_sc_limit, r0 = std::fabs(fhi) <= o.g_sc_limit;
  if (flo * fhi > 0.0) {  // same sign: usable only when an end is already a root
    if (!(l0 || r0)) return false;
    r.omega = std::fabs(flo) <= std::fabs(fhi) ? lo : hi;
  } else {
    r.omega = o.brent ? brent(lo, flo, hi, fhi, f, o) : bisection(lo, flo, hi, fhi, f, o);
  }
  r.residual = f.value(r.omega);
  r.deriv = f.deriv(r.omega);
  r.Z = std::fabs(r.deriv) > 1e-14 ? -1.0 / r.deriv : std::numeric_limits<double>::infinity();
  r.dist = std::fabs(r.omega - ref);
  r.accepted = accept(r, o);
  return true;
}

// the adaptive outward shell scan; returns false when no root was found
inline bool adaptive_core(const Func& f, double w0, double left, double right, const Options& o,
                          std::vector<Root>& acc, std::vector<Root>& rej, double& out) {
  if (left >= right) return false;
  const double sw = shell_width(o);
  double center = w0;
  if (o.gw_sc_iteration == 0) {  // centre on the linearised root when it lies in the window
    const double f0 = f.value(w0), df0 = f.deriv(w0);
    if (std::isfinite(f0) && std::isfinite(df0) && std::fabs(df0) > 1e-6) {
      const double wl = w0 - f0 / df0;
      if (std::isfinite(wl) && wl >= left && wl <= right) center = wl;
    }
  }
  center = std::max(left, std::min(right, center));
  const long n_shells = (long)std::ceil(std::max(center - left, right - center) / sw);

  auto store = [&](double a, double fa, double b, double fb) {
    if (b < a) {
      std::swap(a, b);
      std::swap(fa, fb);
    }
    struct Br {
      double l, fl, r, fr;
    };
    std::vector<Br> brs;
    constexpr int nsub = 12;  // local scan inside the coarse bracket
    if (b > a) {
      const double dx = (b - a) / (double)nsub;
      double xp = a, fp = fa;
      for (int i = 1; i <= nsub; ++i) {
        const double xc = i == nsub ? b : a + (double)i * dx;
        const double fc = i == nsub ? fb : f.value(xc);
        if ((fp < 0.0 && fc > 0.0) || (fp > 0.0 && fc < 0.0)) brs.push_back({xp, fp, xc, fc});
        if (std::fabs(fp) <= o.g_sc_limit && xp < xc) brs.push_back({xp, fp, xc, fc});
        if (std::fabs(fc) <= o.g_sc_limit && xp < xc) brs.push_back({xp, fp, xc, fc});
        xp = xc;
        fp = fc;
      }
    }
    if (!brs.empty()) {  // the sub-bracket whose midpoint is nearest the centre (lower left on ties)
      size_t bi = 0;
      double bd = std::fabs(0.5 * (brs[0].l + brs[0].r) - center);
      for (size_t i = 1; i < brs.size(); ++i) {
        const double dd = std::fabs(0.5 * (brs[i].l + brs[i].r) - center);
        if (dd < bd - 1e-14 || (std::fabs(dd - bd) <= 1e-14 && brs[i].l < brs[bi].l)) {
          bi = i;
          bd = dd;
        }
      }
      a = brs[bi].l;
      fa = brs[bi].fl;
      b = brs[bi].r;
      fb = brs[bi].fr;
    }
    Root r;
    if (!refine(a, fa, b, fb, f, w0, o, r)) return;
    (r.accepted ? acc : rej).push_back(r);
  };

  const double cv = f.value(center);
  double lx = center, lf = cv, rx = center, rf = cv;
  bool la = true, ra = true;
  for (long shell = 1; shell <= n_shells; ++shell) {
    bool added = false;
    const double delta = (double)shell * sw;
    if (la) {
      const double w = center - delta;
      if (w >= left) {
        const double fw = f.value(w);
        added = true;
        if (lf * fw < 0.0) store(w, fw, lx, lf);
        lx = w;
        lf = fw;
      } else {
        la = false;
      }
    }
    if (ra) {
      const double w = center + delta;
      if (w <= right) {
        const double fw = f.value(w);
        added = true;
        if (rf * fw < 0.0) store(rx, rf, w, fw);
        rx = w;
        rf = fw;
      } else {
        ra = false;
      }
    }
    if (!added && !la && !ra) break;
  }
  if (lx > left + 1e-12) {
    const double fe = f.value(left);
    if (fe * lf < 0.0) store(left, fe, lx, lf);
  }
  if (rx < right - 1e-12) {
    const double fe = f.value(right);
    if (rf * fe < 0.0) store(rx, rf, right, fe);
  }
  if (!acc.empty()) {
    out = best_of(acc)->omega;
    return true;
  }
  if (!rej.empty()) {
    out = best_of(rej)->omega;
    return true;
  }
  return false;
}

inline bool windowed_adaptive(const Func& f, double w0, double left, double right, bool allow_rejected,
                              const Options& o, double& out) {
  std::vector<Root> acc, rej;
  const bool found = adaptive_core(f, w0, left, right, o, acc, rej, out);
  if (!acc.empty()) return true;
  if (!rej.empty() && !allow_rejected) return false;
  return found;
}

inline bool windowed_dense(const Func& f, double w0, double left, double right, bool allow_rejected,
                           const Options& o, double& out) {
  std::vector<Root> acc, rej;
  if (left < right) {
    double xp = left, fp = f.value(left);
    const long n = std::max<long>(2, (long)std::ceil((right - left) / o.dense_spacing) + 1);
    for (long i = 1; i < n; ++i) {
      const double x = i == n - 1 ? right : std::min(right, left + (double)i * o.dense_spacing);
      const double v = f.value(x);
      if (fp * v < 0.0) {
        Root r;
        if (refine(xp, fp, x, v, f, w0, o, r)) (r.accepted ? acc : rej).push_back(r);
      }
      xp = x;
      fp = v;
    }
  }
  if (!acc.empty()) {
    out = best_of(acc)->omega;
    return true;
  }
  if (!rej.empty() && allow_rejected) {
    out = best_of(rej)->omega;
    return true;
  }
  return false;
}

inline bool windowed(const Func& f, double w0, double left, double right, bool allow_rejected, const Options& o,
                     double& out) {
  if (o.mode == 1) return windowed_adaptive(f, w0, left, right, allow_rejected, o, out);
  if (o.mode == 2) return windowed_dense(f, w0, left, right, allow_rejected, o, out);
  if (windowed_adaptive(f, w0, left, right, allow_rejected, o, out)) return true;
  return windowed_dense(f, w0, left, right, allow_rejected, o, out);
}

inline bool grid(const Func& f, double w0, bool occupied, const Options& o, double& out) {
  const double fl = w0 - o.half_width, fr = w0 + o.half_width;
  double rl = fl, rr = fr;
  bool restricted = false;
  if (o.restrict_search) {
    if (occupied)
      rr = std::min(fr, -o.zero_margin);
    else
      rl = std::max(fl, o.virtual_min);
    restricted = std::fabs(rl - fl) > 1e-12 || std::fabs(rr - fr) > 1e-12;
  }
  if (restricted && rl < rr) {
    if (windowed(f, w0, rl, rr, false, o, out)) return true;
    return windowed_dense(f, w0, fl, fr, true, o, out);  // the full window, densely
  }
  return windowed(f, w0, fl, fr, true, o, out);
}

inline bool fixed_point(const Func& f, double w0, const Options& o, double& out) {
  double x = w0;
  for (int it = 0; it < o.g_sc_max_iterations; ++it) {
    const double v = f.value(x), dv = f.deriv(x);
    if (std::fabs(dv) < 1e-12) return false;
    const double step = -o.alpha * v / dv;
    if (std::fabs(step) < o.g_sc_limit) {
      out = x;
      return true;
    }
    x += step;
  }
  return false;
}

// GW::SolveQP_Linearisation as the reference evaluates it: Z = 1 - QPFunc::deriv(w0), where
// QPFunc::deriv already carries the -1 (gw.cc:418-423), i.e. a step divided by 2 - dSigma/dw.
inline bool linearisation(const Func& f, double w0, double& out) {
  const double s = f.sigma(w0);
  const double Z = 1.0 - f.deriv(w0);
  if (std::fabs(Z) > 1e-9) {
    out = w0 + (f.offset - w0 + s) / Z;
    return true;
  }
  return false;
}

// GW::SolveQP for qt levels; out / converged per level (an unconverged level keeps its
// linearised or, failing that, its starting frequency).
inline void solve(const BatchEval& ev, long qt, const double* intercepts, const double* freqs, long homo, long qpmin,
                  const Options& o, double* out, int* converged) {
  if (qt <= 0) return;
  const int nw = (int)std::min<long>(qt, 256);
  Batcher batch(ev, nw);
  std::atomic<long> next{0};
  std::mutex emu;
  std::exception_ptr first;
  auto work = [&] {
    try {
      for (;;) {
        const long l = next.fetch_add(1);
        if (l >= qt) break;
        const Func f{&batch, l, intercepts[l]};
        double w = freqs[l];
        int ok = 0;
        double r = 0.0;
        if (o.fixedpoint && fixed_point(f, freqs[l], o, r)) {
          w = r;
          ok = 1;
        } else if (grid(f, freqs[l], l + qpmin <= homo, o, r)) {
          w = r;
          ok = 1;
        } else if (linearisation(f, freqs[l], r)) {
          w = r;
        }
        out[l] = w;
        if (converged) converged[l] = ok;
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(emu);
      if (!first) first = std::current_exception();
    }
    batch.leave();
  };
  std::vector<std::thread> th;
  th.reserve(nw);
  for (int i = 0; i < nw; ++i) th.emplace_back(work);
  for (auto& t : th) t.join();
  if (first) std::rethrow_exception(first);
}

}  // namespace qp
}  // namespace xg
