"""The many-segment C++ surface of include/xtpgpu.hpp on the device: xtpgpu::EpsBatch (eps(i w) of
every segment in one hipGraph-replayed launch) against each segment's RPA::calculate_epsilon_i, and
DavidsonSolver::solve_batched against each segment's own solve and against dense eigh of the
oracle's BSE matrix (the per-segment jobs of ParallelXJobCalc, parallelxjobcalc.cc:52-111)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import xtp_oracle as X
from votca_amd import pyxtp as px

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CPP_BATCH_DRIVER = r"""
#include "xtpgpu.hpp"
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
int main(int argc, char** argv) {
  std::ifstream in(argv[1], std::ios::binary);
  int64_t nseg = 0, nfreq = 0;
  in.read((char*)&nseg, 8);
  in.read((char*)&nfreq, 8);
  std::vector<double> freqs(nfreq);
  in.read((char*)freqs.data(), nfreq * 8);
  xtpgpu::Context ctx(0);
  std::vector<std::unique_ptr<xtpgpu::TCMatrix>> ts;
  std::vector<std::unique_ptr<xtpgpu::RPA>> rpas;
  std::vector<std::unique_ptr<xtpgpu::SingletOperator_TDA>> ops;
  std::vector<const xtpgpu::RPA*> rp;
  std::vector<xtpgpu_bseop> hs;
  std::vector<std::vector<double>> hqps;
  for (int64_t s = 0; s < nseg; ++s) {
    int64_t o, v, P;
    in.read((char*)&o, 8); in.read((char*)&v, 8); in.read((char*)&P, 8);
    const int64_t n = o + v;
    std::vector<double> e(n), M((size_t)n * n * P);
    in.read((char*)e.data(), n * 8);
    in.read((char*)M.data(), M.size() * 8);  // slice m: n x P column-major
    ts.emplace_back(new xtpgpu::TCMatrix(ctx, P, 0, n - 1, 0, n - 1));
    for (int64_t m = 0; m < n; ++m) ts.back()->set_slice(m, M.data() + (size_t)m * n * P, n);
    rpas.emplace_back(new xtpgpu::RPA(ctx, *ts.back()));
    rpas.back()->configure(o - 1, 0, n - 1);
    rpas.back()->setRPAInputEnergies(e);
    rp.push_back(rpas.back().get());
    std::vector<double> H((size_t)n * n, 0.0), eps0inv(P, 1.0);
    for (int64_t i = 0; i < n; ++i) H[i * n + i] = e[i];
    hqps.push_back(H);
    xtpgpu::BSEOperatorOptions opt{o - 1, 0, 0, 0, n - 1};
    ops.emplace_back(new xtpgpu::SingletOperator_TDA(ctx, *ts.back(), eps0inv.data(), hqps.back().data(), n, opt));
    hs.push_back(ops.back()->get());
  }
  // eps: the batch against each segment's own build, max relative difference
  xtpgpu::EpsBatch batch(ctx, rp, freqs);
  batch.autotune();
  batch.run(true);
  batch.run(true);
  ctx.release_scratch();  // the per-segment builds below allocate the kept scratch again
  double eps_rel = 0.0;
  for (int64_t s = 0; s < nseg; ++s) {
    const std::vector<double> b = batch.result((int)s);
    const int64_t P = rp[s]->auxsize();
    for (int64_t w = 0; w < nfreq; ++w) {
      const std::vector<double> r = rp[s]->calculate_epsilon_i(freqs[w]);
      double num = 0.0, den = 0.0;
      for (int64_t i = 0; i < P * P; ++i) {
        const double d = b[w * P * P + i] - r[i];
        num += d * d;
        den += r[i] * r[i];
      }
      eps_rel = std::max(eps_rel, std::sqrt(num / den));
    }
  }
  std::printf("%.6e\n", eps_rel);
  // TDA: the batched solve against each segment's own solve; the batched eigenvalues printed
  xtpgpu::DavidsonSolver ds;
  ds.set_tolerance(1e-9);
  ds.set_iter_max(200);
  ds.set_max_search_space(60);
  std::vector<std::vector<double>> ev;
  std::vector<double> secs;
  if (!ds.solve_batched(hs, 4, 3, ev, &secs)) return 3;
  double dmax = 0.0;
  for (int64_t s = 0; s < nseg; ++s) {
    if (!ds.solve(*ops[s], 4)) return 4;
    for (int j = 0; j < 4; ++j) dmax = std::max(dmax, std::fabs(ds.eigenvalues()[j] - ev[s][j]));
  }
  std::printf("%.6e\n", dmax);
  for (int64_t s = 0; s < nseg; ++s)
    for (int j = 0; j < 4; ++j) std::printf("%.15e\n", ev[s][j]);
  return 0;
}
"""


def test_cpp_eps_batch_and_batched_davidson(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    rng = np.random.default_rng(5)
    freqs = np.array([0.3, 0.9, 2.5])
    segs = []
    for s in range(5):
        o = int(rng.integers(3, 7))
        v = int(rng.integers(18, 30))
        P = 3 * (o + v)
        n = o + v
        e = X.synthetic_energies(500 + s, o, v)
        M = rng.normal(scale=0.03, size=(n, n, P))
        M = 0.5 * (M + M.transpose(1, 0, 2))  # M^P_{mn} = M^P_{nm}
        segs.append((o, v, P, e, M))
    f = tmp_path / "segs.bin"
    with open(f, "wb") as fh:
        np.array([len(segs), len(freqs)], dtype=np.int64).tofile(fh)
        freqs.tofile(fh)
        for o, v, P, e, M in segs:
            np.array([o, v, P], dtype=np.int64).tofile(fh)
            np.asarray(e, dtype=np.float64).tofile(fh)
            # slice m is n x P column-major: element (n', P) at P*n + n'
            np.ascontiguousarray(np.transpose(M, (0, 2, 1))).tofile(fh)
    src = tmp_path / "batch.cc"
    src.write_text(CPP_BATCH_DRIVER)
    exe = tmp_path / "batch"
    libdir = os.path.dirname(px.LIB_PATH)
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", libdir, "-lxtpgpu", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("[cpp-batch] running the C++ driver", flush=True)
    r = subprocess.run([str(exe), str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    vals = [float(x) for x in r.stdout.split()]
    eps_rel, dev_max = vals[0], vals[1]
    assert eps_rel < 1e-12, eps_rel
    assert dev_max < 1e-10, dev_max
    ev = np.array(vals[2:]).reshape(len(segs), 4)
    for s, (o, v, P, e, M) in enumerate(segs):
        n = o + v
        H = X.bse_dense(1, 2, 1, 0, np.ones(P), M, np.diag(e), o - 1, 0, 0, n - 1)
        w = np.linalg.eigvalsh(H)[:4]
        assert np.abs(ev[s] - w).max() < 1e-8, (s, np.abs(ev[s] - w).max())
