"""GPU parity: the HIP path through the C-ABI against the oracle and the reference's
golden fixtures. FP64 must agree with the oracle to round-off; the mixed forms within the
tolerances written in each test."""
import os

import numpy as np
import pytest

from oracle import integrals as I
from oracle import mmio
from oracle import xtp_oracle as X
from votca_amd import pyxtp as px

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS_INV_BSE_OP = np.array([0.999807798016267, 0.994206065211371, 0.917916768047073,
    0.902913813951883, 0.902913745974602, 0.902913584797742, 0.853352878674581, 0.853352727016914,
    0.853352541699637, 0.79703468058566, 0.797034577207669, 0.797034400395582, 0.787701833916331,
    0.518976361745313, 0.518975064844033, 0.518973712898761, 0.459286057710524])


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    return px.default_context()


@pytest.fixture(scope="module")
def ints():
    return I.ao_integrals(os.path.join(G, "threecenter_gwbse/molecule.xyz"),
                          os.path.join(G, "threecenter_gwbse/3-21G.xml"))


def dev_mmn(ctx, M, mmin=0, nmin=0):
    t = px.TCMatrix_gwbse(ctx).Initialize(M.shape[2], mmin, mmin + M.shape[0] - 1, nmin, nmin + M.shape[1] - 1)
    return t.set_tensor(M)


def test_mmn_roundtrip(ctx):
    rng = np.random.default_rng(0)
    M = rng.normal(size=(3, 5, 7))
    t = dev_mmn(ctx, M)
    assert np.array_equal(t.get_tensor(), M)


def test_synthetic_generator_bitexact(ctx):
    t = px.TCMatrix_gwbse(ctx).Initialize(13, 2, 9, 1, 11)
    t.fill_synthetic(42, 0.3, 3.0)
    ref = X.synthetic_mmn(42, 2, 9, 1, 11, 13)
    assert np.array_equal(t.get_tensor(), ref)


def test_multiply_right(ctx):
    rng = np.random.default_rng(1)
    M = rng.normal(size=(6, 9, 70))
    A = rng.normal(size=(70, 70))
    t = dev_mmn(ctx, M)
    t.MultiplyRightWithAuxMatrix(A)
    assert relerr(t.get_tensor(), X.multiply_right(M, A)) < 1e-13


def test_pseudo_invsqrt_gwbse_in_the_library(ctx, ints):  # aomatrix.cc:53-86, test_threecenter_gwbse.cc:36-126
    """AOCoulomb::Pseudo_InvSqrt_GWBSE on the device: with no function removed its result R obeys
    R^T V R = 1 and R R^T = V^-1 (closed forms of ((S^-1/2 V S^-1/2)^-1/2 S^-1/2)^T); it equals the
    oracle's restatement; TCMatrix_gwbse::Fill through it reproduces the reference's fixtures; an
    overlap with an eigenvalue below etol has that function removed and counted."""
    S, Saux, V, T = ints
    R, removed = px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)
    assert removed == 0
    n = V.shape[0]
    assert np.abs(R.T @ V @ R - np.eye(n)).max() < 1e-10
    assert relerr(R @ R.T, np.linalg.inv(V)) < 1e-10
    assert relerr(R, X.pseudo_invsqrt_gwbse(V, Saux, 5e-7)) < 1e-12
    C = mmio.read_matrix(os.path.join(G, "threecenter_gwbse/MOs.mm"))
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 5, 0, 7)
    t.Fill(T, C, R)
    M = t.get_tensor()
    for k in (0, 2, 4):
        assert mmio.isapprox(M[k], mmio.read_matrix(os.path.join(G, f"threecenter_gwbse/ref{k}b.mm")), 1e-5)
    lam, U = np.linalg.eigh(Saux)
    lam[0] = 1e-9  # one aux combination (nearly) linearly dependent
    S2 = (U * lam) @ U.T
    R2, removed2 = px.Pseudo_InvSqrt_GWBSE(V, S2, 5e-7, ctx)
    assert removed2 >= 1
    assert relerr(R2, X.pseudo_invsqrt_gwbse(V, S2, 5e-7)) < 1e-10


def test_fill_ao2mo_matches_threecenter_fixture(ctx, ints):  # test_threecenter_gwbse.cc:36-126
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, "threecenter_gwbse/MOs.mm"))
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 5, 0, 7)
    t.Fill(T, C, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    M = t.get_tensor()
    assert relerr(M, X.fill_mmn(T, C, 0, 5, 0, 7, Saux, V)) < 1e-12
    for k in (0, 2, 4):
        assert mmio.isapprox(M[k], mmio.read_matrix(os.path.join(G, f"threecenter_gwbse/ref{k}b.mm")), 1e-5)


def test_rpa_fixtures_gpu(ctx, ints):  # test_rpa.cc:69-140
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, "rpa/eigenvectors.mm"))
    e = mmio.read_vector(os.path.join(G, "rpa/eigenvals.mm"))
    M = X.fill_mmn(T, C, 0, 16, 0, 16, Saux, V)
    rpa = px.RPA(dev_mmn(ctx, M))
    rpa.configure(4, 0, 16)
    rpa.setRPAInputEnergies(e)
    for name, f, kind in (("i_ref", 0.5, "imag"), ("r_ref", 0.0, "real"), ("r_complex_ref", 0.5 + 0.5j, "complex")):
        g = rpa.calculate_epsilon_i(f) if kind == "imag" else rpa.calculate_epsilon_r(f)
        assert mmio.isapprox(g, mmio.read_matrix(os.path.join(G, f"rpa/{name}.mm")), 1e-4), name
        assert relerr(g, X.rpa_epsilon(M, e, 4, 0, 16, f, kind)) < 1e-13, name


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("bf16x3", 2e-5), ("bf16", 2e-2)])
def test_rpa_random_multi_frequency(ctx, prec, tol):
    rng = np.random.default_rng(2)
    o, v, P = 9, 70, 150
    M = rng.normal(scale=0.3, size=(o + v, o + v, P))
    e = np.concatenate([np.sort(rng.uniform(-1, -0.2, o)), np.sort(rng.uniform(0, 1.5, v))])
    rpa = px.RPA(dev_mmn(ctx, M))
    rpa.configure(o - 1, 0, o + v - 1)
    rpa.setRPAInputEnergies(e)
    freqs = [0.0, 0.3, 1.7, 5.0]
    g = rpa.calculate_epsilon(px.EPS_IMAG, freqs, prec)
    for i, f in enumerate(freqs):
        ref = X.rpa_epsilon(M, e, o - 1, 0, o + v - 1, f, "imag")
        assert relerr(g[:, :, i], ref) < tol, (prec, f, relerr(g[:, :, i], ref))
        assert np.array_equal(g[:, :, i], g[:, :, i].T)


@pytest.mark.parametrize("prec,tol", [("bf16x3", 2e-5), ("bf16", 2e-2)])
@pytest.mark.parametrize("kind,freqs", [("imag", [0.0, 0.3, 1.7]), ("real", [0.0, 0.05, 0.1])])
def test_rpa_mixed_frequency_groups(ctx, monkeypatch, prec, tol, kind, freqs):
    """The mixed eps builds with the frequency weights folded into scaled transition copies, one
    frequency per group (XTPGPU_EPS_SCALE_BYTES=1) and all in one group, against the oracle;
    the real-axis weights change sign across the transitions."""
    rng = np.random.default_rng(5)
    o, v, P = 8, 60, 130
    M = rng.normal(scale=0.3, size=(o + v, o + v, P))
    e = np.concatenate([np.sort(rng.uniform(-1, -0.2, o)), np.sort(rng.uniform(0.3, 1.5, v))])
    rpa = px.RPA(dev_mmn(ctx, M))
    rpa.configure(o - 1, 0, o + v - 1)
    rpa.setRPAInputEnergies(e)
    k = px.EPS_IMAG if kind == "imag" else px.EPS_REAL
    outs = []
    for budget in ("1", None):
        if budget:
            monkeypatch.setenv("XTPGPU_EPS_SCALE_BYTES", budget)
        else:
            monkeypatch.delenv("XTPGPU_EPS_SCALE_BYTES", raising=False)
        g = rpa.calculate_epsilon(k, freqs, prec)
        for i, f in enumerate(freqs):
            ref = X.rpa_epsilon(M, e, o - 1, 0, o + v - 1, f, kind)
            assert relerr(g[:, :, i], ref) < tol, (prec, kind, f, budget)
            assert np.array_equal(g[:, :, i], g[:, :, i].T)
        outs.append(g)
    assert np.array_equal(outs[0], outs[1])  # grouping does not change a single bit


@pytest.mark.parametrize("budget", ["1", None])
def test_rpa_fp64_scaled_copies(ctx, monkeypatch, budget):
    """FP64 eps with the weights in scaled copies (the default), in groups of one frequency and in
    one group, against the oracle (1e-13) and the load-weighted FP64 kernel (XTPGPU_EPS64_SCALED=0)."""
    rng = np.random.default_rng(6)
    o, v, P = 8, 60, 130
    M = rng.normal(scale=0.3, size=(o + v, o + v, P))
    e = np.concatenate([np.sort(rng.uniform(-1, -0.2, o)), np.sort(rng.uniform(0.3, 1.5, v))])
    rpa = px.RPA(dev_mmn(ctx, M))
    rpa.configure(o - 1, 0, o + v - 1)
    rpa.setRPAInputEnergies(e)
    freqs = [0.0, 0.3, 1.7]
    monkeypatch.setenv("XTPGPU_EPS64_SCALED", "0")
    g0 = rpa.calculate_epsilon(px.EPS_IMAG, freqs, "fp64")
    monkeypatch.setenv("XTPGPU_EPS64_SCALED", "1")
    if budget:
        monkeypatch.setenv("XTPGPU_EPS_SCALE_BYTES", budget)
    g = rpa.calculate_epsilon(px.EPS_IMAG, freqs, "fp64")
    for i, f in enumerate(freqs):
        assert relerr(g[:, :, i], X.rpa_epsilon(M, e, o - 1, 0, o + v - 1, f, "imag")) < 1e-13
        assert relerr(g[:, :, i], g0[:, :, i]) < 1e-14
        assert np.array_equal(g[:, :, i], g[:, :, i].T)


def test_release_scratch_then_rebuild(ctx):
    """The context's kept eps scratch freed between builds: the builds after it allocate it again
    and give the same bits as before."""
    rng = np.random.default_rng(8)
    o, v, P = 6, 50, 120
    M = rng.normal(scale=0.3, size=(o + v, o + v, P))
    e = np.concatenate([np.sort(rng.uniform(-1, -0.2, o)), np.sort(rng.uniform(0.3, 1.5, v))])
    rpa = px.RPA(dev_mmn(ctx, M))
    rpa.configure(o - 1, 0, o + v - 1)
    rpa.setRPAInputEnergies(e)
    freqs = [0.0, 0.7]
    before = {p: rpa.calculate_epsilon(px.EPS_IMAG, freqs, p) for p in ("fp64", "bf16x3")}
    for _ in range(2):
        ctx.release_scratch()
        for p, g in before.items():
            assert np.array_equal(rpa.calculate_epsilon(px.EPS_IMAG, freqs, p), g), p


@pytest.mark.parametrize("c", [(1, 2, 1, 0), (1, 0, 1, 0), (0, 2, 0, 0), (0, 0, 1, 0)])
@pytest.mark.parametrize("chunks", [False, True])
def test_grouped_products_match_single(ctx, c, chunks, monkeypatch):
    """The grouped launches of the many-segment solve on one operator (xtpgpu_bse_matmul with
    XTPGPU_MATMUL_GROUPED=1) against the single-operator product and, in FP64, the oracle's dense
    matrix (bse_operator.cc:39-119), with one stage-1 aux chunk and with several (uneven last)."""
    rng = np.random.default_rng(12)
    o, v, P = 9, 70, 150
    n = o + v
    t = px.TCMatrix_gwbse(ctx).Initialize(P, 0, n - 1, 0, n - 1)
    t.fill_synthetic(12, 0.05, 3.0)
    e = X.synthetic_energies(12, o, v)
    eps = rng.uniform(0.3, 1.0, P)
    opt = px.BSEOperator_Options(o - 1, 0, 0, 0, n - 1)
    op = px.BSE_OPERATOR(*c, eps, t, np.diag(e), precisions=("fp64", "bf16", "bf16x3")).configure(opt)
    Xm = rng.normal(size=(op.size, 7))
    H = X.bse_dense(*c, eps, t.get_tensor(), np.diag(e), o - 1, 0, 0, n - 1)
    if chunks:  # FP64 stage-1 block per aux function: 70 x 7 x 16 x 8 bytes; 41 per chunk
        monkeypatch.setenv("XTPGPU_HD_CHUNK_BYTES", str(41 * 70 * 7 * 16 * 8))
    for prec in ("fp64", "bf16x3", "bf16"):
        monkeypatch.delenv("XTPGPU_MATMUL_GROUPED", raising=False)
        y1 = op.matmul(Xm, prec)
        monkeypatch.setenv("XTPGPU_MATMUL_GROUPED", "1")
        yg = op.matmul(Xm, prec)
        scale = np.abs(y1).max()
        if prec == "fp64":
            assert np.abs(yg - y1).max() <= 1e-13 * scale, (c, chunks)
            assert np.abs(yg - H @ Xm).max() <= 1e-12 * scale, (c, chunks)
        else:  # the same BF16 operands; only the FP32 accumulation of Hd stage 2 is split differently
            assert np.abs(yg - y1).max() <= 5e-5 * scale, (c, chunks, prec, np.abs(yg - y1).max() / scale)
    if c[0]:  # a non-diagonal Hqp is not grouped: refused here, run one by one in the batched solve
        Hq = np.diag(e) + 1e-3 * np.ones((n, n))
        opn = px.BSE_OPERATOR(*c, eps, t, Hq, precisions=("fp64",)).configure(opt)
        with pytest.raises(px.XtpGpuError, match="not grouped"):
            opn.matmul(Xm)
        opn.close()
    op.close()
    t.close()


@pytest.mark.parametrize("kern", ["staged32", "staged64"])
def test_rpa_bf16x3_kernel_variants(ctx, kern, monkeypatch):
    """eps(i w) through both register-staged BF16 kernels (K step 32 and 64, on the scaled
    transition copy) against the oracle."""
    rng = np.random.default_rng(3)
    o, v, P = 11, 97, 210
    M = rng.normal(scale=0.3, size=(o + v, o + v, P))
    e = np.concatenate([np.sort(rng.uniform(-1, -0.2, o)), np.sort(rng.uniform(0, 1.5, v))])
    rpa = px.RPA(dev_mmn(ctx, M))
    rpa.configure(o - 1, 0, o + v - 1)
    rpa.setRPAInputEnergies(e)
    freqs = [0.0, 0.9]
    monkeypatch.setenv("XTPGPU_BF16_KERNEL", kern)
    g = rpa.calculate_epsilon(px.EPS_IMAG, freqs, "bf16x3")
    for i, f in enumerate(freqs):
        assert relerr(g[:, :, i], X.rpa_epsilon(M, e, o - 1, 0, o + v - 1, f, "imag")) < 2e-5


@pytest.mark.parametrize("name,c", [("hqp", (1, 0, 0, 0)), ("hx", (0, 1, 0, 0)), ("hd", (0, 0, 1, 0)),
                                    ("hd2", (0, 0, 0, 1))])
def test_bse_operator_fixtures_gpu(ctx, ints, name, c):  # test_bse_operator.cc:37-151
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, "bse_operator/MOs.mm"))
    Hqp = mmio.read_matrix(os.path.join(G, "bse_operator/Hqp.mm"))
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 16, 0, 16)
    t.Fill(T, C, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    t.MultiplyRightWithAuxMatrix(mmio.read_matrix(os.path.join(G, "bse_operator/rpa_op.mm")))
    op = px.BSE_OPERATOR(*c, EPS_INV_BSE_OP, t, Hqp).configure(px.BSEOperator_Options(4, 0, 0, 0, 8))
    H = op.matmul(np.eye(op.size))
    assert mmio.isapprox(H, mmio.read_matrix(os.path.join(G, f"bse_operator/{name}_ref.mm")), 1e-3)
    assert mmio.isapprox(np.diag(H), op.diagonal(), 1e-3)
    ref = X.BSEOperatorOracle(*c, EPS_INV_BSE_OP, t.get_tensor(), Hqp, 4, 0, 0, 0, 8).matmul(np.eye(op.size))
    assert np.abs(H - ref).max() < 1e-12


BSE_CASES = [  # test_bse.cc:40-330: (fixture, psi fixture, operator, use Hqp off-diagonals, roots, full BSE)
    ("singlets_nooffdiag_tda", "singlets_psi_nooffdiag_tda", (1, 2, 1, 0), False, 3, False),
    ("singlets_tda", "singlets_psi_tda", (1, 2, 1, 0), True, 3, False),
    ("triplets_tda", None, (1, 0, 1, 0), True, 1, False),
    ("singlets_btda", "singlets_psi_btda", (1, 2, 1, 0), True, 3, True),
]


@pytest.mark.parametrize("name,psi,c,offdiag,nroots,full", BSE_CASES, ids=[x[0] for x in BSE_CASES])
def test_bse_energy_fixtures_gpu(ctx, ints, name, psi, c, offdiag, nroots, full):
    """The reference's BSE test chain on the device: Fill3cMO with the 3-21G basis as its own
    auxiliary basis, SetupDirectInteractionOperator at the RPA input energies diag(Hqp), then the
    Davidson with the test's options ('lapack' tolerance 1e-9, DPR, 'safe', 50 iterations, search
    space 10 * nmax): energies isApprox 1e-3, coefficient projections 1 within 1e-5."""
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, "bse/MOs.mm"))
    Hqp = mmio.read_matrix(os.path.join(G, "bse/Hqp.mm"))
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 16, 0, 16)
    t.Fill(T, C, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    epsinv = px.setup_screening(t, np.diag(Hqp).copy(), 4, 0, 16)
    Hq = Hqp if offdiag else np.diag(np.diag(Hqp))
    opt = px.BSEOperator_Options(4, 0, 0, 0, 16)
    ds = px.DavidsonSolver()
    ds.set_tolerance("lapack")
    ds.set_correction("DPR")
    ds.set_size_update("safe")
    ds.set_iter_max(50)
    ds.set_max_search_space(10 * nroots)
    A = px.BSE_OPERATOR(*c, epsinv, t, Hq).configure(opt)
    if full:
        B = px.BSE_OPERATOR(0, c[1], 0, 1, epsinv, t, Hq).configure(opt)
        ev, Xv, Yv = ds.solve_full(A, B, nroots)
    else:
        ev, Xv = ds.solve(A, nroots)
    assert ds.converged(), ds.info.max_residual
    assert mmio.isapprox(ev, mmio.read_matrix(os.path.join(G, f"bse/{name}.mm")).ravel(), 1e-3)
    unit = lambda a: a / np.linalg.norm(a, axis=0)  # noqa: E731
    if psi:
        R = unit(mmio.read_matrix(os.path.join(G, f"bse/{psi}.mm")))
        assert np.abs(np.linalg.norm(R.T @ unit(Xv), axis=0) - 1).max() < 1e-5
    if full:
        RA = unit(mmio.read_matrix(os.path.join(G, "bse/singlets_psi_AR_btda.mm")))
        assert np.abs(np.linalg.norm(RA.T @ unit(Yv), axis=0) - 1).max() < 1e-5




def test_bse_adjust_hqp_fixtures_gpu(ctx, ints):  # test_bse.cc:342-374
    """BSE::configure's AdjustHqpSize through the library: the cut (vmin=1, cmax=15) and the
    extension (GW window 1..15 from the cut Hqp, BSE window 0..16) against the reference's
    fixtures, then the singlet TDA on the cut window against the dense oracle of that window."""
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, "bse/MOs.mm"))
    Hqp = mmio.read_matrix(os.path.join(G, "bse/Hqp.mm"))
    rpa_e = np.diag(Hqp).copy()
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 16, 0, 16)
    t.Fill(T, C, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    bse = px.BSE(t).configure(px.BSE_options(4, 0, 16, 0, 16, 1, 15), rpa_e, Hqp)
    ref_cut = mmio.read_matrix(os.path.join(G, "bse/Hqp_cut.mm"))
    assert mmio.isapprox(bse.getHqp(), ref_cut, 1e-3)
    assert np.array_equal(bse.getHqp(), X.adjust_hqp_size(Hqp, rpa_e, 4, 0, 0, 16, 1, 15))
    bse2 = px.BSE(t).configure(px.BSE_options(4, 0, 16, 1, 15, 0, 16), rpa_e, ref_cut)
    assert mmio.isapprox(bse2.getHqp(), mmio.read_matrix(os.path.join(G, "bse/Hqp_extended.mm")), 1e-3)
    with pytest.raises(px.XtpGpuError):
        px.BSE(t).configure(px.BSE_options(4, 0, 16, 0, 16, 1, 17), rpa_e, Hqp)  # cmax past the RPA levels
    epsinv = bse.SetupDirectInteractionOperator()
    A = px.SingletOperator_TDA(epsinv, t, bse.getHqp()).configure(bse.operator_options())
    ds = px.DavidsonSolver()
    ds.set_tolerance("lapack")
    ds.set_max_search_space(30)
    ev, _ = ds.solve(A, 3)
    dense = X.bse_dense(1, 2, 1, 0, epsinv, t.get_tensor(), bse.getHqp(), 4, 0, 1, 15)
    assert np.abs(ev - np.linalg.eigvalsh(dense)[:3]).max() < 1e-9


def test_full_bse_initial_guess_gpu(ctx, ints):
    """The initial-guess overload (davidsonsolver.h:123-152) on the full BSE of the reference's test:
    started from converged [X; Y] it converges at once to the same roots; a guess with fewer
    columns than roots or the wrong row count is refused with the reference's messages."""
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, "bse/MOs.mm"))
    Hqp = mmio.read_matrix(os.path.join(G, "bse/Hqp.mm"))
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 16, 0, 16)
    t.Fill(T, C, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    epsinv = px.setup_screening(t, np.diag(Hqp).copy(), 4, 0, 16)
    opt = px.BSEOperator_Options(4, 0, 0, 0, 16)
    A = px.BSE_OPERATOR(1, 2, 1, 0, epsinv, t, Hqp).configure(opt)
    B = px.BSE_OPERATOR(0, 2, 0, 1, epsinv, t, Hqp).configure(opt)
    ds = px.DavidsonSolver()
    ds.set_tolerance("lapack")
    ds.set_max_search_space(30)
    ev, Xv, Yv = ds.solve_full(A, B, 3)
    assert ds.converged()
    ev2, X2, Y2 = ds.solve_full(A, B, 3, initial_guess=np.vstack([Xv, Yv]))
    assert ds.converged() and ds.num_iterations() <= 1
    assert np.abs(ev2 - ev).max() < 1e-12
    # the three roots are degenerate (T2): compare the spanned subspaces
    Z1, Z2 = np.vstack([Xv, Yv]), np.vstack([X2, Y2])
    assert np.abs(Z1 @ np.linalg.pinv(Z1) - Z2 @ np.linalg.pinv(Z2)).max() < 1e-7
    with pytest.raises(px.XtpGpuError, match="fewer columns"):
        ds.solve_full(A, B, 3, initial_guess=np.vstack([Xv, Yv])[:, :2])
    with pytest.raises(px.XtpGpuError, match="wrong number of rows"):
        ds.solve_full(A, B, 3, initial_guess=Xv)

DYN_CASES = [  # test_bse.cc:177-191, 268-282, 326-340: (fixture, operator, roots, full BSE, isApprox tolerance)
    ("singlets_dynamic_TDA", (1, 2, 1, 0), 3, False, 0.005),
    ("triplets_dynamic_TDA", (1, 0, 1, 0), 1, False, 0.001),
    ("singlets_dynamic_full", (1, 2, 1, 0), 3, True, 0.05),
]


@pytest.mark.parametrize("name,c,nroots,full,tol", DYN_CASES, ids=[x[0] for x in DYN_CASES])
def test_dynamical_screening_fixtures_gpu(ctx, ints, name, c, nroots, full, tol):
    """BSE::Perturbative_DynamicalScreening on the device (screening W(Omega) and the Hd/Hd2
    operators), fed by the device Davidson's eigenpairs of the reference's test: the fixture at the
    test's isApprox tolerance, and the oracle on the same eigenvectors to 1e-9."""
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, "bse/MOs.mm"))
    Hqp = mmio.read_matrix(os.path.join(G, "bse/Hqp.mm"))
    e = np.diag(Hqp).copy()
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 16, 0, 16)
    t.Fill(T, C, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    epsinv = px.setup_screening(t, e, 4, 0, 16)
    opt = px.BSEOperator_Options(4, 0, 0, 0, 16)
    ds = px.DavidsonSolver()
    ds.set_tolerance("lapack")
    ds.set_correction("DPR")
    ds.set_size_update("safe")
    ds.set_iter_max(50)
    ds.set_max_search_space(10 * nroots)
    A = px.BSE_OPERATOR(*c, epsinv, t, Hqp).configure(opt)
    if full:
        B = px.BSE_OPERATOR(0, c[1], 0, 1, epsinv, t, Hqp).configure(opt)
        ev, Xv, Yv = ds.solve_full(A, B, nroots)
        Xv = Xv / np.linalg.norm(Xv, axis=0)
        Yv = Yv / np.linalg.norm(Yv, axis=0)
    else:
        ev, Xv = ds.solve(A, nroots)
        Yv = None
    assert ds.converged(), ds.info.max_residual
    dyn = px.Perturbative_DynamicalScreening(t, e, Hqp, opt, 16, ev, Xv, Yv)
    assert mmio.isapprox(dyn, mmio.read_matrix(os.path.join(G, f"bse/{name}.mm")).ravel(), tol)
    M = X.fill_mmn(T, C, 0, 16, 0, 16, Saux, V)
    ref = X.perturbative_dynamical_screening(M, e, 4, 0, 16, Hqp, 0, 16, ev, Xv, Yv)
    assert np.abs(dyn - ref).max() < 1e-9

def random_bse_case(seed, vt=7, ct=33, P=70, rpamin=0, vmin_off=0, offdiag=False, scale=0.3):
    rng = np.random.default_rng(seed)
    ntot = vmin_off + vt + ct
    M = rng.normal(scale=scale, size=(ntot, ntot, P))
    M = 0.5 * (M + M.transpose(1, 0, 2))
    eps = rng.uniform(0.1, 1.0, P)
    hs = vt + ct
    Hqp = np.diag(np.concatenate([np.sort(rng.uniform(-1, -0.3, vt)), np.sort(rng.uniform(0.1, 2, ct))]))
    if offdiag:
        B = 0.01 * rng.normal(size=(hs, hs))
        Hqp = Hqp + B + B.T
    homo = rpamin + vmin_off + vt - 1
    vmin = rpamin + vmin_off
    cmax = homo + ct
    return M, eps, Hqp, homo, rpamin, vmin, cmax


@pytest.mark.parametrize("c", [(1, 2, 1, 0), (1, 0, 1, 0), (0, 2, 0, 1), (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0),
                               (0, 0, 0, 1), (1, 1, 1, 0), (0, 1, 0, 1)])
def test_bse_operator_random_fp64(ctx, c):
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(10, vmin_off=2, offdiag=True)
    t = dev_mmn(ctx, M, mmin=rpamin, nmin=rpamin)
    op = px.BSE_OPERATOR(*c, eps, t, Hqp).configure(px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    ref_op = X.BSEOperatorOracle(*c, eps, M, Hqp, homo, rpamin, 0, vmin, cmax)
    Xv = np.random.default_rng(11).normal(size=(op.size, 5))
    assert relerr(op.matmul(Xv), ref_op.matmul(Xv)) < 1e-13
    assert relerr(op.diagonal(), ref_op.diagonal()) < 1e-13


@pytest.mark.parametrize("prec,tol", [("bf16x3", 1e-4), ("bf16", 3e-2)])
@pytest.mark.parametrize("c", [(1, 2, 1, 0), (0, 2, 0, 1)])
def test_bse_operator_random_mixed(ctx, c, prec, tol):
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(12, vt=11, ct=90, P=130)
    t = dev_mmn(ctx, M)
    op = px.BSE_OPERATOR(*c, eps, t, Hqp, precisions=("fp64", prec)).configure(
        px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    Xv = np.random.default_rng(13).normal(size=(op.size, 7))
    ref = op.matmul(Xv, "fp64")
    assert relerr(op.matmul(Xv, prec), ref) < tol


@pytest.mark.parametrize("shape", [(11, 90, 130, 7), (40, 200, 300, 8)])
@pytest.mark.parametrize("c", [(0, 0, 1, 0), (0, 0, 0, 1)])
def test_bf16_kernels_agree(ctx, c, shape, monkeypatch):
    """The BF16 GEMM variants (register-staged with 32-, 64- and 128-wide K steps, one or two LDS
    stages, register caps for 3 or 4 waves per SIMD, and the direct-to-LDS pipeline) on the Hd /
    Hd2 stages, split and plain, against each other and FP64; the larger shape takes the
    pipeline's 256x128 tile, the smaller its 128x128 one, both with ragged edges (row and K
    extents that are not tile multiples)."""
    vt, ct, P, k = shape
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(14, vt=vt, ct=ct, P=P)
    t = dev_mmn(ctx, M)
    op = px.BSE_OPERATOR(*c, eps, t, Hqp, precisions=("fp64", "bf16", "bf16x3")).configure(
        px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    Xv = np.random.default_rng(15).normal(size=(op.size, k))
    ref = op.matmul(Xv, "fp64")
    ys, yp = {}, {}
    for kern in ("staged32", "staged64", "wide", "pipe", "big", "tall", "db64", "db128", "k64occ3", "k32occ3",
                 "lds256", "lds128"):
        monkeypatch.setenv("XTPGPU_BF16_KERNEL", kern)
        ys[kern] = op.matmul(Xv, "bf16x3")
        assert relerr(ys[kern], ref) < 2e-5, kern
        yp[kern] = op.matmul(Xv, "bf16")
        assert relerr(yp[kern], ref) < 3e-2, kern
    assert relerr(ys["pipe"], ys["staged32"]) < 2e-5
    assert relerr(ys["staged64"], ys["staged32"]) < 2e-5
    # the plain forms differ only in FP32 summation order
    for kern in yp:
        assert relerr(yp[kern], yp["staged32"]) < 1e-4, kern


@pytest.mark.parametrize("vt,ct,P", [(1, 1, 1), (1, 17, 3), (9, 1, 5), (3, 8, 1)])
@pytest.mark.parametrize("c", [(1, 2, 1, 0), (0, 2, 0, 1), (1, 0, 1, 0)])
def test_bse_operator_degenerate_shapes(ctx, c, vt, ct, P):
    """Single valence / conduction level and single auxiliary function: every padded extent is a
    ragged tile edge. FP64 against the oracle, BF16x3 against FP64; k = 1 and k = 0 columns."""
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(30 + vt + ct + P, vt=vt, ct=ct, P=P, offdiag=True)
    t = dev_mmn(ctx, M)
    op = px.BSE_OPERATOR(*c, eps, t, Hqp, precisions=("fp64", "bf16x3")).configure(
        px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    ref_op = X.BSEOperatorOracle(*c, eps, M, Hqp, homo, rpamin, 0, vmin, cmax)
    assert op.size == vt * ct
    for k in (1, 3):
        Xv = np.random.default_rng(k).normal(size=(op.size, k))
        ref = ref_op.matmul(Xv)
        assert relerr(op.matmul(Xv), ref) < 1e-13
        assert relerr(op.matmul(Xv, "bf16x3"), ref) < 2e-5
    assert relerr(op.diagonal(), ref_op.diagonal()) < 1e-13
    assert op.matmul(np.zeros((op.size, 0))).shape == (op.size, 0)


def test_argument_errors_raise(ctx):
    """Shape and option errors come back as exceptions (the reference throws std::runtime_error)
    and leave the context usable."""
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(31, vt=3, ct=5, P=4)
    t = dev_mmn(ctx, M)
    opt = px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax)
    with pytest.raises(ValueError):
        px.BSE_OPERATOR(0, 0, 1, 1, eps, t, Hqp)
    op = px.BSE_OPERATOR(1, 2, 1, 0, eps, t, Hqp).configure(opt)  # FP64 only
    with pytest.raises(px.XtpGpuError):
        op.matmul(np.ones((op.size, 2)), "bf16x3")
    with pytest.raises(px.XtpGpuError):
        px._check(px.lib().xtpgpu_bse_matmul(op.h, 0, px._dptr(np.ones(op.size)), op.size - 1, 1,
                                             px._dptr(np.ones(op.size)), op.size))
    ds = px.DavidsonSolver()
    with pytest.raises(px.XtpGpuError):
        ds.solve(op, op.size)  # 2*neigen exceeds the operator size
    Xv = np.random.default_rng(0).normal(size=(op.size, 2))
    assert relerr(op.matmul(Xv), X.BSEOperatorOracle(1, 2, 1, 0, eps, M, Hqp, homo, rpamin, 0, vmin, cmax).matmul(Xv)) < 1e-13


def test_device_copy_kernel(ctx):
    rng = np.random.default_rng(5)
    a = rng.normal(size=4096 * 3)
    src, dst = px.DeviceBuffer(ctx, a.nbytes), px.DeviceBuffer(ctx, a.nbytes)
    src.upload(a)
    px._check(px.lib().xtpgpu_copy_device(ctx.h, dst.ptr, src.ptr, a.nbytes))
    assert np.array_equal(dst.download(a.shape), a)
    with pytest.raises(px.XtpGpuError):
        px._check(px.lib().xtpgpu_copy_device(ctx.h, dst.ptr, src.ptr, 8))


def test_davidson_single_root(ctx):
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(32, vt=4, ct=25, P=30, scale=0.05)
    t = dev_mmn(ctx, M)
    op = px.SingletOperator_TDA(eps, t, Hqp).configure(px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    w = np.linalg.eigvalsh(X.bse_dense(1, 2, 1, 0, eps, M, Hqp, homo, rpamin, vmin, cmax))
    ds = px.DavidsonSolver()
    ds.set_tolerance(1e-9)
    ev, vec = ds.solve(op, 1)
    assert ds.converged()
    assert abs(ev[0] - w[0]) < 1e-10


@pytest.mark.parametrize("gs", ["block", "column"])
def test_davidson_fp64_matches_dense(ctx, gs, monkeypatch):
    """FP64 Davidson against the dense spectrum, with the block (projection + Cholesky QR) and the
    column-by-column Gram-Schmidt."""
    monkeypatch.setenv("XTPGPU_GS", gs)
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(20, vt=8, ct=40, P=60, scale=0.05)
    t = dev_mmn(ctx, M)
    op = px.SingletOperator_TDA(eps, t, Hqp).configure(px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    H = X.bse_dense(1, 2, 1, 0, eps, M, Hqp, homo, rpamin, vmin, cmax)
    w = np.linalg.eigvalsh(H)
    ds = px.DavidsonSolver()
    ds.set_tolerance(1e-9)
    ds.set_max_search_space(50)
    ev, vec = ds.solve(op, 5)
    assert ds.converged()
    assert np.abs(ev - w[:5]).max() < 1e-8
    assert np.abs(np.linalg.norm(vec, axis=0) - 1).max() < 1e-12
    r = H @ vec - vec * ev
    assert np.linalg.norm(r, axis=0).max() < 1e-8


@pytest.mark.parametrize("prec", ["bf16", "bf16x3"])
def test_davidson_mixed_refined_matches_dense(ctx, prec):
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(21, vt=8, ct=40, P=60, scale=0.05)
    t = dev_mmn(ctx, M)
    op = px.SingletOperator_TDA(eps, t, Hqp, precisions=("fp64", prec)).configure(
        px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    w = np.linalg.eigvalsh(X.bse_dense(1, 2, 1, 0, eps, M, Hqp, homo, rpamin, vmin, cmax))
    ds = px.DavidsonSolver()
    ds.set_max_search_space(50)
    # the mixed loop runs down to its precision floor, the FP64 rounds take the residual the rest
    # of the way; the energy error is then second order in the FP64 residual
    mixed_tol, refine_tol, etol = {"bf16x3": (1e-6, 1e-7, 1e-10), "bf16": (1e-3, 1e-5, 1e-8)}[prec]
    ds.set_tolerance(mixed_tol)
    ds.set_iter_max(100)
    ds.set_precision(prec, refine_tolerance=refine_tol, refine_max=12)
    ev, _ = ds.solve(op, 5)
    assert ds.converged(), (ds.info.max_residual, ds.info.refine_rounds)
    assert ds.info.max_residual < refine_tol
    assert np.abs(ev - w[:5]).max() < etol


def test_davidson_precision_ladder_matches_dense(ctx):
    """BF16 iterations to 1e-4, restart on the Ritz vectors with BF16x3, FP64 refinement."""
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(21, vt=8, ct=40, P=60, scale=0.05)
    t = dev_mmn(ctx, M)
    op = px.SingletOperator_TDA(eps, t, Hqp, precisions=("fp64", "bf16", "bf16x3")).configure(
        px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    w = np.linalg.eigvalsh(X.bse_dense(1, 2, 1, 0, eps, M, Hqp, homo, rpamin, vmin, cmax))
    ds = px.DavidsonSolver()
    ds.set_max_search_space(50)
    ds.set_tolerance(1e-6)
    ds.set_iter_max(100)
    ds.set_precision("bf16x3", refine_tolerance=1e-7, refine_max=12, start="bf16", switch_tolerance=1e-4)
    ev, vec = ds.solve(op, 5)
    assert ds.converged(), (ds.info.max_residual, ds.info.refine_rounds)
    assert np.abs(ev - w[:5]).max() < 1e-10
    # an operator built without the start precision is refused
    op2 = px.SingletOperator_TDA(eps, t, Hqp, precisions=("fp64", "bf16x3")).configure(
        px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    with pytest.raises(px.XtpGpuError):
        ds.solve(op2, 5)


def test_davidson_forced_failure(ctx):  # test_davidson.cc:107-125
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(22, vt=5, ct=30, P=40)
    t = dev_mmn(ctx, M)
    op = px.SingletOperator_TDA(eps, t, Hqp).configure(px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax))
    ds = px.DavidsonSolver()
    ds.set_tolerance(1e-12)
    ds.set_iter_max(1)
    ev, vec = ds.solve(op, 4)
    assert not ds.converged()


SIGMA_PPM_E = np.array([-0.612601, -0.341755, -0.341755, -0.341755, 0.137304, 0.16678, 0.16678, 0.16678, 0.671592,
                        0.671592, 0.671592, 0.974255, 1.01205, 1.01205, 1.01205, 1.64823, 19.4429])
GW_E = np.array([-10.6784, -0.746424, -0.394948, -0.394948, -0.394948, 0.165212, 0.227713, 0.227713, 0.227713,
                 0.763971, 0.763971, 0.763971, 1.05054, 1.13372, 1.13372, 1.13372, 1.72964])


def ch4_tensor(ctx, ints, mos):
    S, Saux, V, T = ints
    C = mmio.read_matrix(os.path.join(G, mos))
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 16, 0, 16)
    t.Fill(T, C, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    return t


def test_sigma_ppm_fixtures_gpu(ctx, ints):  # test_sigma_ppm.cc:44-121
    t = ch4_tensor(ctx, ints, "sigma_ppm/MOs.mm")
    M = t.get_tensor()
    e = SIGMA_PPM_E
    rpa = px.RPA(t)
    rpa.configure(4, 0, 16)
    rpa.setRPAInputEnergies(e)
    sig = px.Sigma_PPM(t, rpa).configure(4, 0, 16, 0, 16, 1e-3)
    x = sig.CalcExchangeMatrix()
    assert mmio.isapprox(x, mmio.read_matrix(os.path.join(G, "sigma_ppm/x_ref.mm")), 1e-5)
    assert relerr(x, X.sigma_x(M, 4, 0, 0, 16)) < 1e-13
    sig.PrepareScreening()
    phi, w, f = X.ppm_construct(M, e, 4, 0, 16)
    assert np.abs(sig.ppm_weight - w).max() < 1e-10
    assert np.abs(sig.ppm_freq - f).max() < 1e-8
    c = sig.CalcCorrelationOffDiag(e)
    c[np.diag_indices(17)] = sig.CalcCorrelationDiag(e)
    ref = mmio.read_matrix(os.path.join(G, "sigma_ppm/c_ref.mm"))
    assert mmio.isapprox(np.diag(c), np.diag(ref), 1e-5)
    assert mmio.isapprox(c, ref, 1e-5)
    # derivative against a central difference on the device evaluator
    s, d = sig.diag_batch([3, 3, 3], [e[3] - 1e-5, e[3] + 1e-5, e[3]])
    assert abs((s[1] - s[0]) / 2e-5 - d[2]) < 1e-5 * max(1.0, abs(d[2]))



def test_ppm_known_answer_gpu(ctx):  # test_ppm.cc:36-90
    """Plasmon-pole frequencies and weights on the device for CH4 / 3-21G with the core-Hamiltonian
    MOs and energies: the test's values at isApprox 1e-4 and the oracle to 1e-8."""
    xyz, bas = os.path.join(G, "ppm/molecule.xyz"), os.path.join(G, "ppm/3-21G.xml")
    e, Cm = np.linalg.eigh(I.core_hamiltonian(xyz, bas))
    S, Saux, V, T = I.ao_integrals(xyz, bas)
    t = px.TCMatrix_gwbse(ctx).Initialize(17, 0, 16, 0, 16)
    t.Fill(T, Cm, px.Pseudo_InvSqrt_GWBSE(V, Saux, 5e-7, ctx)[0])
    M = t.get_tensor()
    rpa = px.RPA(t)
    rpa.configure(4, 0, 16)
    rpa.setRPAInputEnergies(e)
    sig = px.Sigma_PPM(t, rpa).configure(4, 0, 16, 0, 16, 1e-3)
    sig.PrepareScreening()
    freq = [19.4503, 12.2429, 10.2167, 10.2167, 10.2167, 17.0832, 12.7117, 12.7117, 12.7117, 10.9455, 10.9455,
            10.9455, 11.1861, 9.60843, 9.60843, 9.60843, 9.61518]
    w = [3.59422e-05, 0.00121795, 0.00343632, 0.00343632, 0.00343632, 0.00394659, 0.012066, 0.012066, 0.012066,
         0.0241118, 0.0241118, 0.0241118, 0.0286786, 0.11036, 0.11036, 0.11036, 0.191732]
    assert mmio.isapprox(sig.ppm_freq, np.array(freq), 1e-4)
    assert mmio.isapprox(sig.ppm_weight, np.array(w), 1e-4)
    phi, wo, fo = X.ppm_construct(M, e, 4, 0, 16)
    # the first pole carries a weight of 3.6e-5: its frequency amplifies rounding in eps ~1/w
    assert np.abs(sig.ppm_freq - fo).max() < 1e-6
    assert np.abs(sig.ppm_weight - wo).max() < 1e-10


def test_rpa_h2p_fixture_gpu(ctx, ints):  # test_rpa_h2p.cc:41-110
    """RPA two-particle Hamiltonian on the device: the 60 poles (isApprox 1e-4), |diag(X+Y)|
    (1e-4) and E_c (1e-6 relative) of the reference's test, and the oracle to 1e-10."""
    from test_oracle_golden import H2P_ERPA, H2P_OMEGA, H2P_XPY_DIAG
    t = ch4_tensor(ctx, ints, "rpa/eigenvectors.mm")
    e = mmio.read_vector(os.path.join(G, "rpa/eigenvals.mm"))
    rpa = px.RPA(t)
    rpa.configure(4, 0, 16)
    rpa.setRPAInputEnergies(e)
    sol = rpa.Diagonalize_H2p()
    assert abs(sol.ERPA_correlation - H2P_ERPA) <= 1e-6 * abs(H2P_ERPA)
    assert mmio.isapprox(sol.omega, H2P_OMEGA, 1e-4)
    assert mmio.isapprox(np.abs(np.diag(sol.XpY)), np.abs(H2P_XPY_DIAG), 1e-4)
    om, xpy, er = X.rpa_h2p(t.get_tensor(), e, 4, 0, 16)
    assert np.abs(sol.omega - om).max() < 1e-10 and abs(sol.ERPA_correlation - er) < 1e-12
    # X+Y columns agree through the projector on each group of poles closer than 1e-5 (the
    # degenerate and near-degenerate pairs; a pole split by g moves its vector by ~rounding / g)
    for s0 in range(60):
        grp = np.where(np.abs(om - om[s0]) < 1e-5)[0]
        Pd = sol.XpY[:, grp] @ np.linalg.pinv(sol.XpY[:, grp])
        Po = xpy[:, grp] @ np.linalg.pinv(xpy[:, grp])
        assert np.abs(Pd - Po).max() < 1e-7

@pytest.mark.parametrize("mode,finder", [("adaptive_with_dense_fallback", "bisection"),
                                         ("adaptive_with_dense_fallback", "brent"), ("dense", "bisection"),
                                         ("adaptive", "brent")])
@pytest.mark.parametrize("case", [("mo_eigenvectors.mm", "vxc.mm", "ref.mm"),
                                  ("mo_eigenvectors2.mm", "vxc2.mm", "ref2.mm")])
def test_gw_fixture_gpu(ctx, ints, case, mode, finder):  # test_gw.cc:111-340
    """G0W0-PPM with the grid QP solver on the device evaluator, the reference test's options
    (legacy 601 x 0.005 Ha grid): QP energies and H_QP against gw/ref.mm and gw/ref2.mm (isApprox
    1e-4), and the QP energies against the oracle's GW::SolveQP restatement on its own Sigma (1e-8)."""
    from oracle import qp_solver as Q
    t = ch4_tensor(ctx, ints, f"gw/{case[0]}")
    M = t.get_tensor()
    vxc = mmio.read_matrix(os.path.join(G, f"gw/{case[1]}"))
    kw = dict(qp_solver="grid", qp_grid_steps=601, qp_grid_spacing=0.005, qp_grid_search_mode=mode,
              qp_root_finder=finder)
    gw = px.GW(t, vxc, GW_E).configure(4, 0, 16, 0, 16, eta=1e-3, **kw)
    gw.CalculateGWPerturbation()
    ref = mmio.read_matrix(os.path.join(G, f"gw/{case[2]}"))
    assert gw.qp_converged.all()
    assert mmio.isapprox(gw.getGWAResults(), np.diag(ref), 1e-4)
    e = GW_E
    sx = X.sigma_x(M, 4, 0, 0, 16)
    phi, w, f = X.ppm_construct(M, e, 4, 0, 16)
    Mr = X.multiply_right(M, phi)
    sig = lambda l, om: X.sigma_ppm_diag(Mr, e, w, f, 4, 0, 0, 1e-3, l, om)[0]  # noqa: E731
    dsig = lambda l, om: X.sigma_ppm_diag(Mr, e, w, f, 4, 0, 0, 1e-3, l, om)[1]  # noqa: E731
    oq = Q.from_legacy_grid(601, 0.005, qp_grid_search_mode=mode, qp_root_finder=finder)
    out, _ = Q.solve_qp(sig, dsig, e + np.diag(sx) - np.diag(vxc), e, 4, 0, oq)
    gwa = np.diag(sx) + np.array([sig(l, out[l]) for l in range(17)]) - np.diag(vxc) + e
    assert np.abs(gw.getGWAResults() - gwa).max() < 1e-8
    gw.CalculateHQP()
    assert mmio.isapprox(gw.getHQP(), ref, 1e-4)


def test_gw_cda_qp_gpu(ctx, ints):  # test_sigma_cda.cc inputs; GW::SolveQP on Sigma_CDA (sigmafactory.cc:33)
    """G0W0 with Sigma_c by contour deformation (Legendre order 12) on the CH4 case: Sigma_x, the
    device CDA evaluator and the grid QP solver, against the oracle's SolveQP on the oracle's
    Sigma_CDA with the same central-difference derivative (1e-8 Ha)."""
    from oracle import qp_solver as Q
    from votca_amd import quadrature as QD
    t = ch4_tensor(ctx, ints, "gw/mo_eigenvectors.mm")
    M = t.get_tensor()
    vxc = mmio.read_matrix(os.path.join(G, "gw/vxc.mm"))
    gw = px.GW(t, vxc, GW_E).configure(4, 0, 16, 0, 16, eta=1e-3, sigma_integration="cda", quadrature_order=12,
                                        qp_solver="grid", qp_grid_steps=601, qp_grid_spacing=0.005)
    gw.CalculateGWPerturbation()
    q = QD.create("legendre", 12)
    ora = X.SigmaCDAOracle(M, GW_E, 4, 0, 16, 0, 1e-3, 1e-3, q.points(), q.weights(), q.use_symmetry())
    sig = lambda l, om: ora.diag(om, l)  # noqa: E731
    dsig = lambda l, om: (ora.diag(om + 1e-3, l) - ora.diag(om - 1e-3, l)) / 2e-3  # noqa: E731
    sx = X.sigma_x(M, 4, 0, 0, 16)
    out, conv = Q.solve_qp(sig, dsig, GW_E + np.diag(sx) - np.diag(vxc), GW_E, 4, 0, Q.from_legacy_grid(601, 0.005))
    gwa = np.diag(sx) + np.array([sig(l, out[l]) for l in range(17)]) - np.diag(vxc) + GW_E
    assert list(gw.qp_converged) == list(conv)
    assert np.abs(gw.getGWAResults() - gwa).max() < 1e-8
    # CDA and PPM agree on the (triply degenerate) HOMO to within the plasmon-pole model's error
    # (~0.05 eV here); the deep 2a1 level is where the single-pole model is known to be poor
    ppm = px.GW(ch4_tensor(ctx, ints, "gw/mo_eigenvectors.mm"), vxc, GW_E).configure(
        4, 0, 16, 0, 16, qp_solver="grid", qp_grid_steps=601, qp_grid_spacing=0.005)
    ppm.CalculateGWPerturbation()
    assert np.abs(ppm.getGWAResults()[2:5] - gw.getGWAResults()[2:5]).max() < 0.01


@pytest.mark.parametrize("sigma", ["ppm", "cda"])
def test_gw_mixed_screening_gpu(ctx, ints, sigma):
    """Mixed-precision GW: the screening's eps builds in BF16x3 (split bf16 operands, FP32
    accumulation), inversions / eigensolves / QP equation in FP64. Per-QP-level error against the
    FP64 run on the CH4 case (grid solver): below 1e-5 Ha on every level, and all levels converge."""
    kw = dict(qp_solver="grid", qp_grid_steps=601, qp_grid_spacing=0.005, sigma_integration=sigma,
              quadrature_order=12)
    vxc = mmio.read_matrix(os.path.join(G, "gw/vxc.mm"))
    res = {}
    for prec in ("fp64", "bf16x3"):
        gw = px.GW(ch4_tensor(ctx, ints, "gw/mo_eigenvectors.mm"), vxc, GW_E).configure(
            4, 0, 16, 0, 16, eta=1e-3, screening_precision=prec, **kw)
        gw.CalculateGWPerturbation()
        assert gw.qp_converged.all(), prec
        res[prec] = gw.getGWAResults()
    err = np.abs(res["bf16x3"] - res["fp64"])
    print(f"[mixed GW {sigma}] max |dE_QP| = {err.max():.2e} Ha", flush=True)
    assert err.max() < 1e-5, err


@pytest.mark.parametrize("mix,max_it,converges", [(0, 6, False), (2, 20, True)])
def test_evgw_ppm_gpu(ctx, ints, mix, max_it, converges):  # gw.cc:218-310
    """Eigenvalue-self-consistent GW (plasmon-pole) on the CH4 case to a 1e-5 Ha change of the RPA
    energies: with Anderson mixing (order 2) it converges (9 iterations in the oracle); without
    mixing a virtual level alternates between two roots and the loop stops at its iteration cap.
    The tensor is restored every 5 iterations. Final QP and RPA energies against the oracle's evGW
    loop (1e-7 Ha)."""
    t = ch4_tensor(ctx, ints, "gw/mo_eigenvectors.mm")
    M = t.get_tensor()
    vxc = mmio.read_matrix(os.path.join(G, "gw/vxc.mm"))
    kw = dict(qp_solver="grid", qp_grid_steps=601, qp_grid_spacing=0.005)
    gw = px.GW(t, vxc, GW_E).configure(4, 0, 16, 0, 16, eta=1e-3, gw_sc_max_iterations=max_it, gw_sc_limit=1e-5,
                                        gw_mixing_order=mix, gw_mixing_alpha=0.7, reset_3c=5, **kw)
    gw.CalculateGWPerturbation()
    assert gw.gw_converged == converges and gw.gw_iterations > 1
    from oracle import qp_solver as Q
    gwa, e_rpa, it = X.evgw_ppm(M, GW_E, vxc, 4, 0, 16, 0, 16, 1e-3, Q.from_legacy_grid(601, 0.005), max_it, 1e-5,
                                mix_order=mix, mix_alpha=0.7)
    assert it == gw.gw_iterations
    assert np.abs(gw.getGWAResults() - gwa).max() < 1e-7
    assert np.abs(gw.rpa.getRPAInputEnergies() - e_rpa).max() < 1e-7
    # evGW moves the gap away from G0W0's
    g0 = px.GW(ch4_tensor(ctx, ints, "gw/mo_eigenvectors.mm"), vxc, GW_E).configure(4, 0, 16, 0, 16, **kw)
    g0.CalculateGWPerturbation()
    assert abs((gwa[5] - gwa[4]) - (g0.getGWAResults()[5] - g0.getGWAResults()[4])) > 1e-4


def test_gw_fixed_point_matches_oracle(ctx, ints):
    t = ch4_tensor(ctx, ints, "gw/mo_eigenvectors.mm")
    M = t.get_tensor()
    vxc = mmio.read_matrix(os.path.join(G, "gw/vxc.mm"))
    e = GW_E
    gw = px.GW(t, vxc, e).configure(4, 0, 16, 0, 16, eta=1e-3, qp_solver="fixedpoint")
    gw.CalculateGWPerturbation()
    sx = X.sigma_x(M, 4, 0, 0, 16)
    phi, w, f = X.ppm_construct(M, e, 4, 0, 16)
    Mr = X.multiply_right(M, phi)
    inter = e + np.diag(sx) - np.diag(vxc)
    for l in range(17):
        r = X.qp_fixed_point(lambda x: X.sigma_ppm_diag(Mr, e, w, f, 4, 0, 0, 1e-3, l, x), inter[l], e[l])
        s = X.sigma_ppm_diag(Mr, e, w, f, 4, 0, 0, 1e-3, l, r)[0]
        assert abs(gw.getGWAResults()[l] - (sx[l, l] + s - vxc[l, l] + e[l])) < 1e-9


def test_setup_screening_gpu(ctx):
    rng = np.random.default_rng(30)
    o, v, P = 6, 40, 50
    M = rng.normal(scale=0.1, size=(o + v, o + v, P))
    M = 0.5 * (M + M.transpose(1, 0, 2))
    e = np.concatenate([np.sort(rng.uniform(-1, -0.3, o)), np.sort(rng.uniform(0.1, 2, v))])
    t = dev_mmn(ctx, M)
    ei = px.setup_screening(t, e, o - 1, 0, o + v - 1)
    lam, U = np.linalg.eigh(X.rpa_epsilon(M, e, o - 1, 0, o + v - 1, 0.0, "real"))
    assert np.abs(ei - 1.0 / lam).max() < 1e-12
    # the rotated tensor agrees up to the sign of each eigenvector
    Mr = X.multiply_right(M, U)
    G_ = t.get_tensor()
    sgn = np.sign(np.sum(G_ * Mr, axis=(0, 1)))
    assert relerr(G_ * sgn[None, None, :], Mr) < 1e-10


@pytest.mark.parametrize("gs", ["block", "column"])
def test_full_bse_davidson_matches_dense(ctx, gs, monkeypatch):
    monkeypatch.setenv("XTPGPU_GS", gs)
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(40, vt=6, ct=30, P=50, scale=0.05)
    t = dev_mmn(ctx, M)
    opt = px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax)
    A = px.SingletOperator_TDA(eps, t, Hqp).configure(opt)
    B = px.SingletOperator_BTDA_B(eps, t, Hqp).configure(opt)
    Ad = X.bse_dense(1, 2, 1, 0, eps, M, Hqp, homo, rpamin, vmin, cmax)
    Bd = X.bse_dense(0, 2, 0, 1, eps, M, Hqp, homo, rpamin, vmin, cmax)
    H = np.block([[Ad, Bd], [-Bd, -Ad]])
    w = np.sort(np.linalg.eigvals(H).real)
    w = w[w > 0][:4]
    # closed form for real symmetric A, B: Omega^2 = eig((A-B)^1/2 (A+B) (A-B)^1/2)
    lam, U = np.linalg.eigh(Ad - Bd)
    S = (U * np.sqrt(lam)) @ U.T
    w2 = np.sqrt(np.linalg.eigvalsh(S @ (Ad + Bd) @ S))[:4]
    assert np.abs(w - w2).max() < 1e-10
    ds = px.DavidsonSolver()
    ds.set_tolerance(1e-8)
    ds.set_max_search_space(40)  # 10 * nmax, as BSE::Solve_nonhermitian_Davidson sets it
    ds.set_iter_max(100)
    ev, Xv, Yv = ds.solve_full(A, B, 4)
    assert ds.converged(), ds.info.max_residual
    assert np.abs(ev - w).max() < 1e-7
    # X/Y normalisation |X|^2 - |Y|^2 = 1 and the eigen-equation
    assert np.abs(np.sum(Xv ** 2, 0) - np.sum(Yv ** 2, 0) - 1).max() < 1e-10
    r = Ad @ Xv + Bd @ Yv - Xv * ev
    assert np.abs(r).max() < 1e-6


def _perturb_sym(M, P, v, c, vmin_r, cmin_r, h):
    Mp = M.copy()
    Mp[vmin_r + v, cmin_r + c, P] += h
    Mp[cmin_r + c, vmin_r + v, P] += h
    return Mp


@pytest.mark.parametrize("full", [False, True])
def test_adjoint_matches_finite_differences(ctx, full):
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(50, vt=5, ct=22, P=30, scale=0.05)
    vt, ct = homo - vmin + 1, cmax - homo
    t = dev_mmn(ctx, M)
    opt = px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax)
    A = px.SingletOperator_TDA(eps, t, Hqp).configure(opt)
    B = px.SingletOperator_BTDA_B(eps, t, Hqp).configure(opt) if full else None
    ds = px.DavidsonSolver()
    ds.set_tolerance(1e-10)
    nr = 3
    ds.set_max_search_space(10 * nr)
    ds.set_iter_max(100)
    if full:
        ev, Xv, Yv = ds.solve_full(A, B, nr)
    else:
        ev, Xv = ds.solve(A, nr)
        Yv = None
    assert ds.converged()
    d_eps, d_m = px.bse_adjoint(A, Xv, B, Yv)

    def omegas(Mx, H):
        Ad = X.bse_dense(1, 2, 1, 0, eps, Mx, H, homo, rpamin, vmin, cmax)
        if not full:
            return np.linalg.eigvalsh(Ad)[:nr]
        Bd = X.bse_dense(0, 2, 0, 1, eps, Mx, H, homo, rpamin, vmin, cmax)
        w = np.sort(np.linalg.eigvals(np.block([[Ad, Bd], [-Bd, -Ad]])).real)
        return w[w > 0][:nr]

    assert np.abs(omegas(M, Hqp) - ev).max() < 1e-8
    h = 1e-5
    rng = np.random.default_rng(51)
    for lvl in rng.choice(vt + ct, 6, replace=False):
        Hp, Hm = Hqp.copy(), Hqp.copy()
        Hp[lvl, lvl] += h
        Hm[lvl, lvl] -= h
        fd = (omegas(M, Hp) - omegas(M, Hm)) / (2 * h)
        assert np.abs(fd - d_eps[lvl]).max() < 1e-5 * max(1.0, np.abs(fd).max()), (lvl, fd, d_eps[lvl])
    for _ in range(6):
        P, v, c = rng.integers(30), rng.integers(vt), rng.integers(ct)
        fd = (omegas(_perturb_sym(M, P, v, c, vmin - rpamin, homo + 1 - rpamin, h), Hqp)
              - omegas(_perturb_sym(M, P, v, c, vmin - rpamin, homo + 1 - rpamin, -h), Hqp)) / (2 * h)
        an = d_m[:, P, v, c]
        assert np.abs(fd - an).max() < 1e-5 * max(1e-3, np.abs(fd).max()), (P, v, c, fd, an)


def test_autograd_function_matches_finite_differences(ctx):
    import torch
    from votca_amd.autograd import BSEProblem, excitation_energies
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(60, vt=4, ct=18, P=24, scale=0.05)
    vt, ct = homo - vmin + 1, cmax - homo
    t = dev_mmn(ctx, M)
    prob = BSEProblem(t, eps, homo, rpamin, vmin, cmax, nroots=2, full=True, tol=1e-10)
    e = torch.tensor(np.diag(Hqp).copy(), dtype=torch.float64, requires_grad=True)
    m = torch.tensor(np.transpose(M[0:vt, vt:vt + ct, :], (2, 0, 1)).copy(), dtype=torch.float64,
                     requires_grad=True)
    w = torch.tensor([0.7, -0.3], dtype=torch.float64)
    loss = (excitation_energies(e, m, prob) * w).sum()
    loss.backward()
    h = 1e-5

    def L(ev, mv):
        with torch.no_grad():
            return float((excitation_energies(ev, mv, prob) * w).sum())

    for lvl in (0, vt + 1):
        ep, em = e.detach().clone(), e.detach().clone()
        ep[lvl] += h
        em[lvl] -= h
        fd = (L(ep, m.detach()) - L(em, m.detach())) / (2 * h)
        assert abs(fd - float(e.grad[lvl])) < 1e-5 * max(1.0, abs(fd))
    for (P, v, c) in ((0, 1, 2), (5, 3, 10)):
        mp, mm = m.detach().clone(), m.detach().clone()
        mp[P, v, c] += h
        mm[P, v, c] -= h
        fd = (L(e.detach(), mp) - L(e.detach(), mm)) / (2 * h)
        assert abs(fd - float(m.grad[P, v, c])) < 1e-5 * max(1e-3, abs(fd))
    # the operators were packed once; new QP energies only replaced Hqp, new transitions repacked
    # the transition block (two per M entry)
    assert prob.packs == 1 and prob.refreshes == 4
    # a backward after another forward: the gradient of its own forward
    e2 = e.detach().clone().requires_grad_(True)
    m2 = m.detach().clone().requires_grad_(True)
    l2 = (excitation_energies(e2, m2, prob) * w).sum()
    with torch.no_grad():
        excitation_energies(e.detach(), mp, prob)
    l2.backward()
    assert torch.allclose(m2.grad, m.grad, rtol=1e-8, atol=1e-14) and torch.allclose(e2.grad, e.grad, rtol=1e-8)


def test_operator_set_hqp_and_refresh_match_fresh_pack(ctx):
    """xtpgpu_bse_set_hqp and xtpgpu_bse_refresh give the operator a fresh pack would: the same
    products bit for bit and the same diagonal."""
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(61, vt=5, ct=20, P=30, scale=0.05)
    vt, ct = homo - vmin + 1, cmax - homo
    t = dev_mmn(ctx, M)
    opt = px.BSEOperator_Options(homo, rpamin, 0, vmin, cmax)
    rng = np.random.default_rng(8)
    Xm = rng.normal(size=(vt * ct, 3))
    for c in ((1, 2, 1, 0), (0, 2, 0, 1), (1, 0, 1, 0)):
        op = px.BSE_OPERATOR(*c, eps, t, Hqp, precisions=("fp64", "bf16x3")).configure(opt)
        H2 = Hqp + np.diag(rng.uniform(-0.01, 0.01, Hqp.shape[0]))
        op.set_hqp(H2)
        m_new = 0.9 * np.transpose(M[vmin - rpamin:homo + 1 - rpamin, homo + 1 - rpamin:cmax + 1 - rpamin, :],
                                   (2, 0, 1))
        t.set_block(vmin - rpamin, homo + 1 - rpamin, m_new, symmetric=True)
        op.refresh_transitions()
        ref = px.BSE_OPERATOR(*c, eps, t, H2, precisions=("fp64", "bf16x3")).configure(opt)
        for prec in ("fp64", "bf16x3"):
            assert np.array_equal(op.matmul(Xm, prec), ref.matmul(Xm, prec)), (c, prec)
        assert np.array_equal(op.diagonal(), ref.diagonal()), c
        t.set_tensor(M)


CDA_E = np.array([0.0468207, 0.0907801, 0.0907801, 0.104563, 0.592491, 0.663355, 0.663355, 0.768373, 1.69292,
                  1.97724, 1.97724, 2.50877, 2.98732, 3.4418, 3.4418, 4.81084, 17.1838])


def test_sigma_cda_fixture_gpu(ctx, ints):  # test_sigma_cda.cc:41-113
    t = ch4_tensor(ctx, ints, "sigma_cda/MOs.mm")
    M = t.get_tensor()
    rpa = px.RPA(t)
    rpa.configure(4, 0, 16)
    rpa.setRPAInputEnergies(CDA_E)
    cda = px.Sigma_CDA(t, rpa).configure(4, 0, 16, 0, 16, eta=1e-3, alpha=1e-3, quadrature_scheme="legendre",
                                         order=100)
    x = cda.CalcExchangeMatrix()
    assert mmio.isapprox(x, mmio.read_matrix(os.path.join(G, "sigma_cda/x_ref.mm")), 1e-4)
    cda.PrepareScreening()
    c = cda.CalcCorrelationDiag(CDA_E)
    assert mmio.isapprox(c, np.diag(mmio.read_matrix(os.path.join(G, "sigma_cda/c_ref.mm"))), 1e-5)
    from votca_amd import quadrature as Q
    q = Q.create("legendre", 100)
    ora = X.SigmaCDAOracle(M, CDA_E, 4, 0, 16, 0, 1e-3, 1e-3, q.points(), q.weights(), q.use_symmetry())
    for l, w in ((3, -0.2), (7, 0.9), (12, 2.5)):
        assert abs(cda.diag_batch([l], [w])[0] - ora.diag(w, l)) < 1e-9
    # the GW eta does not enter Sigma_c^CDA (sigma_cda.cc:107,125 pass the RPA eta, rpa.h:90)
    cda2 = px.Sigma_CDA(t, rpa).configure(4, 0, 16, 0, 16, eta=1e-2, alpha=1e-3, quadrature_scheme="legendre",
                                          order=100)
    cda2.PrepareScreening()
    lv, wv = [3, 7, 12], [-0.2, 0.9, 2.5]
    assert np.abs(np.asarray(cda2.diag_batch(lv, wv)) - np.asarray(cda.diag_batch(lv, wv))).max() < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("bf16x3", 2e-5)])
def test_eps_batch_matches_per_segment(ctx, prec, tol):
    rng = np.random.default_rng(70)
    rpas, refs = [], []
    freqs = [0.0, 0.4, 2.0]
    for s_ in range(5):
        o = int(rng.integers(3, 9))
        v = int(o * rng.uniform(5, 8))
        P = 4 * (o + v) // 3
        M = rng.normal(scale=0.2, size=(o + v, o + v, P))
        e = np.concatenate([np.sort(rng.uniform(-1, -0.2, o)), np.sort(rng.uniform(0, 1.5, v))])
        t = dev_mmn(ctx, M)
        r = px.RPA(t)
        r.configure(o - 1, 0, o + v - 1)
        r.setRPAInputEnergies(e)
        rpas.append(r)
        refs.append(np.stack([X.rpa_epsilon(M, e, o - 1, 0, o + v - 1, f, "imag") for f in freqs], axis=2))
    b = px.EpsBatch(rpas, freqs, prec)
    for use_graph in (False, True, True):
        b.run(use_graph)
        for s_ in range(5):
            assert relerr(b.result(s_), refs[s_]) < tol


CPP_DRIVER = r"""
#include "xtpgpu.hpp"
#include <cstdio>
#include <fstream>
#include <vector>
// reads: ntot P vt ct, M[m][n][P] (m-major), eps[P], Hqp[(vt+ct)^2] column-major; prints 5 + 4 roots
int main(int argc, char** argv) {
  std::ifstream in(argv[1], std::ios::binary);
  int64_t hdr[4];
  in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
  const int64_t ntot = hdr[0], P = hdr[1], vt = hdr[2], ct = hdr[3], hs = vt + ct;
  std::vector<double> M(ntot * ntot * P), eps(P), H(hs * hs);
  in.read(reinterpret_cast<char*>(M.data()), M.size() * 8);
  in.read(reinterpret_cast<char*>(eps.data()), eps.size() * 8);
  in.read(reinterpret_cast<char*>(H.data()), H.size() * 8);
  xtpgpu::Context c(0);
  xtpgpu::TCMatrix t(c, P, 0, ntot - 1, 0, ntot - 1);
  std::vector<double> slice(ntot * P);
  for (int64_t m = 0; m < ntot; ++m) {
    for (int64_t n = 0; n < ntot; ++n)
      for (int64_t p = 0; p < P; ++p) slice[p * ntot + n] = M[(m * ntot + n) * P + p];
    t.set_slice(m, slice.data(), ntot);
  }
  xtpgpu::BSEOperatorOptions o{vt - 1, 0, 0, 0, hs - 1};
  xtpgpu::SingletOperator_TDA A(c, t, eps.data(), H.data(), hs, o);
  xtpgpu::SingletOperator_BTDA_B B(c, t, eps.data(), H.data(), hs, o);
  xtpgpu::DavidsonSolver ds;
  ds.set_tolerance(1e-9);
  ds.set_max_search_space(50);
  if (!ds.solve(A, 5)) return 2;
  for (double e : ds.eigenvalues()) std::printf("%.17g\n", e);
  xtpgpu::DavidsonSolver df;
  df.set_tolerance(1e-8);
  df.set_max_search_space(40);
  df.set_iter_max(100);
  if (!df.solve_full(A, B, 4)) return 3;
  for (double e : df.eigenvalues()) std::printf("%.17g\n", e);
  return 0;
}
"""


def test_cpp_wrapper_davidson_gpu(tmp_path):
    """include/xtpgpu.hpp driven from a C++ program (as VOTCA-XTP would link it): TCMatrix filled
    slice by slice, SingletOperator_TDA / _BTDA_B, DavidsonSolver.solve and .solve_full; the roots
    against the dense TDA and full-BSE spectra of the oracle's matrices."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    M, eps, Hqp, homo, rpamin, vmin, cmax = random_bse_case(40, vt=6, ct=30, P=50, scale=0.05)
    ntot, P = M.shape[0], M.shape[2]
    f = tmp_path / "in.bin"
    with open(f, "wb") as fh:
        fh.write(np.array([ntot, P, 6, 30], dtype=np.int64).tobytes())
        fh.write(np.ascontiguousarray(M).tobytes())
        fh.write(np.ascontiguousarray(eps, dtype=np.float64).tobytes())
        fh.write(np.asfortranarray(Hqp).tobytes(order="F"))
    src = tmp_path / "drv.cc"
    src.write_text(CPP_DRIVER)
    exe = tmp_path / "drv"
    libdir = os.path.dirname(px.LIB_PATH)
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", libdir, "-lxtpgpu", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("[cpp-wrapper] running the C++ driver", flush=True)
    r = subprocess.run([str(exe), str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    vals = np.array([float(x) for x in r.stdout.split()])
    Ad = X.bse_dense(1, 2, 1, 0, eps, M, Hqp, homo, rpamin, vmin, cmax)
    Bd = X.bse_dense(0, 2, 0, 1, eps, M, Hqp, homo, rpamin, vmin, cmax)
    assert np.abs(vals[:5] - np.linalg.eigvalsh(Ad)[:5]).max() < 1e-8
    w = np.sort(np.linalg.eigvals(np.block([[Ad, Bd], [-Bd, -Ad]])).real)
    assert np.abs(vals[5:] - w[w > 0][:4]).max() < 1e-7


CPP_SIGMA_DRIVER = r"""
#include "xtpgpu.hpp"
#include <cstdio>
#include <fstream>
#include <vector>
// reads: ntot P homo, M[m][n][P] (m-major), e[ntot]; prints Sigma_x then Sigma_c (ntot^2 each, column-major)
int main(int argc, char** argv) {
  std::ifstream in(argv[1], std::ios::binary);
  int64_t hdr[3];
  in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
  const int64_t ntot = hdr[0], P = hdr[1], homo = hdr[2];
  std::vector<double> M(ntot * ntot * P), e(ntot);
  in.read(reinterpret_cast<char*>(M.data()), M.size() * 8);
  in.read(reinterpret_cast<char*>(e.data()), e.size() * 8);
  xtpgpu::Context c(0);
  xtpgpu::TCMatrix t(c, P, 0, ntot - 1, 0, ntot - 1);
  std::vector<double> slice(ntot * P);
  for (int64_t m = 0; m < ntot; ++m) {
    for (int64_t n = 0; n < ntot; ++n)
      for (int64_t p = 0; p < P; ++p) slice[p * ntot + n] = M[(m * ntot + n) * P + p];
    t.set_slice(m, slice.data(), ntot);
  }
  xtpgpu::Sigma_PPM s(c, t);
  s.configure(homo, 0, ntot - 1, 0, ntot - 1, 1e-3);
  s.setRPAInputEnergies(e);
  std::vector<double> x = s.CalcExchangeMatrix();
  s.PrepareScreening();
  std::vector<double> cmat = s.CalcCorrelationOffDiag(e);
  std::vector<double> d = s.CalcCorrelationDiag(e);
  for (int64_t l = 0; l < ntot; ++l) cmat[l * ntot + l] = d[l];
  for (double v : x) std::printf("%.17g\n", v);
  for (double v : cmat) std::printf("%.17g\n", v);
  if (argc > 2) {  // G0W0: diag(vxc) follows; grid QP solver with the legacy 601 x 0.005 Ha window
    std::ifstream vin(argv[2], std::ios::binary);
    std::vector<double> vxc(ntot), icpt(ntot);
    vin.read(reinterpret_cast<char*>(vxc.data()), ntot * 8);
    for (int64_t l = 0; l < ntot; ++l) icpt[l] = e[l] + x[l * ntot + l] - vxc[l];
    xtpgpu_qp_opts o = xtpgpu::qp_defaults();
    xtpgpu::qp_legacy_grid(o, 601, 0.005);
    std::vector<int> conv;
    std::vector<double> w = s.SolveQP(icpt, e, o, &conv);
    std::vector<double> sc = s.CalcCorrelationDiag(w);
    for (int64_t l = 0; l < ntot; ++l) std::printf("%.17g %d\n", x[l * ntot + l] + sc[l] - vxc[l] + e[l], conv[l]);
  }
  return 0;
}
"""


def test_cpp_wrapper_sigma_ppm_gpu(tmp_path, ints):  # test_sigma_ppm.cc:44-121 through include/xtpgpu.hpp
    """xtpgpu::Sigma_PPM from a C++ program: exchange matrix, PPM screening and the correlation
    matrix against the sigma_ppm fixtures (isApprox 1e-5)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    S, Saux, V, T = ints
    M = X.fill_mmn(T, mmio.read_matrix(os.path.join(G, "sigma_ppm/MOs.mm")), 0, 16, 0, 16, Saux, V)
    f = tmp_path / "in.bin"
    with open(f, "wb") as fh:
        fh.write(np.array([17, M.shape[2], 4], dtype=np.int64).tobytes())
        fh.write(np.ascontiguousarray(M).tobytes())
        fh.write(np.ascontiguousarray(SIGMA_PPM_E, dtype=np.float64).tobytes())
    src = tmp_path / "sig.cc"
    src.write_text(CPP_SIGMA_DRIVER)
    exe = tmp_path / "sig"
    libdir = os.path.dirname(px.LIB_PATH)
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", libdir, "-lxtpgpu", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("[cpp-wrapper] running the Sigma_PPM driver", flush=True)
    r = subprocess.run([str(exe), str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    vals = np.array([float(v) for v in r.stdout.split()])
    x = vals[:289].reshape(17, 17, order="F")
    c = vals[289:].reshape(17, 17, order="F")
    assert mmio.isapprox(x, mmio.read_matrix(os.path.join(G, "sigma_ppm/x_ref.mm")), 1e-5)
    assert mmio.isapprox(c, mmio.read_matrix(os.path.join(G, "sigma_ppm/c_ref.mm")), 1e-5)
    # G0W0 QP energies through Sigma_PPM::SolveQP against gw/ref.mm (test_gw.cc:111-190)
    M = X.fill_mmn(T, mmio.read_matrix(os.path.join(G, "gw/mo_eigenvectors.mm")), 0, 16, 0, 16, Saux, V)
    with open(f, "wb") as fh:
        fh.write(np.array([17, M.shape[2], 4], dtype=np.int64).tobytes())
        fh.write(np.ascontiguousarray(M).tobytes())
        fh.write(np.ascontiguousarray(GW_E, dtype=np.float64).tobytes())
    fv = tmp_path / "vxc.bin"
    fv.write_bytes(np.ascontiguousarray(np.diag(mmio.read_matrix(os.path.join(G, "gw/vxc.mm")))).tobytes())
    print("[cpp-wrapper] running the G0W0 QP solve", flush=True)
    r = subprocess.run([str(exe), str(f), str(fv)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    lines = r.stdout.split("\n")[2 * 289:2 * 289 + 17]
    gwa = np.array([float(v.split()[0]) for v in lines])
    assert all(v.split()[1] == "1" for v in lines)
    assert mmio.isapprox(gwa, np.diag(mmio.read_matrix(os.path.join(G, "gw/ref.mm"))), 1e-4)
