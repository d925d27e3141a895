"""GPU tests of the mixed-precision path: the low-precision trailing update
(mrm_dev_gemm_lowp) and the LU with FP64 iterative refinement
(mrm_dev_lineq_solve_ir, mrm_lineq_solve_ex).

Bars: the update is exact on integer inputs the low precision represents and
within its rounding bound otherwise; every solve reaches the normwise backward
error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) <= 1e-14
(recomputed here in long double by the oracle), and its solution agrees with
the oracle's FP64 solve to the conditioning: ||x - x_oracle|| / ||x|| <=
kappa * 1e-13.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
from mrmath_amd import _lib
from mrmath_amd import linalg as la

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _lowp_update(mode, A, B, C0):
    import torch
    lib = _lib.load()
    M, K = A.shape
    N = B.shape[1]
    dA, dB, dC = _dev(A), _dev(B), _dev(C0)
    wb = lib.mrm_lowp_workspace_bytes(M, N, K, mode)
    W = torch.empty(max(wb, 1), dtype=torch.uint8, device="cuda")
    rc = lib.mrm_dev_gemm_lowp(mode, M, N, K, dA.data_ptr(), K * 8, dB.data_ptr(), N * 8, dC.data_ptr(), N * 8,
                               W.data_ptr(), wb, torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    return dC.cpu().numpy()


@pytest.mark.parametrize("mode", [_lib.SOLVE_FP32_IR, _lib.SOLVE_BF16_IR, _lib.SOLVE_BF16X3_IR])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 5), (300, 257, 256), (513, 260, 100)])
def test_lowp_update_exact_on_small_integers(mode, M, N, K):
    """Integers in [-8, 8] are exact in BF16 and FP32 (the low BF16 piece is 0)
    and their sums of products stay below 2^24, so C - A B must come out
    exactly."""
    rng = np.random.default_rng(M * N + K)
    A = rng.integers(-8, 9, (M, K)).astype(np.float64)
    B = rng.integers(-8, 9, (K, N)).astype(np.float64)
    C0 = rng.integers(-1000, 1000, (M, N)).astype(np.float64)
    assert np.array_equal(_lowp_update(mode, A, B, C0), C0 - orc.generic_mult(A, B))


@pytest.mark.parametrize("mode,u", [(_lib.SOLVE_FP32_IR, 2.0 ** -24), (_lib.SOLVE_BF16_IR, 2.0 ** -8),
                                    (_lib.SOLVE_BF16X3_IR, 2.0 ** -16)])
def test_lowp_update_rounding_bound(mode, u):
    """|error| <= (2u + K' 2^-24) |A||B| elementwise: u the rounding of the
    operands (for the split BF16 pieces: the dropped Al Bl and the rounding of
    the low pieces, each about 2^-18), K' the FP32 sum's length (3K for the
    split mode, whose three products run as one product over 3K)."""
    M, N, K = 700, 600, 256
    A = orc.pm1_matrix(K, M, 11)
    B = orc.pm1_matrix(N, K, 12)
    C0 = orc.pm1_matrix(N, M, 13)
    got = _lowp_update(mode, A, B, C0)
    ref = C0 - orc.generic_mult(A, B)
    kk = 3 * K if mode == _lib.SOLVE_BF16X3_IR else K
    bound = (2 * u + kk * 2.0 ** -24) * (np.abs(A) @ np.abs(B)) + 4 * 2.0 ** -53 * np.abs(C0)
    err = np.abs(got - ref)
    assert np.all(err <= bound), float(np.max(err / bound))
    # and it is a low-precision product, not an FP64 one
    assert np.max(err) > 1e-12


def _system(n, kappa, nrhs, seed):
    rng = np.random.default_rng(seed)
    Uq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Vq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(kappa), n)
    A = (Uq * s) @ Vq.T
    B = rng.standard_normal((n, nrhs))
    return np.ascontiguousarray(A), B


def _solve_ir(mode, A, B, max_iter=50, tol=1e-14):
    import torch
    lib = _lib.load()
    n, nrhs = B.shape
    dA, dB = _dev(A), _dev(B)
    dX = torch.zeros(n, nrhs, dtype=torch.float64, device="cuda")
    it, be, fb = C.c_int32(), C.c_double(), C.c_int32()
    rc = lib.mrm_dev_lineq_solve_ir(mode, dA.data_ptr(), n * 8, n, dB.data_ptr(), nrhs * 8, dX.data_ptr(), nrhs * 8,
                                    nrhs, max_iter, tol, C.byref(it), C.byref(be), C.byref(fb),
                                    torch.cuda.current_stream().cuda_stream)
    return rc, dX.cpu().numpy(), it.value, be.value, fb.value


@pytest.mark.parametrize("n,kappa,nrhs", [(257, 1e2, 3), (1000, 1e6, 8), (2500, 1e2, 64), (2500, 1e6, 64)])
@pytest.mark.parametrize("mode", [_lib.SOLVE_FP64, _lib.SOLVE_FP32_IR, _lib.SOLVE_BF16_IR, _lib.SOLVE_BF16X3_IR])
def test_lu_ir_reaches_fp64_backward_error(n, kappa, nrhs, mode):
    A, B = _system(n, kappa, nrhs, int(n + np.log10(kappa)))
    rc, X, it, be, fb = _solve_ir(mode, A, B)
    print(f"mode={mode} n={n} kappa={kappa:g}: rc={rc} steps={it} berr={be:.3e} fell_back={fb}")
    assert rc == 0
    be_oracle = orc.backward_error(A, X, B)
    assert be <= 1e-14 and be_oracle <= 1e-14, (be, be_oracle)
    if mode == _lib.SOLVE_FP32_IR:
        assert fb == 0  # FP32 updates suffice up to kappa 1e6
    rc, Xo = orc.lineq_solve(A, B, 0)
    rel = np.linalg.norm(X - Xo) / np.linalg.norm(Xo)
    assert rel <= kappa * 1e-13, rel


@pytest.mark.parametrize("n,boost_from", [(9000, None), (6000, 3000)])
def test_lu_ir_pivots_from_many_panel_workgroups(n, boost_from):
    """At these heights the panel kernel spreads its rows over many
    workgroups, and the pivots come from all of them (with boost_from, rows
    below it are 100x larger, so every pivot is far from the diagonal).
    Refinement of an accurate FP32-update factorisation converges in a few
    steps, without the FP64 fallback."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    if boost_from is not None:
        A[boost_from:] *= 100.0
    B = rng.standard_normal((n, 2))
    rc, X, it, be, fb = _solve_ir(_lib.SOLVE_FP32_IR, A, B)
    print(f"n={n} boost={boost_from}: rc={rc} steps={it} berr={be:.3e} fell_back={fb}")
    assert rc == 0 and fb == 0
    assert it <= 6
    assert be <= 1e-14 and orc.backward_error(A, X, B) <= 1e-14


def _solve_ir_stats(mode, A, B, max_iter=50, tol=1e-14):
    import torch
    lib = _lib.load()
    n, nrhs = B.shape
    dA, dB = _dev(A), _dev(B)
    dX = torch.zeros(n, nrhs, dtype=torch.float64, device="cuda")
    st = _lib.SolveStats()
    rc = lib.mrm_dev_lineq_solve_ir_stats(mode, dA.data_ptr(), n * 8, n, dB.data_ptr(), nrhs * 8, dX.data_ptr(),
                                          nrhs * 8, nrhs, max_iter, tol, C.byref(st),
                                          torch.cuda.current_stream().cuda_stream)
    return rc, dX.cpu().numpy(), st


@pytest.mark.parametrize("n,kappa,nrhs", [(1000, 1e2, 4), (2500, 1e2, 64)])
def test_gmres_ir_keeps_bf16_factors(n, kappa, nrhs):
    """BF16 trailing updates leave factors too inexact for plain refinement
    (kappa u_bf16 near 1 at these sizes): the plain step stalls, GMRES
    preconditioned by the same factors supplies the corrections, and the
    solve reaches a 1e-14 backward error without the FP64 fallback."""
    A, B = _system(n, kappa, nrhs, int(n + np.log10(kappa)))
    rc, X, st = _solve_ir_stats(_lib.SOLVE_BF16_IR, A, B)
    print(f"n={n} kappa={kappa:g}: rc={rc} steps={st.iterations} gmres={st.gmres} krylov={st.krylov_steps} "
          f"berr={st.backward_error:.3e} fell_back={st.fell_back}")
    assert rc == 0 and st.fell_back == 0
    assert st.gmres == 1 and st.krylov_steps > 0
    be = orc.backward_error(A, X, B)
    assert st.backward_error <= 1e-14 and be <= 1e-14, (st.backward_error, be)
    rc, Xo = orc.lineq_solve(A, B, 0)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= kappa * 1e-13


@pytest.mark.parametrize("mode,n,kappa,nrhs", [
    (_lib.SOLVE_FP32_IR, 1000, 1e2, 4),     # plain steps only
    (_lib.SOLVE_BF16_IR, 1000, 1e2, 4),     # the plain step stalls, GMRES finishes
    (_lib.SOLVE_BF16X3_IR, 1000, 1e6, 4),   # GMRES with split-BF16 factors
    (_lib.SOLVE_FP32_IR, 1000, 1e7, 2),     # GMRES with FP32 factors (kappa u_fp32 near 1)
    (_lib.SOLVE_BF16_IR, 700, 1e4, 3),      # the low-precision attempt fails: FP64 factors answer
])
def test_ir_path_matches_the_oracle(mode, n, kappa, nrhs):
    """The GPU solve takes the oracle's path (oracle/ir_gmres.py, pinned in
    tests/test_gmres_oracle.py): the same switch to GMRES and the same
    fallback, the same number of outer steps (+-1) and of Krylov steps (within
    a fifth, at least 2: the factors agree to rounding, not bit for bit), and
    the same solution to the conditioning."""
    from oracle import ir_gmres as ig
    A, B = _system(n, kappa, nrhs, int(n + np.log10(kappa)))
    rc, X, st = _solve_ir_stats(mode, A, B)
    rco, Xo, so = ig.solve_ir(A, B, mode)
    print(f"mode={mode} n={n} kappa={kappa:g}: GPU steps={st.iterations} gmres={st.gmres} krylov={st.krylov_steps} "
          f"fell_back={st.fell_back} (low-precision attempt {st.lowp_iterations} steps, {st.lowp_krylov_steps} "
          f"Krylov); oracle {so}")
    assert rc == 0 and rco == 0
    assert st.gmres == so["gmres"] and st.fell_back == so["fell_back"]
    assert abs(st.iterations - so["iterations"]) <= 1
    assert abs(st.krylov_steps - so["krylov_steps"]) <= max(2, so["krylov_steps"] // 5)
    if so["fell_back"]:
        assert st.lowp_iterations >= 1 and st.lowp_backward_error > 1e-14
    be = orc.backward_error(A, X, B)
    assert st.backward_error <= 1e-14 and be <= 1e-14
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= kappa * 1e-13


def test_gmres_ir_stats_without_a_stall():
    """A well-conditioned system with FP32 factors converges by plain
    refinement: no GMRES step, no fallback, and the statistics entry agrees
    with the three-output entry."""
    A, B = _system(600, 1e2, 3, 5)
    rc, X, st = _solve_ir_stats(_lib.SOLVE_FP32_IR, A, B)
    assert rc == 0 and st.gmres == 0 and st.krylov_steps == 0 and st.fell_back == 0
    rc2, X2, it, be, fb = _solve_ir(_lib.SOLVE_FP32_IR, A, B)
    assert rc2 == 0 and it == st.iterations and fb == 0
    assert np.array_equal(X, X2) and be == st.backward_error


def test_lu_ir_ill_conditioned_falls_back_or_converges():
    """kappa = 1e10 is beyond what FP32 factors can refine; the solve must still
    return a 1e-14 backward-error solution, through the FP64 fallback if needed."""
    A, B = _system(800, 1e10, 4, 99)
    for mode in (_lib.SOLVE_FP32_IR, _lib.SOLVE_BF16_IR):
        rc, X, st = _solve_ir_stats(mode, A, B)
        print(f"mode={mode} kappa=1e10: rc={rc} steps={st.iterations} gmres={st.gmres} krylov={st.krylov_steps} "
              f"berr={st.backward_error:.3e} fell_back={st.fell_back} (low-precision attempt: "
              f"{st.lowp_iterations} steps, {st.lowp_krylov_steps} Krylov, berr {st.lowp_backward_error:.3e})")
        assert rc == 0
        assert orc.backward_error(A, X, B) <= 1e-14


def test_lu_ir_singular_and_errors():
    lib = _lib.load()
    A = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 1.0, 1.0]])
    rc, X, it, be, fb = _solve_ir(_lib.SOLVE_FP32_IR, A, np.ones((3, 1)))
    assert rc == _lib.MRM_SINGULAR
    assert lib.mrm_dev_lineq_solve_ir(7, None, 8, 1, None, 8, None, 8, 1, 5, 1e-14, None, None, None, None) != 0
    # max_iter = 0 on an ill-conditioned system with BF16 factors cannot converge in FP64 either
    Ai, Bi = _system(300, 1e10, 2, 5)
    rc, X, it, be, fb = _solve_ir(_lib.SOLVE_BF16_IR, Ai, Bi, max_iter=0, tol=1e-30)
    assert rc == _lib.MRM_NOT_CONVERGED


@pytest.mark.parametrize("mode", [la.smFP32IR, la.smBF16IR])
def test_host_entry_with_mode(mode):
    A, B = _system(600, 1e3, 5, 17)
    X = np.zeros_like(B)
    assert la.MatrixLinEQSolveEx(A, B, X, 0, mode) == la.leOk
    assert orc.backward_error(A, X, B) <= 1e-14
    X2 = np.zeros_like(B)
    assert la.MatrixLinEQSolveEx(A, B, X2, 1, la.smFP64) == la.leOk
    # both are backward stable to 1e-14; forward agreement to kappa * 1e-13
    assert np.linalg.norm(X - X2) / np.linalg.norm(X2) <= 1e3 * 1e-13


def test_rows_per_thread_4_above_16384():
    """n = 16640: the panel kernel packs 4 rows per thread (17 workgroups) and
    the permutation is built in global memory; FP32-update LU-IR still reaches
    1e-14 with pivots from every workgroup (rows below 8000 are 100x larger)."""
    import torch
    lib = _lib.load()
    n = 16640
    g = torch.Generator(device="cuda").manual_seed(16640)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A[8000:] *= 100.0
    B = torch.randn(n, 4, dtype=torch.float64, device="cuda", generator=g)
    X = torch.zeros(n, 4, dtype=torch.float64, device="cuda")
    it, be, fb = C.c_int32(), C.c_double(), C.c_int32()
    rc = lib.mrm_dev_lineq_solve_ir(_lib.SOLVE_FP32_IR, A.data_ptr(), n * 8, n, B.data_ptr(), 32, X.data_ptr(), 32, 4,
                                    50, 1e-14, C.byref(it), C.byref(be), C.byref(fb),
                                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    print(f"n={n}: rc={rc} steps={it.value} berr={be.value:.3e} fell_back={fb.value}")
    assert rc == 0 and fb.value == 0
    R = B - A @ X
    berr = (R.abs().amax(dim=0) / (A.abs().sum(dim=1).amax() * X.abs().amax(dim=0) + B.abs().amax(dim=0))).max()
    assert float(berr) <= 1e-14


@pytest.mark.parametrize("mode", [_lib.SOLVE_FP32_IR, _lib.SOLVE_BF16X3_IR])
def test_lookahead_lowp_matches_the_sequential_schedule(mode):
    """n = 6144 with low-precision trailing updates: the lookahead (the next
    block's update on the update stream, the trailing update beside the panel
    chain) and the sequential schedule (the cooperative relaunch, forced from
    the host with deadline 0) give the same solution and statistics bit for
    bit."""
    import torch
    lib = _lib.load()
    assert lib.mrm_lu_lookahead_available() == 1
    n = 6144
    g = torch.Generator(device="cuda").manual_seed(6144 + mode)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(n, 3, dtype=torch.float64, device="cuda", generator=g)
    s_ = torch.cuda.current_stream().cuda_stream

    def solve():
        X = torch.zeros(n, 3, dtype=torch.float64, device="cuda")
        st = _lib.SolveStats()
        rc = lib.mrm_dev_lineq_solve_ir_stats(mode, A.data_ptr(), n * 8, n, B.data_ptr(), 24, X.data_ptr(), 24, 3,
                                              50, 1e-14, C.byref(st), s_)
        torch.cuda.synchronize()
        return rc, X, (st.iterations, st.gmres, st.krylov_steps, st.fell_back, st.backward_error)

    before = lib.mrm_lu_panel_test_deadline(-1)
    rc0, X0, st0 = solve()
    assert rc0 == 0 and lib.mrm_lu_panel_test_deadline(-1) == before
    try:
        lib.mrm_lu_panel_test_deadline(0)
        rc1, X1, st1 = solve()
    finally:
        lib.mrm_lu_panel_test_deadline(-1)
    assert rc1 == 0 and lib.mrm_lu_panel_test_deadline(-1) == before + 1
    print(f"mode={mode}: {st0} / {st1}")
    assert torch.equal(X0, X1) and st0 == st1
    assert st0[4] <= 1e-14


@pytest.fixture(scope="module")
def system_16384():
    """A = U diag(s) V^T (U, V from Householder QR of Gaussian matrices, s
    log-spaced to kappa = 1e2) and B Gaussian, n = 16384, 64 columns."""
    import torch
    n, nrhs = 16384, 64
    g = torch.Generator(device="cuda").manual_seed(4401)
    Uq, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    Vq, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    B = torch.randn(n, nrhs, dtype=torch.float64, device="cuda", generator=g)
    s = torch.logspace(0, -2, n, dtype=torch.float64, device="cuda")
    A = (Uq * s) @ Vq.t()
    del Uq, Vq
    yield A, B
    del A, B
    torch.cuda.empty_cache()


@pytest.mark.parametrize("mode", [_lib.SOLVE_FP64, _lib.SOLVE_FP32_IR, _lib.SOLVE_BF16_IR, _lib.SOLVE_BF16X3_IR])
def test_n16384_64rhs_independent_residual(mode, system_16384):
    """The LU-IR workload itself (n = 16384, 64 right-hand sides, kappa = 1e2):
    the backward error is recomputed here from an independent FP64 residual
    (torch's matmul), not taken from the library's statistics."""
    import torch
    lib = _lib.load()
    A, B = system_16384
    n, nrhs = B.shape
    X = torch.zeros(n, nrhs, dtype=torch.float64, device="cuda")
    it, be, fb = C.c_int32(), C.c_double(), C.c_int32()
    rc = lib.mrm_dev_lineq_solve_ir(mode, A.data_ptr(), n * 8, n, B.data_ptr(), nrhs * 8, X.data_ptr(), nrhs * 8,
                                    nrhs, 50, 1e-14, C.byref(it), C.byref(be), C.byref(fb),
                                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == 0
    R = B - A @ X
    berr = float((R.abs().amax(dim=0) / (A.abs().sum(dim=1).amax() * X.abs().amax(dim=0) +
                                         B.abs().amax(dim=0))).max())
    print(f"mode={mode}: steps={it.value} library berr={be.value:.3e} independent berr={berr:.3e} "
          f"fell_back={fb.value}")
    assert berr <= 1e-14
    assert berr <= 4 * be.value + 1e-17
