"""GPU parity of the LU factorisation and linear solve against the oracle.

Bars: the reference's known-answer tests bit for bit; on random inputs the
pivot sequence (indx) must equal the oracle's exactly and the factors must
agree to 64*n*2^-53 * max|A| element-wise; solves must reach a normwise
backward error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
<= 1e-14.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import _pins
import oracle as orc
from mrmath_amd import _lib
from mrmath_amd import linalg as la

pytestmark = pytest.mark.gpu

KAT = json.loads((Path(__file__).resolve().parent / "golden" / "kat.json").read_text())
U = 2.0 ** -53


def test_lu_kat():
    k = KAT["lu"]
    A = np.array(k["A"], dtype=np.float64).reshape(2, 2)
    indx = np.zeros(2, dtype=np.int32)
    assert la.MatrixLUDecompInPlace(A, indx) == la.leOk
    assert np.array_equal(A.ravel(), k["LU"])
    assert list(indx) == k["indx"]
    b = np.array(k["b"], dtype=np.float64).reshape(2, 1)
    la.MatrixLUBackSubst(A, indx, b)
    assert np.array_equal(b.ravel(), k["x"])


def test_lu_kat_strided():
    s = KAT["lu"]["strided"]
    raw = np.array(s["A_lw3"], dtype=np.float64)
    A = np.lib.stride_tricks.as_strided(raw, (2, 2), (24, 8))
    indx = np.zeros(2, dtype=np.int32)
    assert la.MatrixLUDecompInPlace(A, indx) == la.leOk
    assert np.array_equal(raw, s["LU_lw3"])
    Braw = np.array(s["B_lw2"], dtype=np.float64)
    la.MatrixLUBackSubst(A, indx, Braw.reshape(2, 2))
    assert np.array_equal(Braw, s["X_lw2"])


def test_singular():
    A = np.array(KAT["singular"]["A"], dtype=np.float64).reshape(3, 3)
    indx = np.zeros(3, dtype=np.int32)
    assert la.MatrixLUDecompInPlace(A.copy(), indx) == la.leSingular
    X = np.zeros((3, 1))
    assert la.MatrixLinEQSolve(A, np.ones((3, 1)), X) == la.leSingular
    with pytest.raises(la.ELinEQSingularException):
        la.SolveLinEQ(A, np.ones((3, 1)))
    Z = np.zeros((40, 40))
    assert la.MatrixLUDecompInPlace(Z, np.zeros(40, dtype=np.int32)) == la.leSingular


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 64, 100, 255, 256, 257, 300, 520, 1000])
def test_lu_matches_oracle(n):
    rng = np.random.default_rng(7000 + n)
    A0 = rng.standard_normal((n, n))
    Ag = A0.copy()
    ig = np.zeros(n, dtype=np.int32)
    assert la.MatrixLUDecompInPlace(Ag, ig) == la.leOk
    Ao = A0.copy()
    rc, io = orc.lu_decomp_inplace(Ao)
    assert rc == 0
    assert np.array_equal(ig, io), "pivot sequence differs from the oracle"
    assert np.max(np.abs(Ag - Ao)) <= 64 * n * U * np.max(np.abs(A0)) * 10
    # P*A = L*U
    L = np.tril(Ag, -1) + np.eye(n)
    Um = np.triu(Ag)
    perm = np.arange(n)
    for i, r in enumerate(ig):
        perm[[i, r]] = perm[[r, i]]
    assert np.max(np.abs(A0[perm] - L @ Um)) <= 16 * n * U * np.max(np.abs(A0)) * 10


def test_lu_padded_linewidth():
    n = 150
    rng = np.random.default_rng(3)
    raw = np.full((n, n + 5), 9.0)
    A = raw[:, :n]
    A[...] = rng.standard_normal((n, n))
    A0 = A.copy()
    ig = np.zeros(n, dtype=np.int32)
    assert la.MatrixLUDecompInPlace(A, ig) == la.leOk
    assert np.all(raw[:, n:] == 9.0)
    _, io = orc.lu_decomp_inplace(A0)
    assert np.array_equal(ig, io)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (5, 1), (48, 3), (300, 4), (777, 2), (1024, 8)])
def test_lineq_solve_backward_error(n, nrhs):
    A = orc.pm1_matrix(n, n, 4400 + n)
    B = orc.pm1_matrix(nrhs, n, 4500 + n)
    X = np.zeros((n, nrhs))
    assert la.MatrixLinEQSolve(A, B, X, 0) == la.leOk
    be = orc.backward_error(A, X, B)
    assert be <= 1e-14, be
    rc, Xo = orc.lineq_solve(A, B, 0)
    # both are backward stable solutions of the same system
    assert np.allclose(X, Xo, rtol=1e-8, atol=1e-10)


@pytest.fixture(scope="module")
def wilkinson_system():
    n = 50
    A = _pins.wilkinson(n)
    B = orc.pm1_matrix(3, n, 7)
    Xe = np.stack([_pins.exact_solve(A, B[:, j]) for j in range(3)], axis=1)
    return A, B, Xe


def test_lineq_solve_refinement_pinned(wilkinson_system):
    """NumRefinments on the GPU (src/LinAlgLU.pas:542-586) against the exact
    rational solution of the n = 50 Wilkinson system (pivot growth 2^49):
    unrefined the forward error is ~1e-3, one FP64 step brings it to the
    rounding level, more steps keep it there; 0 and 1 refinements differ.  A
    refinement that does nothing or has the wrong sign fails."""
    A, B, Xe = wilkinson_system
    X = [np.zeros_like(B) for _ in range(3)]
    for k in range(3):
        assert la.MatrixLinEQSolve(A, B, X[k], k) == la.leOk
    errs = [_pins.forward_error(x, Xe) for x in X]
    print("forward errors by refinement steps:", errs)
    assert errs[0] > 1e-5
    assert errs[1] <= 2e-15 and errs[2] <= 2e-15
    assert np.abs(X[1] - X[0]).max() > 1e3 * np.abs(X[1] - Xe).max()
    assert orc.backward_error(A, X[1], B) < 1e-16
    # single right-hand side in the reference's layout (LineWidthX = 8)
    x1 = np.zeros((50, 1))
    assert la.MatrixLinEQSolve(A, np.ascontiguousarray(B[:, :1]), x1, 1) == la.leOk
    assert _pins.forward_error(x1, Xe[:, :1]) <= 2e-15


def test_refinement_entry_steps_pinned(wilkinson_system):
    """The device refinement driver on the same system: with no step allowed
    it reports the unrefined backward error and MRM_NOT_CONVERGED; allowed
    steps converge in one or two, to the exact solution's rounding level."""
    import ctypes as C
    import torch
    lib = _lib.load()
    A, B, Xe = wilkinson_system
    n, nrhs = B.shape
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    s_ = torch.cuda.current_stream().cuda_stream
    out = {}
    for max_iter in (0, 5):
        dX = torch.zeros(n, nrhs, dtype=torch.float64, device="cuda")
        it, be, fb = C.c_int32(), C.c_double(), C.c_int32()
        rc = lib.mrm_dev_lineq_solve_ir(_lib.SOLVE_FP64, dA.data_ptr(), n * 8, n, dB.data_ptr(), nrhs * 8,
                                        dX.data_ptr(), nrhs * 8, nrhs, max_iter, 1e-15, C.byref(it), C.byref(be),
                                        C.byref(fb), s_)
        torch.cuda.synchronize()
        out[max_iter] = (rc, it.value, be.value, dX.cpu().numpy())
    rc0, it0, be0, X0 = out[0]
    assert rc0 == _lib.MRM_NOT_CONVERGED and it0 == 0 and be0 > 1e-7
    assert be0 == pytest.approx(orc.backward_error(A, X0, B), rel=1e-6)
    rc5, it5, be5, X5 = out[5]
    assert rc5 == _lib.MRM_OK and 1 <= it5 <= 2 and be5 <= 1e-15
    assert _pins.forward_error(X5, Xe) <= 2e-15


def test_nan_right_hand_side_is_not_converged():
    """A NaN in b reaches the convergence test (the residual norms carry it):
    the solve reports MRM_NOT_CONVERGED with a NaN backward error instead of
    a converged NaN solution."""
    import ctypes as C
    import torch
    lib = _lib.load()
    n = 300
    A = torch.from_numpy(orc.pm1_matrix(n, n, 3) + n * np.eye(n)).cuda()
    Bh = orc.pm1_matrix(2, n, 4)
    Bh[17, 1] = np.nan
    B = torch.from_numpy(Bh).cuda()
    s_ = torch.cuda.current_stream().cuda_stream
    for mode in (_lib.SOLVE_FP64, _lib.SOLVE_FP32_IR):
        X = torch.zeros(n, 2, dtype=torch.float64, device="cuda")
        it, be, fb = C.c_int32(), C.c_double(), C.c_int32()
        rc = lib.mrm_dev_lineq_solve_ir(mode, A.data_ptr(), n * 8, n, B.data_ptr(), 16, X.data_ptr(), 16, 2, 5,
                                        1e-14, C.byref(it), C.byref(be), C.byref(fb), s_)
        assert rc == _lib.MRM_NOT_CONVERGED
        assert np.isnan(be.value)
        if mode == _lib.SOLVE_FP32_IR:
            assert fb.value == 1


def test_panel_timeout_relaunches_cooperatively():
    """A panel whose workgroups are not co-resident times out instead of
    hanging; the factorisation then runs again from the original matrix with
    cooperative panel launches.  Forced from the host: deadline 0 makes every
    wait of a plain launch give up at once (every kernel still exits
    normally).  Results equal an undisturbed run bit for bit."""
    import torch
    lib = _lib.load()
    n = 2048
    g = torch.Generator(device="cuda").manual_seed(77)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    s_ = torch.cuda.current_stream().cuda_stream

    def dev_factor():
        LU = A.clone()
        indx = torch.zeros(n, dtype=torch.int32, device="cuda")
        perm = torch.zeros(n, dtype=torch.int32, device="cuda")
        rc = lib.mrm_dev_lu_factor(LU.data_ptr(), n * 8, n, indx.data_ptr(), perm.data_ptr(), s_)
        torch.cuda.synchronize()
        return rc, LU, indx, perm

    rc, LU0, indx0, perm0 = dev_factor()
    assert rc == 0
    Ah = orc.pm1_matrix(300, 300, 5)
    lu_h0 = Ah.copy()
    ih0 = np.zeros(300, dtype=np.int32)
    assert la.MatrixLUDecompInPlace(lu_h0, ih0) == la.leOk
    before = lib.mrm_lu_panel_test_deadline(-1)
    try:
        lib.mrm_lu_panel_test_deadline(0)
        rc, LU1, indx1, perm1 = dev_factor()
        assert rc == 0
        assert lib.mrm_lu_panel_test_deadline(0) == before + 1
        assert torch.equal(LU0, LU1) and torch.equal(indx0, indx1) and torch.equal(perm0, perm1)
        # host entry: restaged from the caller's matrix
        lu_h1 = Ah.copy()
        ih1 = np.zeros(300, dtype=np.int32)
        assert la.MatrixLUDecompInPlace(lu_h1, ih1) == la.leOk
        assert lib.mrm_lu_panel_test_deadline(0) == before + 2
        assert np.array_equal(lu_h0, lu_h1) and np.array_equal(ih0, ih1)
        # solve entry: refactors its copy of A
        B = orc.pm1_matrix(2, 300, 6)
        X = np.zeros((300, 2))
        assert la.MatrixLinEQSolve(Ah, B, X, 1) == la.leOk
        assert orc.backward_error(Ah, X, B) <= 1e-14
        assert lib.mrm_lu_panel_test_deadline(0) == before + 3
    finally:
        lib.mrm_lu_panel_test_deadline(-1)
    # the library stays usable with the default deadline: no more relaunches
    rc, LU2, _, _ = dev_factor()
    assert rc == 0 and torch.equal(LU0, LU2)
    assert lib.mrm_lu_panel_test_deadline(-1) == before + 3


@pytest.mark.parametrize("n", [6144, 7000])
def test_lookahead_matches_the_sequential_schedule(n):
    """From n = 6144 on, outer block k + 1 is updated by block k and factored
    on its own CUs while the other CUs apply block k to the columns beyond
    block k + 1.  The sequential schedule (taken by the cooperative relaunch,
    forced from the host with deadline 0) must give the same pivots and
    factors equal to rounding: the two schedules differ only in how the
    trailing updates are split by column.  n = 7000 ends in a partial outer
    block."""
    import torch
    lib = _lib.load()
    g = torch.Generator(device="cuda").manual_seed(n + 1)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A[n // 3:] *= 10.0  # pivots from the lower rows first
    s_ = torch.cuda.current_stream().cuda_stream

    def dev_factor():
        LU = A.clone()
        indx = torch.zeros(n, dtype=torch.int32, device="cuda")
        perm = torch.zeros(n, dtype=torch.int32, device="cuda")
        rc = lib.mrm_dev_lu_factor(LU.data_ptr(), n * 8, n, indx.data_ptr(), perm.data_ptr(), s_)
        torch.cuda.synchronize()
        return rc, LU, indx, perm

    # the MI355X (256 CUs in 8 XCDs) has the lookahead's streams
    assert lib.mrm_lu_lookahead_available() == 1
    before = lib.mrm_lu_panel_test_deadline(-1)
    rc, LU0, indx0, perm0 = dev_factor()
    assert rc == 0, lib.mrm_last_error_message()
    assert lib.mrm_lu_panel_test_deadline(-1) == before
    try:
        lib.mrm_lu_panel_test_deadline(0)
        rc, LU1, indx1, perm1 = dev_factor()
    finally:
        lib.mrm_lu_panel_test_deadline(-1)
    assert rc == 0 and lib.mrm_lu_panel_test_deadline(-1) == before + 1
    assert torch.equal(indx0, indx1) and torch.equal(perm0, perm1)
    d = float((LU0 - LU1).abs().max())
    print(f"n={n}: max |lookahead - sequential| = {d:.3e}, bit-equal: {torch.equal(LU0, LU1)}")
    assert d <= 64 * n * U * float(A.abs().max())
    # and the factors reproduce P A
    cols = torch.arange(0, n, 97, device="cuda")
    L = torch.tril(LU0, -1) + torch.eye(n, dtype=torch.float64, device="cuda")
    r = A[perm0.long()][:, cols] - L @ torch.triu(LU0)[:, cols]
    assert float(r.abs().max()) <= 64 * n * U * float(A.abs().max())


def test_lookahead_after_shutdown_and_restart():
    """mrm_shutdown frees the lookahead's streams with the rest of the
    library; the next call sets them up again and the factorisation repeats
    bit for bit."""
    import torch
    lib = _lib.load()
    n = 6144
    g = torch.Generator(device="cuda").manual_seed(61)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    s_ = torch.cuda.current_stream().cuda_stream
    out = []
    for _ in range(2):
        assert lib.mrm_lu_lookahead_available() == 1
        LU = A.clone()
        indx = torch.zeros(n, dtype=torch.int32, device="cuda")
        assert lib.mrm_dev_lu_factor(LU.data_ptr(), n * 8, n, indx.data_ptr(), 0, s_) == 0
        torch.cuda.synchronize()
        out.append((LU, indx))
        lib.mrm_shutdown()
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


def test_backsubst_single_column_matches_oracle():
    n = 200
    A = orc.pm1_matrix(n, n, 91)
    lu = A.copy()
    indx = np.zeros(n, dtype=np.int32)
    assert la.MatrixLUDecompInPlace(lu, indx) == la.leOk
    b = orc.pm1_matrix(1, n, 92)
    bo = b.copy()
    la.MatrixLUBackSubst(lu, indx, b)
    orc.lu_backsubst(lu, indx, bo)
    assert np.allclose(b, bo, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("n", [2048, 5000])
def test_device_lu_and_solve(n):
    """Device entries on torch tensors: pivots equal the oracle's (n=2048 only,
    the oracle's recursion is O(n^3) on one core), backward error <= 1e-14 with
    64 right-hand sides."""
    import torch
    lib = _lib.load()
    g = torch.Generator(device="cuda").manual_seed(n)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(n, 64, dtype=torch.float64, device="cuda", generator=g)
    LU = A.clone()
    indx = torch.zeros(n, dtype=torch.int32, device="cuda")
    perm = torch.zeros(n, dtype=torch.int32, device="cuda")
    X = torch.zeros(n, 64, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    assert lib.mrm_dev_lu_factor(LU.data_ptr(), n * 8, n, indx.data_ptr(), perm.data_ptr(), s) == 0
    assert lib.mrm_dev_lu_solve(LU.data_ptr(), n * 8, n, perm.data_ptr(), B.data_ptr(), 64 * 8, X.data_ptr(), 64 * 8,
                                64, s) == 0
    torch.cuda.synchronize()
    An, Bn, Xn = A.cpu().numpy(), B.cpu().numpy(), X.cpu().numpy()
    # perm agrees with indx
    p = np.arange(n)
    for i, r in enumerate(indx.cpu().numpy()):
        p[[i, r]] = p[[r, i]]
    assert np.array_equal(p, perm.cpu().numpy())
    r = Bn - An @ Xn
    be = np.max(np.abs(r).max(axis=0) / (np.abs(An).sum(axis=1).max() * np.abs(Xn).max(axis=0) + np.abs(Bn).max(axis=0)))
    assert be <= 1e-14, be
    if n == 2048:
        A0 = An.copy()
        rc, io = orc.lu_decomp_inplace(A0)
        assert np.array_equal(indx.cpu().numpy(), io)


def test_inverse_and_determinant_kats():
    """Test/TestLineEquations.pas:635-642 and :702-713, bit for bit."""
    A = np.array(KAT["det"]["A"], dtype=np.float64).reshape(2, 2)
    assert la.MatrixDeterminant(A) == KAT["det"]["det"]
    Ai = np.array(KAT["inverse"]["A"], dtype=np.float64).reshape(2, 2)
    assert la.MatrixInverseInPlace(Ai) == la.leOk
    assert np.array_equal(Ai.ravel(), KAT["inverse"]["inv"])
    S = np.array(KAT["singular"]["A"], dtype=np.float64).reshape(3, 3)
    S0 = S.copy()
    assert la.MatrixInverseInPlace(S) == la.leSingular
    assert np.array_equal(S, S0)
    assert la.MatrixDeterminant(S0) == 0.0


@pytest.mark.parametrize("n", [1, 7, 64, 128, 129, 300])
def test_inverse_and_determinant_vs_oracle(n):
    rng = np.random.default_rng(900 + n)
    A = rng.standard_normal((n, n)) + 0.5 * n ** 0.5 * np.eye(n)
    Ag, Ao = A.copy(), A.copy()
    assert la.MatrixInverseInPlace(Ag) == la.leOk
    assert orc.inverse_inplace(Ao) == 0
    # the factors round differently (blocked vs recursive updates)
    assert np.max(np.abs(Ag - Ao)) <= 1e-10 * np.max(np.abs(Ao))
    assert np.max(np.abs(A @ Ag - np.eye(n))) <= 1e-10
    dg, do = la.MatrixDeterminant(A), orc.determinant(A.copy())
    assert dg == do or abs(dg - do) <= 1e-12 * abs(do)
