"""GPU parity of the batched plans (mrm_batched_dgemm_plan / mrm_batched_lu_plan
/ mrm_batched_run): every product against the oracle's GenericMtxMult to the
GEMM bar ||C - Cref||_F / (||A||_F ||B||_F) <= 4 k 2^-53, every solve to the
normwise backward error 1e-14 (oracle, long double residual), repeated graph
replays identical, singular systems reported per system.
"""
import numpy as np
import pytest

import oracle as orc
from mrmath_amd import _lib
from mrmath_amd import batched as bt

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def _gemm_batch(sizes, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    probs = []
    for m, n, k in sizes:
        A = torch.rand(m, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        B = torch.rand(k, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        Cm = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")
        probs.append((A, B, Cm))
    return probs


def _check_gemm(probs, idx):
    for i in idx:
        A, B, Cm = (t.cpu().numpy() for t in probs[i])
        if A.size == 0 or B.size == 0:  # empty inner dimension: C = 0
            assert np.all(Cm == 0.0), i
            continue
        ref = orc.generic_mult(A, B)
        assert orc.gemm_err(Cm, ref, A, B) <= 4 * A.shape[1] * U, i


@pytest.mark.parametrize("tile", [0, 32, 64, 128])
def test_batched_gemm_uneven(tile):
    rng = np.random.default_rng(17)
    sizes = [tuple(int(v) for v in rng.integers(1, 257, 3)) for _ in range(150)]
    sizes += [(1, 1, 1), (16, 16, 16), (255, 257, 3), (256, 256, 256), (3, 200, 1), (40, 1, 129), (7, 9, 0)]
    probs = _gemm_batch(sizes, 5)
    import torch
    plan = bt.dgemm_plan(probs, tile)
    assert plan.tile in (32, 64, 128) and (tile == 0 or plan.tile == tile)
    plan.run()
    torch.cuda.synchronize()
    _check_gemm(probs, range(len(probs)))
    first = [p[2].clone() for p in probs]
    for rep in range(3):  # replays with eager work in between give the same bits
        for p in probs:
            p[2].fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run()
        torch.cuda.synchronize()
        assert all(torch.equal(p[2], f) for p, f in zip(probs, first))
    plan.close()


def test_batched_gemm_full_size():
    """8192 problems, m, n, k uniform in [16, 256] (seed 5101): a sample of 96
    against the oracle, every problem through C x == A (B x)."""
    import torch
    rng = np.random.default_rng(5101)
    sizes = [tuple(int(v) for v in rng.integers(16, 257, 3)) for _ in range(8192)]
    probs = _gemm_batch(sizes, 5101)
    plan = bt.dgemm_plan(probs, 0)
    plan.run()
    torch.cuda.synchronize()
    _check_gemm(probs, rng.choice(len(probs), 96, replace=False))
    worst = 0.0
    for A, B, Cm in probs:
        x = torch.ones(B.shape[1], dtype=torch.float64, device="cuda")
        r = torch.linalg.vector_norm(Cm @ x - A @ (B @ x)) / (torch.linalg.vector_norm(A) *
                                                              torch.linalg.vector_norm(B) *
                                                              torch.linalg.vector_norm(x))
        worst = max(worst, float(r))
    assert worst <= 4 * 256 * U, worst
    plan.close()


def _split_exact_batch(sizes, seed):
    """Operands int + frac 2^-30 with 16 <= |int| <= 64 and |frac| <= 64: the
    FP32 pieces are exactly the integer (frac 2^-30 is below half its FP32
    ulp) and the fraction, every FP32 chunk sum is exact, so the compensated
    product equals int*int + 2^-30 (int*frac + frac*int) exactly (only
    frac*frac 2^-60 is dropped)."""
    import torch
    rng = np.random.default_rng(seed)

    def ints(shape):
        return rng.integers(16, 65, shape) * rng.choice([-1, 1], shape)

    probs, want = [], []
    for m, n, k in sizes:
        ia, fa = ints((m, k)), rng.integers(-64, 65, (m, k))
        ib, fb = ints((k, n)), rng.integers(-64, 65, (k, n))
        A = ia + fa * 2.0 ** -30
        B = ib + fb * 2.0 ** -30
        X = (ia @ ib) * (1 << 30) + (ia @ fb + fa @ ib)   # exact in int64 (< 2^53)
        want.append(X.astype(np.float64) * 2.0 ** -30)
        probs.append((torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                      torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")))
    return probs, want


@pytest.mark.parametrize("tile", [0, 64, 128])
def test_batched_gemm_fp32_compensated_exact_pieces(tile):
    """The FP32 mode forms Ah Bh + (Ah Bl + Al Bh): on operands whose pieces
    and chunk sums are exact in FP32 it must return exactly that (a dropped
    or mis-signed correction term differs by 2^-30 times an integer)."""
    import torch
    rng = np.random.default_rng(41)
    sizes = [tuple(int(v) for v in rng.integers(1, 200, 3)) for _ in range(40)] + [(128, 128, 256), (1, 1, 1)]
    probs, want = _split_exact_batch(sizes, 43)
    plan = bt.dgemm_plan(probs, tile, _lib.SOLVE_FP32_IR)
    assert plan.tile in (64, 128) and (tile == 0 or plan.tile == tile)
    plan.run()
    torch.cuda.synchronize()
    for (A, B, Cm), w in zip(probs, want):
        assert np.array_equal(Cm.cpu().numpy(), w)
    plan.close()


def test_batched_gemm_fp32_compensated_error():
    """Random operands: every element within the bound of FP32 sums over K
    chunks of 16, |C - A B| <= 17 2^-24 (|A||B|) (the dropped Al Bl and the
    FP64 accumulation are far below it); the normwise error against the
    oracle's FP64 GEMM is reported.  Graph replays are bit-identical."""
    import torch
    rng = np.random.default_rng(53)
    sizes = [tuple(int(v) for v in rng.integers(16, 257, 3)) for _ in range(120)]
    probs = _gemm_batch(sizes, 53)
    plan = bt.dgemm_plan(probs, 0, _lib.SOLVE_FP32_IR)
    plan.run()
    torch.cuda.synchronize()
    worst = 0.0
    for A, B, Cm in probs:
        a, b, c = (t.cpu().numpy() for t in (A, B, Cm))
        ref = orc.generic_mult(a, b)
        bound = 17 * 2.0 ** -24 * (np.abs(a) @ np.abs(b))
        assert np.all(np.abs(c - ref) <= bound)
        worst = max(worst, orc.gemm_err(c, ref, a, b))
    print(f"FP32 compensated batched GEMM: worst normwise error {worst:.3e} (FP64 bar {4 * 256 * U:.3e})")
    assert worst > 4 * 256 * U * 1e-6  # a low-precision result, not an FP64 one
    first = [p[2].clone() for p in probs]
    for rep in range(2):
        for p in probs:
            p[2].fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run()
        torch.cuda.synchronize()
        assert all(torch.equal(p[2], f) for p, f in zip(probs, first))
    plan.close()


def _lu_batch(ns, seed, singular=()):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    systems = []
    for i, n in enumerate(ns):
        A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g) + n * torch.eye(
            n, dtype=torch.float64, device="cuda")
        if i in singular:
            A[:, 1] = 0.0  # exactly singular: column 1 stays zero through the elimination
        b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
        x = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        systems.append((A, b, x))
    return systems


@pytest.mark.parametrize("mode", [_lib.SOLVE_FP64, _lib.SOLVE_FP32_IR])
def test_batched_lu_solve(mode):
    import torch
    rng = np.random.default_rng(23)
    ns = [1, 2, 8, 31, 32, 33, 63, 64, 65, 127, 128] + [int(v) for v in rng.integers(1, 129, 200)]
    systems = _lu_batch(ns, 7, singular={5, 9})
    status = torch.full((len(systems),), -1, dtype=torch.int32, device="cuda")
    plan = bt.lu_plan(systems, mode, status)
    plan.run()
    torch.cuda.synchronize()
    st = status.cpu().numpy()
    for i, (A, b, x) in enumerate(systems):
        if i in (5, 9):
            assert st[i] == _lib.MRM_SINGULAR
            continue
        assert st[i] == _lib.MRM_OK, i
        An, bn, xn = A.cpu().numpy(), b.cpu().numpy().reshape(-1, 1), x.cpu().numpy().reshape(-1, 1)
        assert orc.backward_error(An, xn, bn) <= 1e-14, (i, orc.backward_error(An, xn, bn))
        rc, xo = orc.lineq_solve(An, bn, 0)
        assert np.allclose(xn, xo, rtol=1e-12, atol=1e-13), i
    plan.close()


def test_batched_lu_full_size():
    """20000 systems, n uniform in [8, 128], Gaussian + n I (seed 5201), both
    precisions; every status OK, a sample of 150 checked by the oracle."""
    import torch
    rng = np.random.default_rng(5201)
    ns = [int(v) for v in rng.integers(8, 129, 20000)]
    systems = _lu_batch(ns, 5201)
    for mode in (_lib.SOLVE_FP64, _lib.SOLVE_FP32_IR):
        status = torch.full((len(systems),), -1, dtype=torch.int32, device="cuda")
        plan = bt.lu_plan(systems, mode, status)
        plan.run()
        torch.cuda.synchronize()
        assert np.all(status.cpu().numpy() == 0)
        for i in rng.choice(len(systems), 150, replace=False):
            A, b, x = systems[i]
            be = orc.backward_error(A.cpu().numpy(), x.cpu().numpy().reshape(-1, 1), b.cpu().numpy().reshape(-1, 1))
            assert be <= 1e-14, (mode, i, be)
        plan.close()


def test_batched_errors():
    import ctypes as C
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.mrm_batched_dgemm_plan(None, 1, 0, C.byref(h)) == _lib.MRM_BAD_ARG
    assert lib.mrm_batched_dgemm_plan(None, 0, 48, C.byref(h)) == _lib.MRM_BAD_ARG
    assert lib.mrm_batched_dgemm_plan_ex(None, 0, 32, _lib.SOLVE_FP32_IR, C.byref(h)) == _lib.MRM_BAD_ARG
    assert lib.mrm_batched_dgemm_plan_ex(None, 0, 64, _lib.SOLVE_BF16_IR, C.byref(h)) == _lib.MRM_BAD_ARG
    d = (bt.LuDesc * 1)(bt.LuDesc(129, 1, 129 * 8, 1, 1))
    assert lib.mrm_batched_lu_plan(d, 1, 0, None, C.byref(h)) == _lib.MRM_BAD_DIMS
    assert lib.mrm_batched_lu_plan(d, 1, 2, None, C.byref(h)) == _lib.MRM_BAD_ARG
    assert lib.mrm_batched_run(None, None) == _lib.MRM_BAD_ARG
