"""GPU unit checks of the dgetrf building blocks through the C ABI test hooks:
the FP64 matrix-core GEMM update and the unit-lower triangular solve, against
numpy in FP64."""
import numpy as np
import pytest

from dust_amd import _lib as L

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    # column-major matrix a (m, n) -> device buffer holding a.T rows
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (64, 64, 16), (136, 136, 64), (100, 37, 64),
                                   (1, 1, 64), (257, 129, 33)])
def test_gemm_sub(M, N, K):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(M * N + K)
    A, B, Cm = rng.standard_normal((M, K)), rng.standard_normal((K, N)), rng.standard_normal((M, N))
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, Cm)
    print(f"gemm {M}x{N}x{K}", flush=True)
    L.check(L.lib().dust_test_gemm_sub_dev(M, N, K, dA.data_ptr(), M, dB.data_ptr(), K,
                                           dC.data_ptr(), M), "gemm")
    got = dC.cpu().numpy().T
    ref = Cm - A @ B
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("m,n", [(1, 5), (64, 64), (64, 100), (130, 70)])
def test_trsm_llnu(m, n):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(m + n)
    Lm = np.tril(rng.standard_normal((m, m)), -1) * 0.3 + np.eye(m)
    B = rng.standard_normal((m, n))
    dL, dB = _dev(torch, Lm), _dev(torch, B)
    L.check(L.lib().dust_test_trsm_llnu_dev(m, n, dL.data_ptr(), m, dB.data_ptr(), m), "trsm")
    got = dB.cpu().numpy().T
    ref = np.linalg.solve(Lm, B)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("m,n", [(2000, 64), (300, 64), (129, 64), (1000, 20)])
def test_panel_tall(m, n):
    """dgetrf on a single panel (n <= 64) over many row blocks."""
    import scipy.linalg.lapack as lapack
    from dust_amd import linsys as LS
    rng = np.random.default_rng(m)
    A0 = rng.standard_normal((m, n))
    A = np.asfortranarray(A0.copy())
    ipiv, info = LS.dgetrf(A)
    lu, piv, info_s = lapack.dgetrf(A0)
    assert info == info_s
    assert np.array_equal(ipiv - 1, piv)
    assert np.abs(A - lu).max() <= 1e-12 * np.abs(lu).max()


def test_getrs_recovers_after_error_word():
    """The dgetrs error word starts every call at zero: an earlier time-out
    (simulated by writing the word from the host) does not fail later
    solves."""
    import scipy.linalg.lapack as lapack
    from dust_amd import linsys as LS
    rng = np.random.default_rng(5)
    n = 300
    A0 = rng.standard_normal((n, n)) + 0.1 * np.eye(n)
    b = rng.standard_normal((n, 2))
    A = np.asfortranarray(A0.copy())
    ipiv, info = LS.dgetrf(A)
    assert info == 0
    x = np.asfortranarray(b.copy())
    LS.dgetrs(A, ipiv, x)  # allocates the workspace
    L.check(L.lib().dust_test_getrs_set_error_word(1), "poison")
    y = np.asfortranarray(b.copy())
    LS.dgetrs(A, ipiv, y)
    assert np.array_equal(x, y)
    lu, piv, _ = lapack.dgetrf(A0)
    ref, _ = lapack.dgetrs(lu, piv, b)
    assert np.abs(y - ref).max() <= 1e-10 * np.abs(ref).max()


def test_shutdown_then_reuse():
    """dust_shutdown releases the library workspaces; the next calls
    re-create them and give the same results."""
    from dust_amd import linsys as LS
    rng = np.random.default_rng(6)
    n = 200
    A0 = rng.standard_normal((n, n)) + 0.1 * np.eye(n)
    b = rng.standard_normal((n, 1))
    A = np.asfortranarray(A0.copy())
    ipiv, _ = LS.dgetrf(A)
    x = np.asfortranarray(b.copy())
    LS.dgetrs(A, ipiv, x)
    A0 = np.asfortranarray(A0)
    xm = LS.dgesv_mixed(A0, np.asfortranarray(b.copy()), prec="fp16s")[0].copy()
    L.check(L.lib().dust_shutdown(), "shutdown")
    y = np.asfortranarray(b.copy())
    LS.dgetrs(A, ipiv, y)
    assert np.array_equal(x, y)
    ym = LS.dgesv_mixed(A0, np.asfortranarray(b.copy()), prec="fp16s")[0]
    assert np.array_equal(xm, ym)


@pytest.mark.parametrize("case", ["dominant", "mixed", "ties"])
def test_dgetrf_speculative_and_exact_panels(case):
    """Panels whose own-block pivoting matches partial pivoting (the
    speculative path) and panels that need interchanges with rows below
    (factored again exactly), against LAPACK: same pivots, same factors."""
    import scipy.linalg.lapack as lapack
    from dust_amd import linsys as LS
    rng = np.random.default_rng(17)
    n = 700
    A0 = rng.standard_normal((n, n)) + 4.0 * np.sqrt(n) * np.eye(n)
    if case == "mixed":
        # rows of the third 64-column panel need a row from far below
        A0[[140, 600]] = A0[[600, 140]]
    if case == "ties":
        # a row below the block whose candidate equals the block's pivot
        A0[300, 0] = A0[0, 0]
    A = np.asfortranarray(A0.copy())
    ipiv, info = LS.dgetrf(A)
    lu, piv, info_s = lapack.dgetrf(A0)
    assert info == info_s == 0
    assert np.array_equal(ipiv - 1, piv)
    assert np.abs(A - lu).max() <= 1e-12 * np.abs(lu).max()
