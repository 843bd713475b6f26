"""GPU parity of the AIC assembly and LU path (libdust_gpu through the C ABI)
against the CPU oracle and the host LAPACK (scipy) on identical inputs."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from dust_amd import linsys as LS
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def small_mesh():
    return W.ellipsoid(24, 32)


def test_aic_surfpan_vs_oracle(orc, small_mesh):
    ver, nv = small_mesh
    print(f"aic surfpan n={len(nv)}", flush=True)
    A, B = LS.aic_surfpan(ver, nv)
    G = orc.geo(ver, nv)
    Ao, Bo = orc.aic_surfpan(G)
    assert rel(A, Ao) <= 1e-12
    assert rel(B, Bo) <= 1e-12


def test_aic_vortlatt_vs_oracle(orc, small_mesh):
    ver, nv = small_mesh
    A, B = LS.aic_vortlatt(ver, nv, rankine_rad=0.1, cutoff_rad=0.001)
    G = orc.geo(ver, nv)
    Ao, Bo = orc.aic_vortlatt(G, 10.0, 0.1, 0.001)
    assert rel(A, Ao) <= 1e-12
    assert np.array_equal(B, Bo)


def test_rhs_vs_oracle(orc, small_mesh):
    ver, nv = small_mesh
    A, B = LS.aic_surfpan(ver, nv)
    U = W.freestreams(3)
    b = LS.rhs_static(ver, nv, B, U)
    G = orc.geo(ver, nv)
    for r in range(3):
        bo = orc.rhs_static(G, B, U[r])
        assert rel(b[:, r], bo) <= 1e-12


@pytest.mark.parametrize("n", [1, 7, 64, 65, 200, 1000, 2113])
def test_dgetrf_dgetrs_vs_lapack(n):
    rng = np.random.default_rng(n)
    A0 = rng.standard_normal((n, n)) + 0.1 * np.eye(n)
    A = np.asfortranarray(A0.copy())
    print(f"getrf n={n}", flush=True)
    ipiv, info = LS.dgetrf(A)
    lu, piv, info_s = lapack.dgetrf(A0)
    assert info == info_s == 0
    X = np.asfortranarray(rng.standard_normal((n, 3)))
    B = np.asfortranarray(A0 @ X)
    LS.dgetrs(A, ipiv, B)
    assert rel(B, X) <= 1e-10
    # the factorisation reproduces P A = L U
    Lm = np.tril(A, -1) + np.eye(n)
    Um = np.triu(A)
    PA = A0.copy()
    for i, p in enumerate(ipiv):
        PA[[i, p - 1]] = PA[[p - 1, i]]
    assert np.linalg.norm(PA - Lm @ Um) / np.linalg.norm(A0) <= 1e-13


def test_dgetrf_rectangular_and_singular():
    rng = np.random.default_rng(1)
    A0 = rng.standard_normal((300, 180))
    A = np.asfortranarray(A0.copy())
    ipiv, info = LS.dgetrf(A)
    lu, piv, info_s = lapack.dgetrf(A0)
    assert info == 0
    assert np.allclose(np.abs(np.diag(A)), np.abs(np.diag(lu)), rtol=1e-10)
    S = np.asfortranarray(np.ones((50, 50)))
    ipiv, info = LS.dgetrf(S)
    assert info == lapack.dgetrf(np.ones((50, 50)))[2] > 0


def test_surfpan_solve_end_to_end(orc, small_mesh):
    """Static surface-panel system: GPU assembly + GPU LU solve against the
    oracle assembly + LAPACK solve."""
    ver, nv = small_mesh
    A, B = LS.aic_surfpan(ver, nv)
    b = LS.rhs_static(ver, nv, B, W.freestreams(2))
    A0 = A.copy()
    ipiv, info = LS.dgetrf(A)
    assert info == 0
    x = np.asfortranarray(b.copy())
    LS.dgetrs(A, ipiv, x)
    G = orc.geo(ver, nv)
    Ao, Bo = orc.aic_surfpan(G)
    xo = np.linalg.solve(Ao, b)
    assert rel(x, xo) <= 1e-10
    assert np.linalg.norm(A0 @ x - b) / np.linalg.norm(b) <= 1e-12


@pytest.mark.parametrize("ns,nm", [(300, 0), (0, 250), (200, 130), (70, 5)])
def test_solve_linsys_block_path(ns, nm):
    """The static/moving block elimination (mod_linsys.f90:352-483) gives the
    solution of the full system; the static block is pre-factored by dgetrf
    as initialize_linsys does (mod_linsys.f90:236)."""
    rng = np.random.default_rng(ns * 1000 + nm)
    n = ns + nm
    A0 = rng.standard_normal((n, n)) + 2.0 * np.eye(n)
    b = rng.standard_normal(n)
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(n, np.int32)
    if ns > 0:
        As = np.asfortranarray(A[:ns, :ns].copy())
        p, info = LS.dgetrf(As)
        assert info == 0
        A[:ns, :ns] = As
        ipiv[:ns] = p
    res, info = LS.solve_linsys(A, ipiv, ns, b)
    assert info == 0
    x = np.linalg.solve(A0, b)
    assert rel(res, x) <= 1e-10
    # the factors left in A solve the same system again (dgetrs on them)
    B = np.asfortranarray(b.reshape(n, 1).copy())
    LS.dgetrs(A, ipiv, B)
    assert rel(B[:, 0], x) <= 1e-10
