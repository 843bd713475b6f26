"""dust_dgetrs_noalloc_dev (static device workspace, no allocation) against
dust_dgetrs_dev through the C ABI: bit-identical solutions on two calls in a
row, the second with a smaller system, and a clear error beyond the
workspace's maximum size."""
import numpy as np
import pytest

from dust_amd import _lib as L

pytestmark = pytest.mark.gpu


def _factor(torch, n, seed):
    rng = np.random.default_rng(seed)
    A0 = rng.standard_normal((n, n)) + 0.2 * n ** 0.5 * np.eye(n)
    dA = torch.from_numpy(np.ascontiguousarray(A0.T)).cuda()  # column-major
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = L.lib().dust_dgetrf_dev(n, n, dA.data_ptr(), n, ipiv.data_ptr(), None)
    assert info == 0
    return A0, dA, ipiv


@pytest.mark.parametrize("sizes", [(2000, 700), (6144, 130)])
def test_noalloc_getrs_matches_getrs(sizes):
    torch = pytest.importorskip("torch")
    lib = L.lib()
    for k, (n, nrhs) in enumerate(zip(sizes, (9, 3))):
        print(f"noalloc getrs n={n} nrhs={nrhs}", flush=True)
        A0, dA, ipiv = _factor(torch, n, 10 + k)
        B = np.random.default_rng(20 + k).standard_normal((nrhs, n))  # rows = columns of B
        x1 = torch.from_numpy(B.copy()).cuda()
        x2 = torch.from_numpy(B.copy()).cuda()
        L.check(lib.dust_dgetrs_dev(n, nrhs, dA.data_ptr(), n, ipiv.data_ptr(), x1.data_ptr(), n,
                                    None), "getrs")
        L.check(lib.dust_dgetrs_noalloc_dev(n, nrhs, dA.data_ptr(), n, ipiv.data_ptr(),
                                            x2.data_ptr(), n, None), "getrs_noalloc")
        a, b = x1.cpu().numpy(), x2.cpu().numpy()
        assert np.array_equal(a, b)
        r = np.abs(A0 @ b.T - B.T).max() / (np.abs(A0).sum(1).max() * np.abs(b).max())
        assert r <= 1e-13


def test_noalloc_getrs_rejects_large_systems():
    torch = pytest.importorskip("torch")
    n = 6145
    d = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = L.lib().dust_dgetrs_noalloc_dev(n, 1, d.data_ptr(), n, d.data_ptr(), d.data_ptr(), n, None)
    assert rc == L.DUST_E_CAPACITY
    assert "6144" in L.lib().dust_last_error().decode()
