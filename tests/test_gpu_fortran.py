"""The Fortran boundary end to end: tests/fortran/dust_gpu_driver (built by
build() with AMD flang against dust_amd/fortran/mod_dust_gpu.f90) calls
initialize_octree -> sort_particles -> calculate_multipole -> apply_multipole
and dgetrf + dgetrs through ISO_C_BINDING, as DUST's mod_octree / mod_linsys
would; results are compared with the CPU oracle and with LAPACK."""
import os
import subprocess

import numpy as np
import pytest

from dust_amd import linsys as LS
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "fortran", "dust_gpu_driver")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_fmm_and_lu(orc, tmp_path):
    cen, alpha, rv, tree = W.helix_wake(6000, seed=4)
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    nlev, min_part, degree = 4, 64, 4
    ver, nv = W.ellipsoid(12, 16)
    A, Bs = LS.aic_surfpan(ver, nv)
    b = LS.rhs_static(ver, nv, Bs, W.freestreams(1))[:, 0]
    n, nlu = len(cen), A.shape[0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, nlev, min_part, degree, nlu], np.int32).tofile(f)
        np.array([tree["box_length"], *tree["origin"], tree["delta"]], np.float64).tofile(f)
        for arr in (cen, dr, mag, rv):                 # Fortran cen(3,n): row-major (n,3)
            np.ascontiguousarray(arr, np.float64).tofile(f)
        np.asfortranarray(A).T.tofile(f)                # column-major A(nlu,nlu)
        b.astype(np.float64).tofile(f)
    print(f"fortran driver n={n} nlu={nlu}", flush=True)
    p = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    with open(fout, "rb") as f:
        nleaves, info = np.fromfile(f, np.int32, 2)
        vel = np.fromfile(f, np.float64, 3 * n).reshape(n, 3)
        st = np.fromfile(f, np.float64, 3 * n).reshape(n, 3)
        x = np.fromfile(f, np.float64, nlu)
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], nlev, min_part, degree,
                   tree["delta"])
    v, s, _ = c.run(cen, dr, mag, rv)
    assert nleaves == len(c.leaves()[0])
    assert rel(vel, v) <= 1e-10
    assert rel(st, s) <= 1e-10
    assert info == 0, p.stdout
    assert rel(x, np.linalg.solve(A, b)) <= 1e-10
