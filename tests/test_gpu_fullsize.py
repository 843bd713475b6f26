"""Parity at the full configured sizes, through the C ABI against the oracle
and host LAPACK on identical inputs.

* FMM: the 1,000,000-particle 3-helix wake (default seed), 7 levels,
  min_part 64, expansion order 4, velocity + stretching.  FP64 form <= 1e-10
  relative L2 on velocity and stretching, leaf set and membership
  bit-exact; tuned form <= 1e-6 on velocity and <= 2e-5 on stretching.
* AIC: the 16,384-panel perturbed ellipsoid (128 x 128, pole rows of
  triangles).  Full rows of A and B against the oracle for 384 panels
  (both pole-triangle rings and 128 rows spread over the quads) <= 1e-12;
  the FP64 getrf/getrs solution and both mixed-precision solutions of the
  8 right-hand sides against the host LAPACK dgetrf/dgetrs solution
  <= 1e-10 relative L2.
"""
import numpy as np
import pytest

from dust_amd import linsys as LS
from dust_amd import octree as O
from dust_amd import workloads as W

from test_gpu_fmm import check_leaves, rel

pytestmark = pytest.mark.gpu


def _fmm(parts, tree, form):
    g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                            tree["min_part"], 4, tree["delta"], capacity=parts.n, use_vs=True,
                            precision=form)
    O.sort_particles(parts, g)
    O.calculate_multipole(parts, g)
    O.apply_multipole(parts, g)
    return g, parts.vel.copy(), parts.stretch.copy()


def test_helix_1M_vs_oracle(orc):
    cen, alpha, rv, tree = W.helix_wake(1_000_000)
    assert tree["nlevels"] == 7 and tree["min_part"] == 64
    parts = O.VortParts.from_alpha(cen, alpha, rv)
    print(f"helix n={parts.n} levels={tree['nlevels']} p=4 fp64", flush=True)
    g, v64, s64 = _fmm(parts, tree, "fp64")
    print("helix tuned", flush=True)
    gt, vt, stt = _fmm(parts, tree, "tuned")
    gt.destroy()
    print("helix oracle", flush=True)
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                   tree["min_part"], 4, tree["delta"])
    v, s, _ = c.run(parts.cen, parts.dir, parts.mag, parts.r_vortex)
    ev, es = rel(v64, v), rel(s64, s)
    et, ets = rel(vt, v), rel(stt, s)
    print(f"fp64 vel {ev:.2e} stretch {es:.2e}; tuned vel {et:.2e} stretch {ets:.2e}")
    assert ev <= 1e-10 and es <= 1e-10
    assert et <= 1e-6 and ets <= 2e-5
    check_leaves(parts, g, c)
    assert np.array_equal(g.work(), c.work())
    g.destroy()


@pytest.fixture(scope="module")
def aic16k():
    ver, nv = W.ellipsoid(128, 128)
    assert len(nv) == 16384
    print("aic n=16384", flush=True)
    A, B = LS.aic_surfpan(ver, nv)
    return ver, nv, A, B


def test_aic16384_rows_vs_oracle(orc, aic16k):
    ver, nv, A, B = aic16k
    n = len(nv)
    rng = np.random.default_rng(20261204)
    quads = np.flatnonzero(nv == 4)
    rows = np.concatenate([np.arange(128), np.arange(n - 128, n),
                           np.sort(rng.choice(quads, 128, replace=False))])
    assert np.all(nv[:128] == 3) and np.all(nv[-128:] == 3)
    G = orc.geo(ver, nv)
    Ar, Br = orc.aic_surfpan_rows(G, rows)
    ea, eb = rel(A[rows], Ar), rel(B[rows], Br)
    print(f"rows {len(rows)}: A {ea:.2e} B {eb:.2e}")
    assert ea <= 1e-12 and eb <= 1e-12


def test_aic16384_solves_vs_lapack(aic16k):
    import scipy.linalg.lapack as lapack
    ver, nv, A, B = aic16k
    b = LS.rhs_static(ver, nv, B, W.freestreams(8))
    print("host LAPACK dgetrf/dgetrs n=16384", flush=True)
    lu, piv, info = lapack.dgetrf(A)
    assert info == 0
    xr, info = lapack.dgetrs(lu, piv, b)
    assert info == 0
    del lu
    print("gpu fp64 getrf/getrs", flush=True)
    F = np.asfortranarray(A.copy())
    ipiv, info = LS.dgetrf(F)
    assert info == 0
    x64 = LS.dgetrs(F, ipiv, np.asfortranarray(b.copy()))
    del F
    e64 = rel(x64, xr)
    print(f"fp64 x vs LAPACK {e64:.2e}", flush=True)
    assert e64 <= 1e-10
    for prec in ("fp16s", "fp32"):
        x, iters, berr = LS.dgesv_mixed(A, np.asfortranarray(b.copy()), prec=prec)
        e = rel(x, xr)
        print(f"{prec}: iters {iters} backward error {berr:.2e} x vs LAPACK {e:.2e}", flush=True)
        assert iters >= 0 and berr <= 1e-12
        assert e <= 1e-10
