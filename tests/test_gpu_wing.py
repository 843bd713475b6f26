"""The wing configuration (a rectangular NACA0012 wing of 1,600 surface panels
with an 8,000-particle trailing sheet, order 4): the FMM of the sheet and the
wing's surface-panel system on the GPU against the CPU oracle, in both
forms."""
import numpy as np
import pytest

from dust_amd import linsys as LS
from dust_amd import octree as O
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def wing():
    return W.wing_case()


@pytest.fixture(scope="module")
def sheet_oracle(orc, wing):
    ver, nv, cen, alpha, rv, tree = wing
    parts = O.VortParts.from_alpha(cen, alpha, rv)
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                   tree["min_part"], 4, tree["delta"], use_vs=True)
    v, s, _ = c.run(parts.cen, parts.dir, parts.mag, parts.r_vortex)
    return c, v, s


@pytest.mark.parametrize("precision", ["fp64", "tuned"])
def test_sheet_fmm(wing, sheet_oracle, precision):
    ver, nv, cen, alpha, rv, tree = wing
    c, v, s = sheet_oracle
    parts = O.VortParts.from_alpha(cen, alpha, rv)
    print(f"wing sheet n={parts.n} {precision}", flush=True)
    g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                            tree["min_part"], 4, tree["delta"], capacity=parts.n, use_vs=True,
                            precision=precision)
    O.sort_particles(parts, g)
    O.calculate_multipole(parts, g)
    O.apply_multipole(parts, g)
    ev, es = rel(parts.vel, v), rel(parts.stretch, s)
    print(f"{precision}: vel {ev:.2e} stretch {es:.2e}")
    if precision == "fp64":
        assert ev <= 1e-10 and es <= 1e-10
        lev_g, ijk_g, np_g, mem_g = g.leaves(parts.n)
        lev_c, ijk_c, np_c, mem_c = c.leaves()
        assert sorted(zip(lev_g.tolist(), map(tuple, ijk_g.tolist()), np_g.tolist())) == \
            sorted(zip(lev_c.tolist(), map(tuple, ijk_c.tolist()), np_c.tolist()))
    else:
        assert ev <= 1e-6 and es <= 2e-5
    g.destroy()


def test_wing_system(orc, wing):
    """AIC of the 1,600 wing panels and the static free-stream system: FP64
    LU and the tuned mixed solve against the oracle assembly and LAPACK."""
    ver, nv = wing[0], wing[1]
    print(f"wing aic n={len(nv)}", flush=True)
    A, B = LS.aic_surfpan(ver, nv)
    G = orc.geo(ver, nv)
    Ao, Bo = orc.aic_surfpan(G)
    assert rel(A, Ao) <= 1e-12
    assert rel(B, Bo) <= 1e-12
    u = np.array([[30.0 * np.cos(np.radians(5.0)), 0.0, 30.0 * np.sin(np.radians(5.0))]])
    b = LS.rhs_static(ver, nv, B, u)
    bo = orc.rhs_static(G, Bo, u[0])
    assert rel(b.ravel(), bo) <= 1e-12
    xo = np.linalg.solve(Ao, bo)
    LU = np.asfortranarray(A.copy())
    ipiv, info = LS.dgetrf(LU)
    assert info == 0
    x = np.asfortranarray(b.copy())
    LS.dgetrs(LU, ipiv, x)
    assert rel(x.ravel(), xo) <= 1e-10
    for prec in ("fp32", "fp16s"):
        xm, it, be = LS.dgesv_mixed(np.asfortranarray(A), np.asfortranarray(b.copy()), prec=prec,
                                    tol=1e-12, max_iter=40)
        print(f"{prec}: iters {it} berr {be:.2e}")
        assert it >= 1 and be <= 1e-12
        assert rel(xm.ravel(), xo) <= 1e-10
