"""The trailing-edge description of the time step, restated twice (the
product's dust_amd.step.trailing_edge / te_geometry and the oracle pipeline's
strip_trailing_edge / te_rows_input), pinned by the wing's known layout: the
stripes, their shared points and orientations, and the first wake row along
the chord of the pitched wing (the mean of the mirrored upper and lower
trailing-edge panels lies on the chord line); and the pipeline's first step,
which with an empty wake is the steady system with one implicit wake panel
of length |u| dt along the chord."""
import numpy as np

from dust_amd import step as ST
from dust_amd import workloads as W
from oracle import pipeline as PL


def test_trailing_edge_tables_of_the_wing():
    ns, nc = 6, 5
    ver, nv, *_ = W.wing_case(nc, ns, 8.0, 5.0, nrows=1)
    a = ST.trailing_edge(ver, ns, nc)
    b = PL.strip_trailing_edge(ver, ns, nc)
    for k in ("up", "lo", "start", "neigh", "o", "gen"):
        assert np.array_equal(a[k], b[k]), k
    # vertex 1 of the upper panel of strip j is the trailing-edge point at
    # y_{j+1}, vertex 4 the one at y_j: points in the order y_1, y_0, y_2, ...
    ys = np.linspace(-4.0, 4.0, ns + 1)
    pts, _ = ST.te_geometry(ver, a)
    assert np.allclose(pts[a["start"][:, 0] - 1, 1], ys[1:])
    assert np.allclose(pts[a["start"][:, 1] - 1, 1], ys[:-1])
    # stripe j shares its first point with stripe j+1 (as that one's second)
    for j in range(ns):
        assert a["neigh"][j, 0] == (j + 2 if j < ns - 1 else 0)
        assert a["neigh"][j, 1] == (j if j > 0 else 0)
    assert np.all(a["o"][a["neigh"] > 0] == 1)


def test_first_row_follows_the_chord():
    ver, nv, *_ = W.wing_case(8, 4, 8.0, 5.0, nrows=1)
    ca, sa = np.cos(np.radians(5.0)), np.sin(np.radians(5.0))
    for pts, d in (ST.te_geometry(ver, ST.trailing_edge(ver, 4, 8)),
                   PL.te_rows_input(ver, PL.strip_trailing_edge(ver, 4, 8))):
        assert np.allclose(d, np.tile([ca, 0.0, -sa], (len(d), 1)), atol=1e-12)


def test_first_step_is_the_short_implicit_wake(orc):
    import linsys_cases  # noqa: F401  (tests on the path)
    from dust_amd import linsys as LS
    ns, nc = 6, 8
    ver, nv, cen, alpha, rv, tree = W.wing_case(nc, ns, 8.0, 5.0, nrows=1)
    u, dt = np.array([30.0, 0.0, 0.0]), 1.0 / 600.0
    o = PL.OracleStep((ver, nv, cen[:0], alpha[:0], rv[:0], tree), (ns, nc), 100, u, dt)
    r = o.step()
    te = PL.strip_trailing_edge(ver, ns, nc)
    pts, d = PL.te_rows_input(ver, te)
    st = te["start"] - 1
    L = 30.0 * dt
    wv = np.stack([pts[st[:, 0]], pts[st[:, 1]], pts[st[:, 1]] + L * d[st[:, 1]],
                   pts[st[:, 0]] + L * d[st[:, 0]]], 1)
    wake = dict(ver=wv, n_ver=np.full(ns, 4, np.int32), mag=np.zeros(ns), n_impl=ns,
                impl_ind=te["gen"])
    S = LS.System(np.zeros(len(nv), np.int32), ver, nv, len(nv), wake=wake, u_inf=tuple(u))
    s = orc.initialize_linsys(S)
    b, _ = orc.assemble_linsys(S, s)
    x = b.copy()
    orc.lib().ao_dgetrs(S.rank, 1, orc._p(s["A"]), S.rank, orc._p(s["ipiv"]), orc._p(x), S.rank)
    assert np.abs(r["mu"] - x).max() <= 1e-12 * np.abs(x).max()
