"""The oracle's restatement of initialize_linsys / assemble_linsys with the
panel wake (oracle/linsys_oracle.c), pinned without the reference:

* lift of a vortex-lattice flat plate and of a surface-panel NACA0012 wing
  with the implicit first wake row against lifting-surface theory
  (Helmbold's lift slope): a wrong sign or a missing implicit row moves the
  lift far outside the band;
* an explicit wake panel equals its two halves (solid angle additivity for
  surface-panel rows, cancelling shared edges for vortex-lattice rows);
* an implicit wake panel on the left-hand side equals an explicit one of
  strength mu(ind1) - mu(ind2) on the right-hand side;
* a body moving with velocity V equals the body at rest in the wind u - V;
* a static/moving split of the rows gives the same system as all rows
  moving; static explicit elements through L_static equal moving ones.
"""
import numpy as np
import pytest

from dust_amd import linsys as LS

import linsys_cases as LC

U = 10.0
AL = np.radians(5.0)


def solve_static(orc, S):
    st = orc.initialize_linsys(S)
    assert st["info"] == 0
    b, bp = orc.assemble_linsys(S, st)
    x = b.copy()
    orc.lib().ao_dgetrs(S.rank, 1, orc._p(st["A"]), S.rank, orc._p(st["ipiv"]), orc._p(x),
                        S.rank)
    return x


def test_vlm_plate_lift(orc):
    nc, ns = 6, 24
    ver, nv, dy, wake = LC.vlm_plate(nc=nc, ns=ns)
    S = LC.system(ver, nv, LS.ROW_VORTLATT, len(nv), wake=wake,
                  u_inf=(U * np.cos(AL), 0.0, U * np.sin(AL)), ff_ratio=1e9, rankine=1e-4)
    x = solve_static(orc, S)
    gte = x[nc - 1::nc]  # trailing-edge ring of every strip = wake strength
    assert np.all(gte < 0) or np.all(gte > 0)
    cl = 2.0 * np.abs(gte).sum() * dy / (U * 8.0)
    ratio = cl / (LC.helmbold_cla(8.0) * AL)
    print(f"VLM plate CL {cl:.4f}, {ratio:.3f} of Helmbold")
    assert 0.90 <= ratio <= 1.02


def test_morino_wing_lift(orc):
    ver, nv, dy, wake, up, lo = LC.morino_wing()
    S = LC.system(ver, nv, LS.ROW_SURFPAN, len(nv), wake=wake, u_inf=(U, 0.0, 0.0))
    x = solve_static(orc, S)
    g = x[up] - x[lo]  # Kutta condition: wake strength = mu_upper - mu_lower
    assert np.all(g < 0) or np.all(g > 0)
    cl = 2.0 * np.abs(g).sum() * dy / (U * 8.0)
    ratio = cl / (LC.helmbold_cla(8.0) * AL)
    print(f"surface-panel wing CL {cl:.4f}, {ratio:.3f} of Helmbold")
    assert 0.93 <= ratio <= 1.06
    # the wake panels traversed the other way round do not continue the
    # trailing-edge panels: the lift collapses
    wf = dict(wake, ver=wake["ver"][:, [1, 0, 3, 2]])
    xf = solve_static(orc, LC.system(ver, nv, LS.ROW_SURFPAN, len(nv), wake=wf,
                                     u_inf=(U, 0.0, 0.0)))
    assert 2.0 * np.abs(xf[up] - xf[lo]).sum() * dy / (U * 8.0) < 0.2 * cl


def _moving(orc, S):
    """All rows moving: A and b of one assembly (nothing factorised)."""
    st = orc.initialize_linsys(S)
    b, bp = orc.assemble_linsys(S, st)
    return st, b, bp


@pytest.mark.parametrize("kind", [LS.ROW_SURFPAN, LS.ROW_VORTLATT])
def test_explicit_wake_panel_equals_its_halves(orc, kind):
    ver, nv, dy, wake, up, lo = LC.morino_wing(nc=6, ns=4)
    w = wake["ver"]
    mid1, mid2 = 0.5 * (w[:, 0] + w[:, 3]), 0.5 * (w[:, 1] + w[:, 2])
    whole = dict(ver=w, n_ver=wake["n_ver"], mag=np.linspace(0.5, 1.5, len(w)), n_impl=0)
    halves = dict(ver=np.concatenate([np.stack([w[:, 0], w[:, 1], mid2, mid1], 1),
                                      np.stack([mid1, mid2, w[:, 2], w[:, 3]], 1)]),
                  n_ver=np.full(2 * len(w), 4, np.int32),
                  mag=np.tile(whole["mag"], 2), n_impl=0)
    kw = dict(u_inf=(U, 0.0, 0.0), ff_ratio=1e9, rankine=1e-4)
    _, b1, _ = _moving(orc, LC.system(ver, nv, kind, 0, wake=whole, **kw))
    _, b2, _ = _moving(orc, LC.system(ver, nv, kind, 0, wake=halves, **kw))
    _, b0, _ = _moving(orc, LC.system(ver, nv, kind, 0, **kw))
    assert np.abs((b1 - b0) - (b2 - b0)).max() <= 1e-10 * np.abs(b1 - b0).max()


@pytest.mark.parametrize("kind", [LS.ROW_SURFPAN, LS.ROW_VORTLATT])
def test_implicit_wake_equals_explicit(orc, kind):
    ver, nv, dy, wake, up, lo = LC.morino_wing(nc=6, ns=4)
    if kind == LS.ROW_VORTLATT:
        wake = dict(wake, impl_ind=np.stack([up + 1, np.zeros_like(up)], 1).astype(np.int32))
    mu = np.random.default_rng(3).normal(size=len(nv))
    i1, i2 = wake["impl_ind"][:, 0] - 1, wake["impl_ind"][:, 1] - 1
    strength = mu[i1] - np.where(i2 >= 0, mu[np.maximum(i2, 0)], 0.0)
    expl = dict(wake, n_impl=0, mag=strength)
    kw = dict(u_inf=(U, 0.0, 0.0))
    st1, b1, _ = _moving(orc, LC.system(ver, nv, kind, 0, wake=wake, **kw))
    st2, b2, _ = _moving(orc, LC.system(ver, nv, kind, 0, wake=expl, **kw))
    r1, r2 = st1["A"] @ mu - b1, st2["A"] @ mu - b2
    assert np.abs(r1 - r2).max() <= 1e-12 * np.abs(r1).max()
    assert not np.allclose(st1["A"], st2["A"])


def test_moving_body_equals_body_at_rest_in_relative_wind(orc):
    ver, nv, dy, wake, up, lo = LC.morino_wing(nc=6, ns=4)
    V = np.array([2.0, -1.0, 0.5])
    n = len(nv)
    uv = np.random.default_rng(4).normal(size=(n, 3))
    _, b1, _ = _moving(orc, LC.system(ver, nv, LS.ROW_SURFPAN, 0, wake=wake, ub=np.tile(V, (n, 1)),
                                      uvort=uv, u_inf=(U, 0.0, 0.0)))
    _, b2, _ = _moving(orc, LC.system(ver, nv, LS.ROW_SURFPAN, 0, wake=wake, uvort=uv,
                                      u_inf=tuple(np.array([U, 0.0, 0.0]) - V)))
    assert np.abs(b1 - b2).max() <= 1e-12 * np.abs(b1).max()


def test_static_moving_split_equals_all_moving(orc):
    """Rows split into static and moving blocks (the split between strips, so
    no implicit wake panel straddles it): the assembled moving blocks, the
    right-hand side, the Bernoulli right-hand side and the static block (from
    its LU factors) equal the all-moving assembly."""
    import scipy.linalg as sla
    c = LC.mixed_case()
    nst = 3 * c["strip"]
    Sm = LC.mixed_system(c, 0, 0, 0)
    stm, bm, bpm = _moving(orc, Sm)
    Ss = LC.mixed_system(c, nst, nst, 2)
    sts = orc.initialize_linsys(Ss)
    assert sts["info"] == 0
    bs, bps = orc.assemble_linsys(Ss, sts)
    tol = 1e-12 * np.abs(bm).max()
    assert np.abs(bs - bm).max() <= tol
    assert np.abs(bps - bpm).max() <= 1e-12 * np.abs(bpm).max()
    A = stm["A"]
    assert np.array_equal(sts["A"][:, nst:], A[:, nst:])
    assert np.array_equal(sts["A"][nst:, :nst], A[nst:, :nst])
    # static block: P L U of the factors
    f = sts["A"][:nst, :nst]
    Lf, Uf = np.tril(f, -1) + np.eye(nst), np.triu(f)
    perm = np.arange(nst)
    for i, p in enumerate(sts["ipiv"][:nst] - 1):
        perm[[i, p]] = perm[[p, i]]
    rec = np.empty_like(f)
    rec[perm] = Lf @ Uf
    assert np.abs(rec - A[:nst, :nst]).max() <= 1e-12 * np.abs(A[:nst, :nst]).max()
    # the pressure matrix: blocks of A before the wake correction, the static
    # one factorised
    ids = c["id_sp"]
    A0 = _moving(orc, LS.System(c["kind"], c["ver"], c["n_ver"], 0, uvort=c["uvort"],
                                u_inf=(U, 0.0, 1.0)))[0]["A"]
    Ap = sts["A_pres"]
    assert np.array_equal(Ap[:nst, nst:], A0[np.ix_(ids[:nst], ids[nst:])])
    assert np.array_equal(Ap[nst:, :], A0[np.ix_(ids[nst:], ids)])
    P, Lp, Up = sla.lu(A0[np.ix_(ids[:nst], ids[:nst])])
    assert np.abs(np.tril(Ap[:nst, :nst], -1) + np.eye(nst) - Lp).max() <= 1e-12
    assert np.abs(np.triu(Ap[:nst, :nst]) - Up).max() <= 1e-12 * np.abs(Up).max()


def test_static_explicit_elements_through_l_static(orc):
    c = LC.mixed_case()
    nst = 3 * c["strip"]
    b = {}
    for nse in (0, 2):
        S = LC.mixed_system(c, nst, nst, nse)
        st = orc.initialize_linsys(S)
        b[nse], _ = orc.assemble_linsys(S, st)
    assert np.abs(b[0] - b[2]).max() <= 1e-12 * np.abs(b[0]).max()
