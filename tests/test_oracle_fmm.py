"""Pins of the CPU oracle's FMM restatement (CPU only).

Each check compares the oracle with something it does not contain: a closed
form, a symbolic derivative, a brute-force sum, an exactness identity of the
expansion operators or an independent characterisation of the tree."""
import itertools
import os

import numpy as np
import pytest

from dust_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


# ---- monomial tables (set_degree_polyexp, mod_multipole.f90:454-501) -------
def test_monomial_order_known_answer(orc):
    pwr, nmd = orc.pexp_table(2)
    # graded by order, then i descending, then j descending (read from the
    # reference loops do o; do i=o,0,-1; do j=o-i,0,-1)
    expect = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1),
              (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    assert [tuple(r) for r in pwr] == expect
    assert list(nmd) == [1, 4, 10]


@pytest.mark.parametrize("deg", [0, 3, 8, 18])
def test_monomial_table_is_graded_bijection(orc, deg):
    pwr, nmd = orc.pexp_table(deg)
    n = len(pwr)
    assert n == (deg + 1) * (deg + 2) * (deg + 3) // 6
    assert len({tuple(r) for r in pwr}) == n
    orders = pwr.sum(axis=1)
    assert np.all(np.diff(orders) >= 0)
    for o in range(deg + 1):
        assert nmd[o] == np.count_nonzero(orders <= o)


# ---- kernel derivatives (compute_der_ker, mod_multipole.f90:388-450) -------
@pytest.mark.parametrize("diff,delta", [((0.7, -0.4, 1.3), 0.0), ((-2.0, 0.5, 0.25), 0.3),
                                        ((0.0, 0.0, 3.0), 0.02)])
def test_kernel_derivatives_vs_sympy(orc, diff, delta):
    import sympy as sp
    deg = 6  # pexp_der degree of an order-2 expansion
    x, y, z = sp.symbols("x y z", real=True)
    f = 1 / (4 * sp.pi * sp.sqrt(x**2 + y**2 + z**2 + sp.Rational(str(delta)) ** 2))
    D, Dc = orc.der_ker(deg, diff, delta)
    pw, _ = orc.pexp_table(deg)
    subs = {x: sp.Rational(str(diff[0])), y: sp.Rational(str(diff[1])), z: sp.Rational(str(diff[2]))}
    var = (x, y, z)
    for m in range(D.shape[0]):
        i, j, k = pw[m]
        # the recurrence differentiates with respect to the source point:
        # bk(m) = (-1)^|m| (1/m!) d^m f, hence
        # D(c, m) = (-1)^|m| (1/m!) d^m d_c f and Dc(c, ii, m) = (-1)^|m| (1/m!) d^m d_ii d_c f
        af = float(sp.factorial(i) * sp.factorial(j) * sp.factorial(k)) * (-1) ** (i + j + k)
        base = sp.diff(f, x, i, y, j, z, k) if (i + j + k) else f
        for c in range(3):
            ref = float(sp.diff(base, var[c]).subs(subs)) / af
            assert D[m, c] == pytest.approx(ref, rel=1e-11, abs=1e-14)
        if m < Dc.shape[0]:
            for c in range(3):
                for ii in range(3):
                    ref = float(sp.diff(base, var[c], var[ii]).subs(subs)) / af
                    assert Dc[m, c, ii] == pytest.approx(ref, rel=1e-11, abs=1e-14)


# ---- particle kernels (mod_vortpart.f90:139-336) ------------------------------
def test_biot_savart_closed_form(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        xs, d, x = rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3)
        d /= np.linalg.norm(d)
        mag, r = rng.uniform(0.1, 2.0), rng.uniform(0.01, 0.5)
        v, s, df = orc.pair(xs, d, mag, r, x, np.zeros(3))
        rr = x - xs
        R3 = (rr @ rr + r * r) ** 1.5
        # u = (1/4pi) alpha x (x - x_s) / (|x - x_s|^2 + r^2)^(3/2), returned times 4 pi
        assert np.allclose(v, np.cross(mag * d, rr) / R3, rtol=1e-13, atol=0)


def test_stretching_is_transposed_velocity_gradient(orc):
    """compute_stretch = (grad u)^T alpha_t (the 'transpose' scheme), checked
    against a finite-difference gradient of the oracle's own velocity."""
    rng = np.random.default_rng(1)
    for _ in range(10):
        xs, d, x, at = (rng.standard_normal(3) for _ in range(4))
        d /= np.linalg.norm(d)
        mag, r = rng.uniform(0.5, 2.0), rng.uniform(0.05, 0.5)
        h = 1e-5
        G = np.zeros((3, 3))  # G[i, j] = d u_j / d x_i
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            G[i] = (orc.pair(xs, d, mag, r, x + e, at)[0] - orc.pair(xs, d, mag, r, x - e, at)[0]) / (2 * h)
        s = orc.pair(xs, d, mag, r, x, at)[1]
        assert np.allclose(s, G @ at, rtol=1e-7, atol=1e-9 * np.abs(G).max())
        assert not np.allclose(s, G.T @ at, rtol=1e-3)


def test_pse_kernel_moments(orc):
    """etaeps integrates to 4 and has second moment 3 eps^2 in 3-D (closed forms
    via Beta functions); checked by radial quadrature of the pair diffusion."""
    from scipy.integrate import quad
    eps = 0.3
    xs, d = np.zeros(3), np.array([1.0, 0, 0])
    vol = 4.0 / 3.0 * np.pi * eps**3

    def eta(rho):
        # pair diffusion with alpha_t = 0 and mag 1: (1/eps^2) vol dir eta
        _, _, df = orc.pair(xs, d, 1.0, eps, np.array([rho, 0, 0]), np.zeros(3))
        return df[0] * eps**2 / vol

    m0 = quad(lambda r: 4 * np.pi * r**2 * eta(r), 0, 50 * eps, limit=200)[0]
    m2 = quad(lambda r: 4 * np.pi * r**4 * eta(r), 0, 200 * eps, limit=400)[0]
    assert m0 == pytest.approx(4.0, rel=1e-8)
    assert m2 == pytest.approx(3.0 * eps**2, rel=1e-6)


def test_self_term_zero_and_antisymmetry(orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(3)
    d = np.array([0.0, 0.6, 0.8])
    v, s, df = orc.pair(x, d, 1.5, 0.1, x, np.array([0.3, 0.2, 0.1]))
    assert np.all(v == 0) and np.allclose(s, -(-np.cross([0.3, 0.2, 0.1], 1.5 * d) / 0.1**3))
    # velocity of a source seen from +r and -r is opposite
    r = rng.standard_normal(3)
    v1 = orc.pair(x, d, 1.0, 0.1, x + r, np.zeros(3))[0]
    v2 = orc.pair(x, d, 1.0, 0.1, x - r, np.zeros(3))[0]
    assert np.allclose(v1, -v2, rtol=1e-14)


# ---- expansion operators ---------------------------------------------------------
def _cluster(rng, n, c, h):
    pc = c + rng.uniform(-h, h, (n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return pc, d, rng.uniform(0.5, 1.5, n)


@pytest.mark.parametrize("p", [2, 4, 8])
def test_m2m_exact(orc, p):
    """M2M of raw moments is exact: shifting the children's P2M to the parent
    equals the parent's P2M of all particles (mod_multipole.f90:256-278)."""
    rng = np.random.default_rng(p)
    cen = np.array([0.1, -0.2, 0.3])
    a_par = np.zeros(((p + 1) * (p + 2) * (p + 3) // 6, 3))
    allp = []
    for ch in itertools.product((-0.25, 0.25), repeat=3):
        cc = cen + np.array(ch)
        pc, d, m = _cluster(rng, 30, cc, 0.25)
        allp.append((pc, d, m))
        a_par = orc.m2m(p, cen, orc.p2m(p, cc, pc, d, m), cc, a_par)
    pc = np.concatenate([q[0] for q in allp])
    d = np.concatenate([q[1] for q in allp])
    m = np.concatenate([q[2] for q in allp])
    direct = orc.p2m(p, cen, pc, d, m)
    assert np.allclose(a_par, direct, rtol=1e-12, atol=1e-14 * np.abs(direct).max())


@pytest.mark.parametrize("p", [2, 4, 8])
def test_l2l_exact(orc, p):
    """L2L re-centres a degree-p polynomial exactly (mod_multipole.f90:360-384)."""
    rng = np.random.default_rng(10 + p)
    nm = (p + 1) * (p + 2) * (p + 3) // 6
    b = rng.standard_normal((nm, 3))
    c = rng.standard_normal((nm, 3, 3))
    pc, cc = np.array([0.0, 0.0, 0.0]), np.array([0.25, -0.25, 0.25])
    bc, ccc = orc.l2l(p, cc, b, c, pc)
    for _ in range(5):
        x = cc + rng.uniform(-0.25, 0.25, 3)
        v1, g1 = orc.l2p(p, pc, b, c, x)
        v2, g2 = orc.l2p(p, cc, bc, ccc, x)
        assert np.allclose(v1, v2, rtol=1e-12, atol=1e-12 * np.abs(v1).max())
        assert np.allclose(g1, g2, rtol=1e-12, atol=1e-12 * np.abs(g1).max())


def test_m2l_converges_to_direct(orc):
    """P2M -> M2L -> L2P approaches the direct regularised sum geometrically
    in the order p (delta equal to the particle radius)."""
    rng = np.random.default_rng(4)
    delta = 0.05
    src_c, tgt_c = np.zeros(3), np.array([2.0, 0.0, 0.0])
    pc, d, m = _cluster(rng, 40, src_c, 0.5)
    tx = tgt_c + rng.uniform(-0.5, 0.5, (10, 3))
    rv = np.full(len(m), delta)
    at = rng.standard_normal((10, 3))
    vel_d = np.zeros((10, 3))
    str_d = np.zeros((10, 3))
    for t in range(10):
        for q in range(len(m)):
            v, s, _ = orc.pair(pc[q], d[q], m[q], delta, tx[t], at[t])
            vel_d[t] += v / (4 * np.pi)
            str_d[t] += s / (4 * np.pi)
    errs = []
    for p in (2, 4, 6, 8):
        a = orc.p2m(p, src_c, pc, d, m)
        b, c = orc.m2l(p, tgt_c - src_c, delta, a)
        vv = np.zeros((10, 3))
        ss = np.zeros((10, 3))
        for t in range(10):
            v, g = orc.l2p(p, tgt_c, b, c, tx[t])
            vv[t] = v
            ss[t] = g @ at[t]  # str_j = sum_i grad(j, i) alpha_i
        errs.append((np.linalg.norm(vv - vel_d) / np.linalg.norm(vel_d),
                     np.linalg.norm(ss - str_d) / np.linalg.norm(str_d)))
    ev = [e[0] for e in errs]
    es = [e[1] for e in errs]
    assert all(ev[i + 1] < 0.5 * ev[i] for i in range(3)), ev
    assert all(es[i + 1] < 0.6 * es[i] for i in range(3)), es
    assert ev[-1] < 1e-3 and es[-1] < 3e-3


# ---- tree and full FMM ---------------------------------------------------------------
def _tree_leaves_independent(cen, tree, nlev):
    """Leaves from an independent characterisation of the reference rule: a
    cell is a branch iff it is above the finest level and one of its children
    holds >= min_part particles; a leaf is a non-branch cell whose parent is a
    branch (or a level-1 cell that is not a branch)."""
    org = np.asarray(tree["origin"])
    box = tree["box_length"]
    nb = np.asarray(tree["nbox"])
    cs = box / 2 ** (nlev - 1)
    idx = np.ceil((cen - org) / cs).astype(np.int64) - 1  # 0-based finest cells
    counts = [dict() for _ in range(nlev)]
    for l in range(nlev):
        sh = nlev - 1 - l
        keys, cnt = np.unique(idx >> sh, axis=0, return_counts=True)
        counts[l] = {tuple(k): int(c) for k, c in zip(keys, cnt)}

    def npart(l, c):
        return counts[l].get(c, 0)

    def branch(l, c):
        if l == nlev - 1:
            return False
        kids = [(2 * c[0] + a, 2 * c[1] + b, 2 * c[2] + e) for e in (0, 1) for b in (0, 1) for a in (0, 1)]
        return any(npart(l + 1, k) >= tree["min_part"] for k in kids)

    leaves = {}
    stack = [(0, (i, j, k)) for k in range(nb[2]) for j in range(nb[1]) for i in range(nb[0])]
    while stack:
        l, c = stack.pop()
        if branch(l, c):
            for e in (0, 1):
                for b in (0, 1):
                    for a in (0, 1):
                        stack.append((l + 1, (2 * c[0] + a, 2 * c[1] + b, 2 * c[2] + e)))
        else:
            leaves[(l + 1, c[0] + 1, c[1] + 1, c[2] + 1)] = npart(l, c)
    return leaves


@pytest.mark.parametrize("nlev,minp", [(3, 8), (5, 64), (6, 5)])
def test_leaf_set_independent(orc, nlev, minp):
    cen, alpha, rv, tree = W.helix_wake(6000, seed=3)
    tree = dict(tree, nlevels=nlev, min_part=minp)
    mag = np.linalg.norm(alpha, axis=1)
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], nlev, minp, 2, tree["delta"])
    c.sort(cen, alpha / mag[:, None], mag, rv)
    lev, ijk, npart, mem = c.leaves()
    got = {(int(a), *map(int, b)): int(n) for a, b, n in zip(lev, ijk, npart)}
    assert got == _tree_leaves_independent(cen, tree, nlev)
    assert sorted(mem.tolist()) == list(range(len(cen)))


@pytest.mark.parametrize("nlev", [1, 2])
def test_fmm_without_far_field_equals_direct(orc, nlev):
    """With 1 or 2 levels every cell is a neighbour: the FMM is pure P2P and
    must equal the direct sum (mod_wake.f90:1050-1113) to round-off."""
    d = np.load(os.path.join(HERE, "golden", "reg_test_b_particles.npz"))
    cen = d["cen"][:1500]
    alpha = d["alpha"][:1500]
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    rv = np.full(len(mag), 0.05)
    c = orc.Octree(6.0, (1, 1, 1), (-3.0, -3.0, -5.0), nlev, 5, 2, 0.005, use_vs=True)
    v, s, _ = c.run(cen, dr, mag, rv)
    vd, sd = orc.direct(cen, dr, mag, rv, True, False, 0.0)
    assert np.linalg.norm(v - vd) / np.linalg.norm(vd) < 1e-13
    assert np.linalg.norm(s - sd) / np.linalg.norm(sd) < 1e-13


def test_fmm_converges_with_order(orc):
    cen, alpha, rv, tree = W.helix_wake(8000, seed=4)
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    vd, sd = orc.direct(cen, dr, mag, rv)
    ev = []
    for p in (2, 4, 6):
        c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], 4, 64, p, tree["delta"])
        v, s, _ = c.run(cen, dr, mag, rv)
        ev.append(np.linalg.norm(v - vd) / np.linalg.norm(vd))
    assert ev[1] < 0.3 * ev[0] and ev[2] < 0.3 * ev[1] and ev[2] < 1e-4, ev


# ---- apply_multipole_panels (mod_octree.f90:1434-1510) ------------------------
def _panel_points(rng, cen, m, lo=-2.0, hi=2.0):
    """Points inside the particle cloud (near particles, so inside leaves),
    kept inside the box [lo, hi]^3."""
    p = cen[rng.choice(len(cen), m, replace=False)] + 0.01 * rng.standard_normal((m, 3))
    return np.clip(p, lo + 1e-3, hi - 1e-3)


@pytest.mark.parametrize("nlev", [1, 2])
def test_panels_without_far_field_equal_direct(orc, nlev):
    """One or two levels: every panel's leaf sees all particles through its
    neighbour lists, so the result is the direct get_vort_vel sum (plus the
    uvort it started from)."""
    d = np.load(os.path.join(HERE, "golden", "reg_test_b_particles.npz"))
    cen, alpha = d["cen"][:1500], d["alpha"][:1500]
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    rv = np.full(len(mag), 0.05)
    c = orc.Octree(6.0, (1, 1, 1), (-3.0, -3.0, -5.0), nlev, 5, 2, 0.005, use_vs=True)
    c.run(cen, dr, mag, rv)
    pts = _panel_points(np.random.default_rng(2), cen, 200, -3.0, 1.0)
    u0 = np.random.default_rng(3).standard_normal((200, 3))
    u = c.apply_panels(pts, u0)
    ud = orc.vort_vel(pts, cen, dr, mag, rv)
    assert np.linalg.norm(u - u0 - ud) / np.linalg.norm(ud) < 1e-13


def test_panels_converge_with_order(orc):
    cen, alpha, rv, tree = W.helix_wake(8000, seed=4)
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    pts = _panel_points(np.random.default_rng(5), cen, 300)
    ud = orc.vort_vel(pts, cen, dr, mag, rv)
    ev = []
    for p in (2, 4, 6):
        c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], 4, 64, p, tree["delta"])
        c.run(cen, dr, mag, rv)
        ev.append(np.linalg.norm(c.apply_panels(pts) - ud) / np.linalg.norm(ud))
    assert ev[1] < 0.3 * ev[0] and ev[2] < 0.3 * ev[1] and ev[2] < 1e-4, ev


def test_panels_outside_the_box(orc):
    """A centre outside the box is an error (use_fmm_pan, mod_octree.f90:785-788);
    inside, every point lies under one leaf (a level-1 cell whose children are
    all inactive is itself a leaf, :953-957), so a point far from the
    particles still receives their far field."""
    cen, alpha, rv, tree = W.helix_wake(3000, seed=6)
    mag = np.linalg.norm(alpha, axis=1)
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], 5, 64, 4, tree["delta"])
    c.run(cen, alpha / mag[:, None], mag, rv)
    with pytest.raises(ValueError):
        c.apply_panels(np.array([[100.0, 0.0, 0.0]]))
    corner = np.asarray(tree["origin"]) + 1e-3
    u = c.apply_panels(corner[None])
    ud = orc.vort_vel(corner[None], cen, alpha / mag[:, None], mag, rv)
    assert np.linalg.norm(u - ud) < 0.1 * np.linalg.norm(ud)  # p = 4 at a leaf corner
