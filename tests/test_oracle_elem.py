"""Pins of the CPU oracle's element <-> particle terms (oracle/elem_oracle.c)
against mathematics that does not share its code (CPU only):

- velocity_calc_sou_surfpan (src/aeroel/mod_surfpan.f90:268-357) is the
  gradient of the source potential potential_calc_sou_surfpan, itself pinned
  to quadrature (finite differences, near and far field);
- compute_vel_vortline (src/aeroel/mod_vortline.f90:126-170) is the
  Biot-Savart integral of the segment (Gauss-Legendre quadrature);
- a vortex ring (velocity_calc_doublet, mod_doublet.f90:178-248) is the sum
  of its edges as vortex lines, and so is its gradient;
- gradient_calc_doublet / compute_grad_vortline (mod_doublet.f90:254-305,
  mod_vortline.f90:174-236) are the Jacobian d v_r / d x_c (row = velocity
  component) of the segment velocity: the unregularised line velocity for a
  vanishing core, and a(l.R/|R|)/(a^2 + del^2)^1.5 (a = l x R) for a finite
  one -- a line's field is not curl-free, so a transposed gradient fails;
- a closed surface of constant-strength doublet panels induces no velocity
  off the surface (edges cancel pairwise), a closed surface of uniform
  sources the field of a point source outside it (Gauss);
- get_vort_vel (mod_surfpan.f90:962-977) sums compute_vel_vortpart, whose
  pair kernel is pinned in test_oracle_fmm.py, and the end-vortex lines;
- the stretching term of apply_multipole (mod_octree.f90:1380-1401) is
  transpose(grad) alpha / (4 pi), with grad pinned as above.
"""
import numpy as np
import pytest

QUAD = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.2, 0.9, 0.0], [0.1, 1.1, 0.0]])
TRI = np.array([[0.0, 0.0, 0.0], [1.0, 0.2, 0.0], [0.3, 0.8, 0.0]])
NEAR = [np.array(p) for p in ([0.3, 0.4, 0.5], [2.0, -1.0, 0.7], [0.5, 0.5, -0.3],
                              [-0.4, 0.2, 0.05], [0.6, 0.3, 0.2])]
FAR = [np.array([9.0, 7.0, -6.0]), np.array([-12.0, 3.0, 5.0])]
LINE = np.array([[0.2, -0.3, 0.1], [1.1, 0.5, -0.4]])


def panel(orc, ver):
    v = np.zeros((1, 4, 3))
    v[0, :len(ver)] = ver
    return orc.geo(v, np.array([len(ver)]))[0]


def fd_jacobian(f, p, h=1e-5):
    """J[r, c] = d f_r / d x_c by central differences."""
    return np.stack([(f(p + h * e) - f(p - h * e)) / (2 * h) for e in np.eye(3)], axis=1)


@pytest.mark.parametrize("ver", [QUAD, TRI], ids=["quad", "tri"])
def test_source_velocity_is_gradient_of_potential(orc, ver):
    g = panel(orc, ver)
    for p in NEAR + FAR:
        v = orc.vel_source(g, p)

        def pot(x):
            return orc.pot_source(g, orc.pot_doublet(g, x, 10.0), x)
        grad = np.array([(pot(p + 1e-5 * e) - pot(p - 1e-5 * e)) / 2e-5 for e in np.eye(3)])
        assert np.allclose(v, grad, rtol=1e-6, atol=1e-9), (p, v, grad)


def gauss_segment(x1, x2, p, n=200):
    s, w = np.polynomial.legendre.leggauss(n)
    L = np.linalg.norm(x2 - x1)
    l = (x2 - x1) / L
    y = x1[None] + np.outer(0.5 * (s + 1) * L, l)
    r = p[None] - y
    f = np.cross(l[None], r) / np.linalg.norm(r, axis=1)[:, None] ** 3
    return 0.5 * L * (w[:, None] * f).sum(axis=0)


def test_vortline_is_biot_savart_integral(orc):
    for p in NEAR + FAR:
        v = orc.vel_vortline(LINE, p, rankine=1e-6, cutoff=1e-9)
        assert np.allclose(v, gauss_segment(LINE[0], LINE[1], p), rtol=1e-9, atol=1e-12)


def test_vortline_core_and_cutoff(orc):
    x1, x2 = LINE
    l = (x2 - x1) / np.linalg.norm(x2 - x1)
    nrm = np.cross(l, [0.0, 0.0, 1.0])
    nrm /= np.linalg.norm(nrm)
    mid = 0.5 * (x1 + x2)
    L = np.linalg.norm(x2 - x1)
    # inside the Rankine core the 1/h^2 factor becomes 1/(rankine L)^2: about
    # linear in h, and continuous across the core radius
    v1 = orc.vel_vortline(LINE, mid + 0.01 * L * nrm, rankine=0.1, cutoff=0.001)
    v2 = orc.vel_vortline(LINE, mid + 0.02 * L * nrm, rankine=0.1, cutoff=0.001)
    assert np.allclose(v2, 2.0 * v1, rtol=2e-3)
    vin = orc.vel_vortline(LINE, mid + 0.1 * L * (1 - 1e-9) * nrm, rankine=0.1, cutoff=0.001)
    vout = orc.vel_vortline(LINE, mid + 0.1 * L * (1 + 1e-9) * nrm, rankine=0.1, cutoff=0.001)
    assert np.allclose(vin, vout, rtol=1e-7)
    # at an end point (inside the cut-off) the line induces nothing
    assert np.all(orc.vel_vortline(LINE, x1 + 1e-7 * l, rankine=0.1, cutoff=0.001) == 0.0)


@pytest.mark.parametrize("ver", [QUAD, TRI], ids=["quad", "tri"])
def test_ring_is_sum_of_edge_lines(orc, ver):
    g = panel(orc, ver)
    nv = len(ver)
    for p in NEAR:
        v = orc.vel_doublet(g, p, 1e9, 0.1, 0.001)
        gr = orc.grad_doublet(g, p, 0.05)
        vs = sum(orc.vel_vortline(np.array([ver[i], ver[(i + 1) % nv]]), p, 0.1, 0.001)
                 for i in range(nv))
        gs = sum(orc.grad_vortline(np.array([ver[i], ver[(i + 1) % nv]]), p, 0.05)
                 for i in range(nv))
        assert np.allclose(v, vs, rtol=1e-12, atol=1e-14)
        assert np.allclose(gr, gs, rtol=1e-12, atol=1e-14)


def test_line_gradient_is_velocity_jacobian(orc):
    for p in NEAR:
        J = fd_jacobian(lambda x: orc.vel_vortline(LINE, x, 1e-6, 1e-9), p)
        G = orc.grad_vortline(LINE, p, 0.0)
        assert np.allclose(G, J, rtol=1e-6, atol=1e-7), (G, J)
        assert not np.allclose(G, G.T, rtol=1e-3)  # the transposed convention would fail


def test_line_gradient_with_core(orc):
    """For del > 0 the gradient is the exact Jacobian of the regularised
    segment velocity sum_{k=1,2} (-1)^(k+1) a_k (l.R_k/|R_k|) |a| / (|a|^2 + del^2)^1.5."""
    x1, x2 = LINE
    l = (x2 - x1) / np.linalg.norm(x2 - x1)
    d = 0.3

    def v_reg(x):
        R1, R2 = x - x1, x - x2
        a = np.linalg.norm(np.cross(l, R1))
        h = a / (a * a + d * d) ** 1.5
        return h * (np.cross(l, R1) * l.dot(R1) / np.linalg.norm(R1)
                    - np.cross(l, R2) * l.dot(R2) / np.linalg.norm(R2))
    for p in NEAR:
        assert np.allclose(orc.grad_vortline(LINE, p, d), fd_jacobian(v_reg, p),
                           rtol=1e-6, atol=1e-8)


def test_ring_gradient_is_velocity_jacobian(orc):
    g = panel(orc, QUAD)
    for p in NEAR:
        J = fd_jacobian(lambda x: orc.vel_doublet(g, x, 1e9, 1e-6, 1e-9), p)
        G = orc.grad_doublet(g, p, 0.0)
        assert np.allclose(G, J, rtol=1e-6, atol=1e-7)
        assert abs(np.trace(G)) < 1e-9 * np.abs(G).max()  # solenoidal


def icosphere_quads():
    """A closed cube-sphere of quads (outward normals)."""
    n = 6
    faces = []
    t = np.linspace(-1, 1, n + 1)
    for ax in range(3):
        for s in (-1.0, 1.0):
            for i in range(n):
                for j in range(n):
                    q = []
                    for a, b in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)):
                        p = np.zeros(3)
                        p[ax] = s
                        p[(ax + 1) % 3] = t[a]
                        p[(ax + 2) % 3] = t[b]
                        q.append(p / np.linalg.norm(p))
                    if s < 0:
                        q = q[::-1]
                    faces.append(q)
    return np.array(faces)


def test_closed_doublet_surface_induces_nothing(orc):
    ver = icosphere_quads()
    grp = [dict(kind=orc.RING, ver=ver, n_ver=np.full(len(ver), 4), mag=np.full(len(ver), 0.7))]
    pts = np.array([[0.1, 0.2, -0.3], [1.7, 0.4, 0.2], [-3.0, 2.0, 1.0]])
    vel, _ = orc.elem_apply(grp, pts, ff_ratio=1e9, rankine=1e-6, cutoff=1e-9)
    one, _ = orc.elem_apply([dict(grp[0], ver=ver[:1], n_ver=[4], mag=[0.7])], pts,
                            ff_ratio=1e9, rankine=1e-6, cutoff=1e-9)
    assert np.abs(vel).max() < 1e-11 * np.abs(one).max() * len(ver)


def test_closed_source_surface_is_point_source(orc):
    """Uniform sigma on a closed surface: outside, the velocity is that of a
    point source of the total strength (compute_vel_surfpan = -vsou sigma)."""
    ver = icosphere_quads()
    G = orc.geo(ver, np.full(len(ver), 4))
    area = G[:, orc.G_AREA].sum()
    grp = [dict(kind=orc.SURFPAN, ver=ver, n_ver=np.full(len(ver), 4), mag=np.zeros(len(ver)),
                sigma=np.ones(len(ver)))]
    pts = np.array([[2.5, 0.4, 0.2], [-3.0, 2.0, 1.0], [0.0, 0.0, 6.0]])
    vel, _ = orc.elem_apply(grp, pts)
    cen = G[:, orc.G_CEN:orc.G_CEN + 3].mean(axis=0)
    for p, v in zip(pts, vel):
        r = p - cen
        expect = area * r / np.linalg.norm(r) ** 3 / (4 * np.pi)
        assert np.allclose(v, expect, rtol=5e-3)


def test_stretching_is_transposed_gradient(orc):
    ver = np.array([QUAD, QUAD + [0.0, 0.0, 0.5]])
    lines = np.array([LINE, LINE[::-1] + 0.3])
    pts = np.array(NEAR)
    alpha = np.random.default_rng(3).standard_normal((len(pts), 3))
    groups = [dict(kind=orc.RING, ver=ver, n_ver=[4, 4], mag=[0.4, -1.1]),
              dict(kind=orc.LINE, ver=lines, n_ver=[2, 2], mag=[0.9, 0.3]),
              dict(kind=orc.RING, ver=ver, n_ver=[4, 4], mag=[5.0, 5.0], stretch=False)]
    _, st = orc.elem_apply(groups, pts, alpha=alpha, vortex_rad=0.05)
    for k, p in enumerate(pts):
        G = sum(m * orc.grad_doublet(orc.geo(ver[i:i + 1], [4])[0], p, 0.05)
                for i, m in enumerate([0.4, -1.1]))
        G = G + sum(m * orc.grad_vortline(lines[i], p, 0.05) for i, m in enumerate([0.9, 0.3]))
        assert np.allclose(st[k], G.T @ alpha[k] / (4 * np.pi), rtol=1e-12, atol=1e-15)


def test_vort_vel_particles_and_lines(orc):
    rng = np.random.default_rng(11)
    cen = rng.uniform(-1, 1, (7, 3))
    pc = rng.uniform(-1, 1, (40, 3))
    pd = rng.standard_normal((40, 3))
    pm = rng.uniform(0.5, 2.0, 40)
    pr = rng.uniform(0.01, 0.1, 40)
    lv = np.array([LINE, LINE + 0.2])
    lm = np.array([0.8, -0.4])
    u = orc.vort_vel(cen, pc, pd, pm, pr, lv, lm)
    for i, x in enumerate(cen):
        up = sum(orc.pair(pc[j], pd[j], pm[j], pr[j], x, np.zeros(3))[0] for j in range(40))
        ul = sum(lm[k] * orc.vel_vortline(lv[k], x) for k in range(2))
        assert np.allclose(u[i], (up + ul) / (4 * np.pi), rtol=1e-12, atol=1e-15)
    # no strength associated with the lines: they induce nothing
    assert np.allclose(orc.vort_vel(cen, pc, pd, pm, pr, lv, None), orc.vort_vel(cen, pc, pd, pm, pr))


def test_particles_along_a_line_match_the_line(orc):
    """Many small particles with alpha = Gamma dl along a segment induce, away
    from it, the line's velocity (both kernels are Biot-Savart)."""
    x1, x2 = LINE
    n = 4000
    s = (np.arange(n) + 0.5) / n
    pc = x1[None] + np.outer(s, x2 - x1)
    L = np.linalg.norm(x2 - x1)
    pd = np.tile((x2 - x1) / L, (n, 1))
    pm = np.full(n, L / n)
    pts = np.array(FAR + [np.array([0.6, 1.5, 0.9])])
    u = orc.vort_vel(pts, pc, pd, pm, np.full(n, 1e-4))
    ul = orc.vort_vel(pts, lver=np.array([LINE]), lmag=np.array([1.0]))
    assert np.allclose(u, ul, rtol=1e-6)
