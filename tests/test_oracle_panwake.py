"""The oracle's restatement of the panel wake (oracle/panwake_oracle.c),
pinned by hand-worked cases: the second row of points from the trailing
edge, intensities travelling back one row per step, points convected by
v dt, a uniform sheet shedding only its tip vortices, the closed-loop
identity of a single stripe (partvec = Gamma (P1 - P2)), the refined
sub-particles adding up to the stripe's vortex vector, the junction of two
trailing edges (join_te) and a full particle pool."""
import numpy as np
import pytest

import panwake_cases as PC


def test_second_row_of_points(orc):
    w, spts, gdir, scal, vte = PC.straight(orc, ns=3, nmax=4, dt=0.1, ndt=2)
    u = np.array([10.0, 0.0, 1.0])
    w.start(True, spts, gdir, scal, vte, u)
    assert np.array_equal(w.pts[0], spts)
    rel = np.linalg.norm(u - vte, axis=1)
    expect = spts + gdir * (scal * rel * 0.2 / np.linalg.norm(gdir, axis=1))[:, None]
    assert np.allclose(w.pts[1], expect, rtol=0, atol=1e-14)
    # below min_vel_at_te: the initial row is not normalised by |dir|, the
    # later ones are (the reference's two branches)
    w2, spts, gdir, scal, vte = PC.straight(orc, ns=3, nmax=4, dt=0.1, ndt=2, min_vel=50.0)
    w2.start(True, spts, gdir, scal, vte, u)
    assert np.allclose(w2.pts[1], spts + gdir * (scal * 50.0 * 0.2)[:, None], atol=1e-13)
    w2.start(False, spts, gdir, scal, vte, u)
    n = np.linalg.norm(gdir, axis=1)
    assert np.allclose(w2.pts[1], spts + gdir * (scal * 50.0 * 0.2 / n)[:, None], atol=1e-13)


def test_intensities_travel_back_one_row(orc):
    nmax = 4
    w, spts, gdir, scal, vte = PC.straight(orc, ns=3, nmax=nmax)
    w.start(True, spts, gdir, scal, vte, (10.0, 0.0, 0.0))
    g = []
    ends = []
    for k in range(1, 8):
        mu = np.arange(1, 7, dtype=float) * k        # upper elements 1..3, lower 4..6
        g.append(mu[:3] - mu[3:])
        vel = np.zeros((w.len * w.s.np, 3))
        before = w.idou[nmax - 1].copy()
        w.update(mu, vel)
        assert w.len == min(k + 1, nmax)
        assert np.array_equal(w.idou[0], g[-1])
        for r in range(1, w.len):
            assert np.array_equal(w.idou[r], g[max(len(g) - r, 0)] if len(g) - r >= 1 else g[0])
        if w.full:
            ends.append((w.end_idou.copy(), before))
    # once full, the intensity leaving the last row is the one that sat there
    assert ends and all(np.array_equal(e, b) for e, b in ends)


def test_points_convected(orc):
    w, spts, gdir, scal, vte = PC.straight(orc, ns=2, nmax=5, dt=0.05, ndt=3)
    w.start(True, spts, gdir, scal, vte, (10.0, 0.0, 0.0))
    V = np.array([7.0, -1.0, 0.5])
    for k in range(6):
        old = w.pts.copy()
        w.update(np.zeros(4), np.tile(V, (w.len * w.s.np, 1)))
        np1 = min(w.len + 1, 6)
        for row in range(3, np1 + 1):  # 1-based rows
            dtr = 0.05 if row == 3 else 0.15
            assert np.allclose(w.pts[row - 1], old[row - 2] + V * dtr, atol=1e-13)
        w.start(False, spts, gdir, scal, vte, (10.0, 0.0, 0.0))
    assert w.full
    assert np.allclose(w.pend, old[5] + V * 0.15, atol=1e-13)


def test_uniform_sheet_sheds_tip_vortices(orc):
    ns, gam = 4, 2.5
    w, P1, P2, E1, E2 = PC.full_last_row(orc, ns, end=np.full(ns, gam), last=np.full(ns, gam))
    pool = PC.pool(64)
    assert w.shed(pool, (-1e9,) * 3, (1e9,) * 3) == 0
    assert pool["n_prt"] == ns
    pv = pool["dir"][:ns] * pool["mag"][:ns, None]
    assert np.abs(pv[1:-1]).max() <= 1e-14
    assert np.allclose(pv[0], gam * (P1[0] - E1[0]), atol=1e-14)
    assert np.allclose(pv[-1], gam * (E2[-1] - P2[-1]), atol=1e-14)


def test_single_stripe_closed_loop(orc):
    gam = 1.5
    w, P1, P2, E1, E2 = PC.full_last_row(orc, 1, end=np.array([gam]), last=np.zeros(1),
                                        kvr=0.3)
    pool = PC.pool(4)
    w.shed(pool, (-1e9,) * 3, (1e9,) * 3)
    pv = pool["dir"][0] * pool["mag"][0]
    assert np.allclose(pv, gam * (P1[0] - P2[0]), atol=1e-14)   # left + right + end sides
    assert np.allclose(pool["cen"][0], (P1[0] + P2[0] + E1[0] + E2[0]) / 4, atol=1e-15)
    area = np.linalg.norm(np.cross(E1[0] - P2[0], E2[0] - P2[0]))
    assert pool["rv"][0] == pytest.approx(0.3 * np.sqrt(2 * area), rel=1e-15)
    assert w.last_idou[0] == gam


def test_refined_particles_add_up(orc):
    gam = 1.0
    w, P1, P2, E1, E2 = PC.full_last_row(orc, 1, end=np.array([gam]), last=np.zeros(1),
                                        chord=1.0, span=0.25, dz=0.0, refine=True, k_refine=2)
    pool = PC.pool(64)
    w.shed(pool, (-1e9,) * 3, (1e9,) * 3)
    n_chord, n_span = 8, 2   # ceil(1 / (0.25 / 2)), k_refine
    assert pool["n_prt"] == n_chord * n_span
    pv = (pool["dir"][:16] * pool["mag"][:16, None]).sum(0)
    assert np.allclose(pv, gam * (P1[0] - P2[0]), atol=1e-14)
    cen = pool["cen"][:16].reshape(n_chord, n_span, 3)
    ic, js = np.meshgrid(np.arange(1, 9) - 0.5, np.arange(1, 3) - 0.5, indexing="ij")
    expect = (P1[0] + (ic / 8)[..., None] * (E1[0] - P1[0]) + (js / 2)[..., None] *
              (P2[0] - P1[0]))
    assert np.allclose(cen, expect, atol=1e-14)


def test_join_te(orc):
    """Two separate trailing edges whose facing ends are closer than the
    tolerance: their first two rows of points meet half way, and the shed
    intensity of the joined sides is the oriented mean."""
    w, spts, gdir, scal, vte = PC.two_edges(orc, gap=0.01, join_fact=0.5)
    w.start(True, spts, gdir, scal, vte, (10.0, 0.0, 0.0))
    assert np.allclose(w.pts[0, 1], w.pts[0, 2]) and np.allclose(w.pts[1, 1], w.pts[1, 2])
    assert np.allclose(w.pts[0, 1], 0.5 * (spts[1] + spts[2]))
    assert w.joined[0, 0, 1] == 2 and w.joined[0, 1, 0] == 1
    w2, spts2, gdir2, scal2, vte2 = PC.two_edges(orc, gap=0.5, join_fact=0.5)
    w2.start(True, spts2, gdir2, scal2, vte2, (10.0, 0.0, 0.0))
    assert not w2.joined.any()


def test_full_pool_is_reported(orc):
    w, *_ = PC.full_last_row(orc, 3, end=np.ones(3), last=np.zeros(3))
    pool = PC.pool(2)
    assert w.shed(pool, (-1e9,) * 3, (1e9,) * 3) == -4


def test_joined_intensity_is_oriented_mean(orc):
    w, spts, gdir, scal, vte = PC.two_edges(orc, gap=0.01, join_fact=0.5, nmax=3)
    w.start(True, spts, gdir, scal, vte, (10.0, 0.0, 0.0))
    nmax = 3
    w.joined[nmax] = w.joined[0]          # the junction of the panels now at the end
    y = spts[:, 1]
    w.pts[nmax] = np.stack([np.full(4, 4.0), y, np.zeros(4)], 1)
    w.pend[:] = np.stack([np.full(4, 5.0), y, np.zeros(4)], 1)
    w.idou[nmax - 1] = [3.0, 1.0]         # intensities of the last row (get_joined_intensity)
    w.end_idou[:] = [2.0, 0.5]
    w.s.len, w.s.full = nmax, 1
    pool = PC.pool(8)
    w.shed(pool, (-1e9,) * 3, (1e9,) * 3)
    P1, P2, E1, E2 = w.pts[nmax][0], w.pts[nmax][1], w.pend[0], w.pend[1]
    # stripe 0: its free left side takes the last row's own intensity (3), its
    # right side the mean with stripe 1 of the same orientation ((3 - 1) / 2),
    # the end side end - last
    expect = (P1 - E1) * 3.0 + (E2 - P2) * 1.0 + (E1 - E2) * 2.0
    assert np.allclose(pool["dir"][0] * pool["mag"][0], expect, atol=1e-14)
    # stripe 1: joined left side (1 - 3) / 2, free right side 1
    Q1, Q2, F1, F2 = w.pts[nmax][2], w.pts[nmax][3], w.pend[2], w.pend[3]
    expect = (Q1 - F1) * -1.0 + (F2 - Q2) * 1.0 + (F1 - F2) * 0.5
    assert np.allclose(pool["dir"][1] * pool["mag"][1], expect, atol=1e-14)
