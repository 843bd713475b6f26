"""Pins of the oracle's variable_wind (src/aeroel/mod_wind.f90:66-121) and
of its use as the FMM's free-stream term (mod_octree.f90:1376), CPU only.

- no gust: the uniform u_inf;
- 'ACM' gust, a one-minus-cosine profile behind a front that moves with
  front_speed along the normalised front_direction: u_inf + u_des in every
  component at penetration s = gradient, u_des / 2 at s = gradient / 2,
  nothing ahead of the front (s < 0) or from s = 2 gradient on; the front
  reaches a point at time start + distance / speed; the direction's length
  does not matter;
- 'linear': u_inf + (0, 0, 0.1 x);
- the FMM's velocities with a gust differ from those without by exactly
  variable_wind - u_inf at each particle (to round-off).
"""
import numpy as np
import pytest

from dust_amd import workloads as W

U = np.array([30.0, 1.0, -2.0])


def test_no_gust_is_uniform(orc):
    pos = np.random.default_rng(1).normal(size=(20, 3))
    g = orc.wind_params(U, front_speed=5.0, u_des=3.0, gradient=0.5)
    assert np.array_equal(orc.variable_wind(0, g, pos, 0.3), np.tile(U, (20, 1)))


def test_acm_profile_behind_a_moving_front(orc):
    H, ud, c, t0 = 0.4, 2.5, 10.0, 0.1
    g = orc.wind_params(U, origin=(1.0, 0.0, 0.0), front_direction=(3.0, 0.0, 0.0), front_speed=c,
                        u_des=ud, gradient=H, start_time=t0)
    t = 0.25                               # front at x = 1 + c (t - t0) = 2.5
    xf = 1.0 + c * (t - t0)
    s = np.array([-0.3, 0.0, H / 2, H, 1.5 * H, 2 * H - 1e-9, 2 * H, 3 * H])
    pos = np.stack([xf - s, np.full_like(s, 0.7), np.full_like(s, -0.2)], 1)
    w = orc.variable_wind(1, g, pos, t) - U
    assert np.allclose(w[:, 0], w[:, 1]) and np.allclose(w[:, 0], w[:, 2])   # scalar added
    expect = [0.0, 0.0, ud / 2, ud, ud / 2, 0.0, 0.0, 0.0]
    assert np.allclose(w[:, 0], expect, atol=1e-12)
    # the front reaches x = 4 at time t0 + (4 - 1) / c
    p = np.array([[4.0 - 1e-9, 0.0, 0.0]])
    assert np.allclose(orc.variable_wind(1, g, p, t0 + 3.0 / c - 1e-6), U, atol=0)
    assert orc.variable_wind(1, g, np.array([[4.0 - H, 0, 0]]), t0 + 3.0 / c)[0, 0] == \
        pytest.approx(U[0] + ud, abs=1e-12)


def test_linear_form(orc):
    pos = np.random.default_rng(2).normal(size=(10, 3))
    w = orc.variable_wind(2, orc.wind_params(U), pos, 1.0)
    assert np.allclose(w, U + np.outer(pos[:, 0], [0.0, 0.0, 0.1]), rtol=0, atol=1e-15)


def test_fmm_free_stream_is_the_wind(orc):
    cen, alpha, rv, tree = W.helix_wake(3000, seed=6)
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    x0 = float(np.median(cen[:, 0]))
    g = orc.wind_params(U, origin=(x0, 0, 0), front_direction=(1, 0, 0), front_speed=2.0, u_des=1.5,
                        gradient=0.3, start_time=0.0)

    def run(gust):
        t = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], 4, 64, 4, tree["delta"],
                       u_inf=tuple(U))
        if gust:
            t.set_wind(1, g, 0.05)
        return t.run(cen, dr, mag, rv)[0]

    d = run(True) - run(False)
    wind = orc.variable_wind(1, g, cen, 0.05) - U
    assert np.count_nonzero(wind[:, 0]) > 100
    assert np.allclose(d, wind, rtol=0, atol=1e-13)
