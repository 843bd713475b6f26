"""Pins of the CPU oracle's particle-pool steps (oracle/wake_oracle.c) by
properties that do not restate its loops (CPU only):

- the particle loop of complete_wake (src/aeroel/mod_wake.f90:1305-1355):
  free slots untouched; a live particle either moves by exactly vel dt and
  stays inside the box, or its slot is freed (and it is not moved); mag
  unchanged; the new direction is the unit vector along dir mag + stretch dt
  (or unchanged when that vanishes); the divergence filter leaves stretch
  alone when rotu is parallel to dir and otherwise subtracts
  filt_eta mag (dir - rotu/|rotu|);
- prepare_wake's list (:826-835) is the ascending list of non-free slots;
- shedding (:1888-1967): the inside particles occupy, in order, the lowest
  free slots; outside ones take none; a full pool is an error;
- avoid_collision (:2350-2534): untouched far away, below a surface panel or
  moving away; a particle crossing a surface panel within dt lands on its
  plane after dt*ndt with its speed kept; inside the layer the damped
  correction; the vortex-lattice branch's elementwise normal distance and
  its layer thickness doubling per panel, each from a hand-worked case.
"""
import numpy as np
import pytest

BOX = (np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, 1.0]))


def pool(n, seed, free_frac=0.3):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return dict(free=(rng.random(n) < free_frac).astype(np.uint8),
                cen=rng.uniform(-0.95, 0.95, (n, 3)), dir=d, mag=rng.uniform(0.1, 1.0, n),
                rv=rng.uniform(0.01, 0.05, n), vel=rng.standard_normal((n, 3)),
                stretch=rng.standard_normal((n, 3)), rotu=rng.standard_normal((n, 3)))


def copy(p):
    return {k: v.copy() for k, v in p.items()}


@pytest.mark.parametrize("divfilt", [False, True])
def test_update_properties(orc, divfilt):
    p0 = pool(2000, 1)
    p = copy(p0)
    dt, eta = 0.05, 0.3
    nfreed = orc.wake_update(p, *BOX, dt, eta, use_vs=True, use_divfilt=divfilt)
    was_free = p0["free"] == 1
    for k in ("cen", "dir", "stretch"):
        assert np.array_equal(p[k][was_free], p0[k][was_free])
    assert np.array_equal(p["mag"], p0["mag"])
    live = ~was_free
    moved = p0["cen"] + p0["vel"] * dt
    inside = np.all((moved >= BOX[0]) & (moved <= BOX[1]), axis=1)
    assert nfreed == np.count_nonzero(live & ~inside) > 0
    assert np.all(p["free"][live & ~inside] == 1) and np.all(p["free"][live & inside] == 0)
    assert np.array_equal(p["cen"][live & inside], moved[live & inside])
    assert np.array_equal(p["cen"][live & ~inside], p0["cen"][live & ~inside])
    sel = live & inside
    s = p0["stretch"].copy()
    if divfilt:
        r = p0["rotu"] / np.linalg.norm(p0["rotu"], axis=1)[:, None]
        s = s - eta * p0["mag"][:, None] * (p0["dir"] - r)
        assert np.allclose(p["stretch"][sel], s[sel], rtol=1e-14, atol=1e-15)
    a = p0["dir"] * p0["mag"][:, None] + s * dt
    u = a / np.linalg.norm(a, axis=1)[:, None]
    assert np.allclose(p["dir"][sel], u[sel], rtol=1e-14, atol=1e-15)
    assert np.allclose(np.linalg.norm(p["dir"][sel], axis=1), 1.0, rtol=1e-14)


def test_divergence_filter_parallel_rotu_is_neutral(orc):
    p = pool(500, 2, free_frac=0.0)
    p["rotu"] = 3.7 * p["dir"]
    p["vel"][:] = 0.0
    s0 = p["stretch"].copy()
    orc.wake_update(p, *BOX, 0.01, 0.5, use_vs=True, use_divfilt=True)
    assert np.allclose(p["stretch"], s0, rtol=0, atol=1e-15)


def test_update_without_stretching_only_moves(orc):
    p0 = pool(500, 3)
    p = copy(p0)
    orc.wake_update(p, *BOX, 0.02, use_vs=False, use_vd=False)
    assert np.array_equal(p["dir"], p0["dir"]) and np.array_equal(p["stretch"], p0["stretch"])


def test_zero_alpha_keeps_direction(orc):
    p = pool(10, 4, free_frac=0.0)
    p["vel"][:] = 0.0
    dt = 0.1
    p["stretch"] = -p["dir"] * p["mag"][:, None] / dt   # alpha = 0 exactly
    d0 = p["dir"].copy()
    orc.wake_update(p, *BOX, dt, use_vs=True)
    a = d0 * p["mag"][:, None] + p["stretch"] * dt
    zero = np.all(a == 0.0, axis=1)
    assert zero.any()
    assert np.array_equal(p["dir"][zero], d0[zero])


def test_live_list(orc):
    free = (np.random.default_rng(5).random(1000) < 0.4).astype(np.uint8)
    assert np.array_equal(orc.wake_live(free), np.flatnonzero(free == 0))


def test_shed_takes_lowest_free_slots(orc):
    p0 = pool(300, 6, free_frac=0.2)
    p = copy(p0)
    rng = np.random.default_rng(7)
    k = 40
    ncen = rng.uniform(-1.2, 1.2, (k, 3))
    npv = rng.standard_normal((k, 3))
    npv[3] = 1e-15  # below the threshold: dir = partvec
    nrv = rng.uniform(0.01, 0.1, k)
    nvel = rng.standard_normal((k, 3))
    added = orc.wake_shed(p, ncen, npv, nrv, nvel, *BOX)
    inside = np.all((ncen >= BOX[0]) & (ncen <= BOX[1]), axis=1)
    slots = np.flatnonzero(p0["free"] == 1)[: np.count_nonzero(inside)]
    assert added == len(slots)
    assert np.array_equal(np.flatnonzero((p0["free"] == 1) & (p["free"] == 0)), slots)
    src = np.flatnonzero(inside)
    assert np.array_equal(p["cen"][slots], ncen[src])
    assert np.array_equal(p["vel"][slots], nvel[src])
    assert np.array_equal(p["rv"][slots], nrv[src])
    m = np.linalg.norm(npv[src], axis=1)
    assert np.allclose(p["mag"][slots], m, rtol=1e-15)
    big = m > 1e-13
    assert np.allclose(np.linalg.norm(p["dir"][slots][big], axis=1), 1.0, rtol=1e-15)
    assert np.array_equal(p["dir"][slots][~big], npv[src][~big])
    # untouched elsewhere
    other = np.ones(300, bool)
    other[slots] = False
    for key in ("cen", "dir", "mag", "vel", "rv"):
        assert np.array_equal(p[key][other], p0[key][other])


def test_shed_full_pool_is_an_error(orc):
    p = pool(50, 8, free_frac=0.0)
    p["free"][[3, 9]] = 1
    k = 3
    assert orc.wake_shed(p, np.zeros((k, 3)), np.ones((k, 3)), np.full(k, 0.1), np.zeros((k, 3)),
                         *BOX) == -1


# ---- avoid_collision (mod_wake.f90:2350-2534) --------------------------------
PA = dict(dt=0.01, dtw=0.02, u_ref=10.0, rad_mult=1.0, elrad_mult=1.0)


def flat(kind, z=0.0, maxedge=0.4):
    return (np.array([kind]), np.array([[0.0, 0.0, z]]), np.array([[0.0, 0.0, 1.0]]),
            np.array([maxedge]), np.zeros((1, 3)))


def test_collision_far_or_leaving_is_untouched(orc):
    el = flat(0)
    v = np.array([0.3, -0.2, -2.0])
    far = orc.avoid_collision(*el, np.array([5.0, 0.0, 0.05]), 0.02, v, **PA)
    assert np.array_equal(far, v)
    below = orc.avoid_collision(*el, np.array([0.05, 0.0, -0.05]), 0.02, v, **PA)
    assert np.array_equal(below, v)
    away = orc.avoid_collision(*el, np.array([0.05, 0.0, 0.05]), 0.02, -v, **PA)
    assert np.array_equal(away, -v)


def test_collision_lands_on_the_panel_with_the_same_speed(orc):
    """Surface panel: a particle that would cross the panel within dt gets the
    normal velocity that brings it exactly onto the plane in dt*ndt, and the
    tangential part is rescaled so that the speed is unchanged."""
    el = flat(0)
    h, w = 0.01, 2.0                       # crosses after h/w = 0.005 < dt
    v = np.array([0.3, -0.2, -w])
    out = orc.avoid_collision(*el, np.array([0.05, 0.02, h]), 0.02, v, **PA)
    assert out[2] == pytest.approx(-h / PA["dtw"], rel=1e-14)
    assert np.linalg.norm(out) == pytest.approx(np.linalg.norm(v), rel=1e-14)
    assert np.allclose(out[:2] / np.linalg.norm(out[:2]), v[:2] / np.linalg.norm(v[:2]))


def test_collision_damped_inside_the_layer(orc):
    """Surface panel, not reached within dt but ending inside the layer of
    thickness r_vortex: normvel + (b - d2)/dtw * d2/b."""
    el = flat(0)
    b, h, w = 0.02, 0.025, 1.0             # d2 = h - w dt = 0.015 <= b
    v = np.array([0.0, 0.0, -w])
    out = orc.avoid_collision(*el, np.array([0.0, 0.0, h]), b, v, **PA)
    d2 = h - w * PA["dt"]
    assert out[2] == pytest.approx(-w + (b - d2) / PA["dtw"] * (d2 / b), rel=1e-13)
    assert out[0] == 0.0 and out[1] == 0.0


def test_vortex_lattice_layer_doubles_per_panel(orc):
    """Vortex-lattice branch: the layer thickness doubles each time the branch
    reaches its test (the reference's local blthick).  With the particle
    ending at d2 in (2b, 4b], the first of two coincident panels leaves the
    velocity alone and the second corrects it with thickness 4b."""
    b, h, w = 0.01, 0.04, 1.0              # d2 = 0.03 in (0.02, 0.04]
    el = tuple(np.concatenate([a, a]) for a in flat(1))
    v = np.array([0.0, 0.0, -w])
    out = orc.avoid_collision(*el, np.array([0.0, 0.0, h]), b, v, **PA)
    d2 = h - w * PA["dt"]
    assert out[2] == pytest.approx(-w + (4 * b - d2) / PA["dtw"] * (d2 / (4 * b)), rel=1e-13)
    one = orc.avoid_collision(*flat(1), np.array([0.0, 0.0, h]), b, v, **PA)
    assert np.array_equal(one, v)


def test_vortex_lattice_normal_distance_is_elementwise_norm(orc):
    """Vortex-lattice branch: the normal distance is |dist * n| (elementwise),
    so with a tilted normal it differs from the dot product; a particle on the
    far side is corrected too (the test is on the sign of (dist.n)(v.n))."""
    n = np.array([0.0, 0.6, 0.8])
    el = (np.array([1]), np.zeros((1, 3)), n[None], np.array([0.4]), np.zeros((1, 3)))
    pos = -0.01 * n                         # below, moving up towards the panel
    v = 2.0 * n
    out = orc.avoid_collision(*el, pos, 0.02, v, **PA)
    d1 = np.linalg.norm(pos * n)            # elementwise, as the reference
    assert np.dot(out, n) == pytest.approx(-d1 / PA["dtw"], rel=1e-13)
