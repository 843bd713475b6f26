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
  free slots; outside ones take none; a full pool is an error.
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
