"""The Gaussian particle core (the commented alternative of kernel_coeffs,
src/aeroel/mod_vortpart.f90:283-291) in the oracle, pinned without its own
formula: the velocity of a Gaussian vortex blob is the singular Biot-Savart
velocity times the blob's mass inside the radius, a radial integral done by
quadrature; the stretching coefficient d follows from c through the
transposed velocity gradient (finite differences of the oracle's own
velocity); far from the core it is the singular kernel, and the FMM without
far field equals the direct sum."""
import os

import numpy as np
import pytest
from scipy import integrate

HERE = os.path.dirname(os.path.abspath(__file__))


def blob_mass(r, s):
    """Mass of the unit 3-D Gaussian of deviation s inside radius r."""
    f = lambda q: 4.0 * np.pi * q * q * np.exp(-q * q / (2.0 * s * s)) / (2.0 * np.pi * s * s) ** 1.5
    return integrate.quad(f, 0.0, r, epsabs=0.0, epsrel=1e-13, limit=200)[0]


@pytest.mark.parametrize("rho", [0.1, 0.5, 1.0, 2.0, 3.5, 6.0])
def test_velocity_is_biot_savart_times_blob_mass(orc, rho):
    s = 0.02
    d = np.array([0.3, -0.4, np.sqrt(1 - 0.25)])
    x = np.array([1.0, 2.0, -0.5]) / np.sqrt(5.25) * rho * s
    v, _, _ = orc.pair(np.zeros(3), d, 1.7, s, x, np.zeros(3), kernel=orc.GAUSSIAN)
    r = np.linalg.norm(x)
    # returned times 4 pi: u = alpha x (x - x_s) M(r) / r^3
    ref = np.cross(1.7 * d, x) * blob_mass(r, s) / r ** 3
    tol = 1e-12 if rho >= 0.5 else 1e-9  # the closed form cancels as rho^2 at small rho
    assert np.allclose(v, ref, rtol=tol, atol=0)


def test_stretching_is_transposed_gradient_of_gaussian_velocity(orc):
    rng = np.random.default_rng(3)
    for _ in range(10):
        xs, d, at = (rng.standard_normal(3) for _ in range(3))
        d /= np.linalg.norm(d)
        s = rng.uniform(0.05, 0.5)
        x = xs + rng.standard_normal(3) * s * rng.uniform(0.5, 3.0)
        mag, h = rng.uniform(0.5, 2.0), 1e-6
        G = np.zeros((3, 3))
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            vp = orc.pair(xs, d, mag, s, x + e, at, kernel=orc.GAUSSIAN)[0]
            vm = orc.pair(xs, d, mag, s, x - e, at, kernel=orc.GAUSSIAN)[0]
            G[i] = (vp - vm) / (2 * h)
        st = orc.pair(xs, d, mag, s, x, at, kernel=orc.GAUSSIAN)[1]
        assert np.allclose(st, G @ at, rtol=1e-6, atol=1e-8 * np.abs(G).max())


def test_far_field_and_coincident_limits(orc):
    s, d = 0.02, np.array([0.0, 0.0, 1.0])
    x = np.array([0.3, 0.1, 0.2])  # 18.7 core radii
    v, st, _ = orc.pair(np.zeros(3), d, 1.0, s, x, np.ones(3), kernel=orc.GAUSSIAN)
    r = np.linalg.norm(x)
    assert np.allclose(v, np.cross(d, x) / r ** 3, rtol=1e-14, atol=0)
    v0, s0, _ = orc.pair(np.zeros(3), d, 1.0, s, np.full(3, 1e-15), np.ones(3),
                         kernel=orc.GAUSSIAN)
    assert not v0.any() and not s0.any()
    # the switch is scoped: the Rosenhead core afterwards
    vr, _, _ = orc.pair(np.zeros(3), d, 1.0, s, x, np.zeros(3))
    assert np.allclose(vr, np.cross(d, x) / (r * r + s * s) ** 1.5, rtol=1e-13, atol=0)


def test_gaussian_fmm_without_far_field_equals_direct(orc):
    d = np.load(os.path.join(HERE, "golden", "reg_test_b_particles.npz"))
    cen, alpha = d["cen"][:1500], d["alpha"][:1500]
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    rv = np.full(len(mag), 0.05)
    c = orc.Octree(6.0, (1, 1, 1), (-3.0, -3.0, -5.0), 2, 5, 2, 0.005, use_vs=True,
                   kernel=orc.GAUSSIAN)
    v, s, _ = c.run(cen, dr, mag, rv)
    vd, sd = orc.direct(cen, dr, mag, rv, True, False, 0.0, kernel=orc.GAUSSIAN)
    vr, _ = orc.direct(cen, dr, mag, rv, True, False, 0.0)
    assert np.linalg.norm(v - vd) / np.linalg.norm(vd) < 1e-13
    assert np.linalg.norm(s - sd) / np.linalg.norm(sd) < 1e-13
    assert np.linalg.norm(vr - vd) / np.linalg.norm(vd) > 1e-3  # the cores differ
