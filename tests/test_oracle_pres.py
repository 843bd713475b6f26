"""Pins of the CPU oracle's pressure right-hand side (oracle/pres_oracle.c,
assemble_pressure_sys, src/linsys/mod_pressure_equation.f90:139-236) by
hand-worked values and properties that do not restate its loops (CPU only):

- a particle: d.(dir x vel) mag / (|d|^2 + r^2)^1.5 from hand-worked cases,
  zero for vel parallel to dir, odd under the exchange of dir and vel,
  linear in mag;
- an end-vortex line: a hand-worked symmetric case;
- a wake panel: the sum of its edges taken as lines (vertex i to i+1, with
  the two vertices' velocities), quads and triangles -- which pins the edge
  loop and its wrap against the hand-pinned line;
- b_pres is accumulated (b0 + terms).
"""
import numpy as np
import pytest


def test_particle_hand_worked(orc):
    # d = (0,0,2), dir x vel = x x y = z, d.z = 2, |d|^3 = 8
    b = orc.pres_rhs_rot([[0.0, 0.0, 0.0]], pcen=[[0.0, 0.0, 2.0]], pdir=[[1.0, 0.0, 0.0]],
                         pmag=[0.5], prv=[0.0], pvel=[[0.0, 1.0, 0.0]])
    assert b[0] == pytest.approx(2.0 * 0.5 / 8.0, rel=1e-15)
    # core radius 1: (4 + 1)^1.5
    b = orc.pres_rhs_rot([[0.0, 0.0, 0.0]], pcen=[[0.0, 0.0, 2.0]], pdir=[[1.0, 0.0, 0.0]],
                         pmag=[0.5], prv=[1.0], pvel=[[0.0, 1.0, 0.0]])
    assert b[0] == pytest.approx(2.0 * 0.5 / 5.0 ** 1.5, rel=1e-15)
    # the target's offset enters as pcen - cen
    b = orc.pres_rhs_rot([[1.0, 2.0, 3.0]], pcen=[[1.0, 2.0, 1.0]], pdir=[[1.0, 0.0, 0.0]],
                         pmag=[1.0], prv=[0.0], pvel=[[0.0, 1.0, 0.0]])
    assert b[0] == pytest.approx(-2.0 / 8.0, rel=1e-15)


def test_particle_properties(orc):
    rng = np.random.default_rng(3)
    n, m = 50, 7
    cen = rng.standard_normal((m, 3))
    pc, pd, pv = (rng.standard_normal((n, 3)) for _ in range(3))
    pm, pr = rng.uniform(0.1, 1.0, n), rng.uniform(0.0, 0.2, n)
    b = orc.pres_rhs_rot(cen, pcen=pc, pdir=pd, pmag=pm, prv=pr, pvel=pv)
    swapped = orc.pres_rhs_rot(cen, pcen=pc, pdir=pv, pmag=pm, prv=pr, pvel=pd)
    assert np.allclose(swapped, -b, rtol=1e-14, atol=0.0)
    twice = orc.pres_rhs_rot(cen, pcen=pc, pdir=pd, pmag=2.0 * pm, prv=pr, pvel=pv)
    assert np.allclose(twice, 2.0 * b, rtol=1e-14, atol=0.0)
    par = orc.pres_rhs_rot(cen, pcen=pc, pdir=pd, pmag=pm, prv=pr, pvel=3.0 * pd)
    assert np.abs(par).max() <= 1e-14 * np.abs(b).max()  # zero up to the cross's rounding
    b0 = rng.standard_normal(m)
    acc = orc.pres_rhs_rot(cen, b=b0, pcen=pc, pdir=pd, pmag=pm, prv=pr, pvel=pv)
    assert np.allclose(acc, b0 + b, rtol=1e-13, atol=1e-15)


def test_line_hand_worked(orc):
    """Line (-1,0,1) -> (1,0,1), both vertex velocities y, target at the
    origin: d1 x y = (-1,0,-1), d2 x y = (-1,0,1), |d|^3 = 2 sqrt 2, e = (2,0,0):
    e.(...) = -sqrt 2, and b -= mag/2 (-sqrt 2)."""
    lver = np.array([[[-1.0, 0.0, 1.0], [1.0, 0.0, 1.0]]])
    lvel = np.array([[[0.0, 1.0, 0.0], [0.0, 1.0, 0.0]]])
    b = orc.pres_rhs_rot([[0.0, 0.0, 0.0]], lver=lver, lmag=[1.5], lvel=lvel)
    assert b[0] == pytest.approx(0.75 * np.sqrt(2.0), rel=1e-15)
    # not associated: nothing
    b = orc.pres_rhs_rot([[0.0, 0.0, 0.0]], lver=lver, lmag=None, lvel=lvel)
    assert b[0] == 0.0


@pytest.mark.parametrize("nv", [4, 3])
def test_wake_panel_is_its_edges(orc, nv):
    rng = np.random.default_rng(nv)
    m, nw = 9, 5
    cen = rng.standard_normal((m, 3))
    wver = np.zeros((nw, 4, 3))
    wvel = np.zeros((nw, 4, 3))
    wver[:, :nv] = 3.0 + rng.standard_normal((nw, nv, 3))
    wvel[:, :nv] = rng.standard_normal((nw, nv, 3))
    wmag = rng.standard_normal(nw)
    b = orc.pres_rhs_rot(cen, wver=wver, wnver=np.full(nw, nv), wmag=wmag, wvel=wvel)
    ref = np.zeros(m)
    for w in range(nw):
        for i in range(nv):
            j = (i + 1) % nv
            ref = orc.pres_rhs_rot(cen, b=ref, lver=[[wver[w, i], wver[w, j]]], lmag=[wmag[w]],
                                   lvel=[[wvel[w, i], wvel[w, j]]])
    assert np.allclose(b, ref, rtol=1e-13, atol=1e-15)
