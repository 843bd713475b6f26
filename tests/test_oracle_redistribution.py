"""Pins of the oracle's particle redistribution (sort_particles,
mod_octree.f90:827-894), CPU only.

Tree: unit box, 4 levels (finest cells 0.125), lvl_solid = 3 (cells 0.25).
One element at (0.1, 0.1, 0.1) puts the level-3 cells with indices 1..2 in
every direction near the wall, i.e. the region [0, 0.5)^3.

- hand-worked single cells: the freed particle's share (mag * dir) / nprt
  reaches the others (a fraction 1/nprt of it is dropped, as in the
  reference); the walk sees the magnitudes the earlier removals raised (a
  particle first below the threshold survives); a vector that cancels to zero
  gets dir = 0 and is freed next;
- the near-wall neighbourhood at lvl_solid: cells at Chebyshev distance 1
  from the element's cell are left alone, distance 2 is not; an element
  outside the box marks nothing;
- the tree after a redistribution is the tree of the survivors: leaves,
  membership and the FMM results equal those of a sort without
  redistribution over the kept particles with their new strengths.
"""
import numpy as np
import pytest

ELEM = np.array([[0.1, 0.1, 0.1]])
FAR = (0.8125, 0.8125, 0.8125)    # centre of a finest cell far from the wall


def tree(orc, nlevels=4, min_part=1, degree=2):
    return orc.Octree(1.0, (1, 1, 1), (0.0, 0.0, 0.0), nlevels, min_part, degree, 0.01)


def cell_particles(centre, n, spread=0.03, seed=0):
    rng = np.random.default_rng(seed)
    return np.asarray(centre) + rng.uniform(-spread, spread, size=(n, 3))


def run(orc, cen, dir, mag, elem=ELEM, lvl=3, ratio=3.0):
    t = tree(orc)
    t.set_redistribution(lvl, ratio, elem)
    t.sort(cen, dir, mag, np.full(len(mag), 0.02))
    return t, t.redistributed()


def test_freed_share_goes_to_the_others(orc):
    cen = cell_particles(FAR, 3)
    dir = np.array([[1.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]])
    mag = np.array([1.0, 1.0, 0.1])
    # mean 0.7, threshold 0.7 / 3: the third is freed; share (0, 0.1, 0) / 3
    _, (keep, d, m, nk) = run(orc, cen, dir, mag)
    assert keep.tolist() == [True, True, False] and nk == 2
    s = 0.1 / 3
    for i in (0, 1):
        assert m[i] == pytest.approx(np.sqrt(1.0 + s * s), rel=1e-15)
        assert d[i] == pytest.approx(np.array([1.0, s, 0.0]) / np.sqrt(1.0 + s * s), rel=1e-15)
    # the freed particle keeps its strength; total vorticity loses 1/3 of it
    assert m[2] == 0.1 and d[2].tolist() == [0, 1.0, 0]
    tot = (d[:2] * m[:2, None]).sum(axis=0)
    assert tot == pytest.approx([2.0, 2 * s, 0.0], rel=1e-15)


def test_walk_sees_raised_magnitudes(orc):
    cen = cell_particles(FAR, 4, seed=1)
    dir = np.tile([0.0, 0.0, 1.0], (4, 1))
    mag = np.array([0.1, 0.17, 1.0, 1.0])
    # mean 0.5675, threshold 0.18917: the first is freed, its share 0.025
    # lifts the second to 0.195, above the threshold: it stays
    _, (keep, d, m, nk) = run(orc, cen, dir, mag)
    assert keep.tolist() == [False, True, True, True] and nk == 3
    assert m[1:] == pytest.approx([0.195, 1.025, 1.025], rel=1e-15)
    assert np.array_equal(d[1:], np.tile([0.0, 0.0, 1.0], (3, 1)))


def test_zero_vector_gets_zero_direction_and_is_freed(orc):
    cen = cell_particles(FAR, 4, seed=2)
    dir = np.array([[1.0, 0, 0], [-1.0, 0, 0], [0, 1.0, 0], [0, 1.0, 0]])
    mag = np.array([0.3, 0.075, 5.0, 5.0])
    # the first is freed (share (0.075, 0, 0), exactly 0.3 / 4): the second
    # cancels to zero, dir = 0, and is freed in turn with a zero share
    _, (keep, d, m, nk) = run(orc, cen, dir, mag)
    assert keep.tolist() == [False, False, True, True] and nk == 2
    assert m[1] == 0.0 and d[1].tolist() == [0, 0, 0]
    big = np.sqrt(25.0 + 0.075 ** 2)
    assert m[2:] == pytest.approx([big, big], rel=1e-15)
    assert d[2] == pytest.approx(np.array([0.075, 5.0, 0]) / big, rel=1e-15)


@pytest.mark.parametrize("centre,untouched", [
    ((0.3125, 0.3125, 0.3125), True),    # level-3 cell (2,2,2): a neighbour of (1,1,1)
    ((0.5625, 0.0625, 0.0625), False),   # level-3 cell (3,1,1): two cells away
    ((0.0625, 0.0625, 0.0625), True),    # the element's own cell
])
def test_near_wall_neighbourhood(orc, centre, untouched):
    cen = cell_particles(centre, 4, spread=0.02, seed=3)
    dir = np.tile([0.0, 1.0, 0.0], (4, 1))
    mag = np.array([0.1, 0.17, 1.0, 1.0])
    _, (keep, d, m, nk) = run(orc, cen, dir, mag)
    if untouched:
        assert keep.all() and np.array_equal(m, mag) and np.array_equal(d, dir)
    else:
        assert keep.tolist() == [False, True, True, True]


def test_element_outside_the_box_marks_nothing(orc):
    cen = cell_particles((0.0625, 0.0625, 0.0625), 4, spread=0.02, seed=4)
    dir = np.tile([0.0, 1.0, 0.0], (4, 1))
    mag = np.array([0.1, 0.17, 1.0, 1.0])
    _, (keep, _, _, _) = run(orc, cen, dir, mag, elem=np.array([[2.0, 2.0, 2.0]]))
    assert keep.tolist() == [False, True, True, True]
    _, (keep, _, _, _) = run(orc, cen, dir, mag, elem=None)
    assert keep.tolist() == [False, True, True, True]


def test_ratio_bounds_the_survivors(orc):
    rng = np.random.default_rng(7)
    n = 4000
    cen = rng.uniform(0.001, 0.999, size=(n, 3))
    a = rng.normal(size=(n, 3))
    dir = a / np.linalg.norm(a, axis=1)[:, None]
    mag = rng.lognormal(0.0, 1.5, size=n)
    for ratio in (1.5, 3.0, 10.0):
        _, (keep, d, m, nk) = run(orc, cen, dir, mag, ratio=ratio)
        assert nk == keep.sum() and 0 < nk < n
        near = np.all(cen < 0.5, axis=1)
        assert keep[near].all() and np.array_equal(m[near], mag[near])
        assert np.allclose(np.linalg.norm(d[keep], axis=1), 1.0, atol=1e-15)
    # a larger ratio frees fewer particles
    counts = [run(orc, cen, dir, mag, ratio=r)[1][3] for r in (1.5, 3.0, 10.0)]
    assert counts[0] < counts[1] < counts[2]


def test_tree_after_redistribution_is_the_tree_of_the_survivors(orc):
    rng = np.random.default_rng(11)
    n = 3000
    cen = rng.uniform(0.001, 0.999, size=(n, 3))
    a = rng.normal(size=(n, 3))
    dir = a / np.linalg.norm(a, axis=1)[:, None]
    mag = rng.lognormal(0.0, 1.0, size=n)
    rv = np.full(n, 0.02)
    t = orc.Octree(1.0, (1, 1, 1), (0.0, 0.0, 0.0), 4, 8, 4, 0.01)
    t.set_redistribution(3, 3.0, ELEM)
    v1, s1, _ = t.run(cen, dir, mag, rv)
    keep, d, m, nk = t.redistributed()
    lev1, ijk1, np1, mem1 = t.leaves()
    assert 0 < nk < n and not v1[~keep].any()
    idx = np.flatnonzero(keep)
    u = orc.Octree(1.0, (1, 1, 1), (0.0, 0.0, 0.0), 4, 8, 4, 0.01)
    v2, s2, _ = u.run(cen[idx], d[idx], m[idx], rv[idx])
    lev2, ijk2, np2, mem2 = u.leaves()
    assert np.array_equal(lev1, lev2) and np.array_equal(ijk1, ijk2)
    assert np.array_equal(np1, np2) and np.array_equal(mem1[:nk], idx[mem2])
    assert np.array_equal(v1[idx], v2) and np.array_equal(s1[idx], s2)
