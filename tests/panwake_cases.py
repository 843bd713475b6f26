"""Panel-wake configurations for the oracle and GPU tests: a straight
trailing edge of ns stripes along y, a full last panel row set by hand, and
two separate trailing edges that can be joined."""
import numpy as np


def tables(ns, offset=0, left=0, right=0):
    """1-based start points, neighbours (+1 orientation) and generating
    elements (upper 1..ns, lower ns+1..2ns) of a straight edge."""
    start = np.array([[offset + i + 1, offset + i + 2] for i in range(ns)], np.int32)
    neigh = np.array([[i if i > 0 else left, i + 2 if i < ns - 1 else right]
                      for i in range(ns)], np.int32)
    neigh[:, :] = np.where(neigh > 0, neigh + (offset if offset else 0) * 0, 0)
    o = np.where(neigh > 0, 1, 0).astype(np.int32)
    gen = np.array([[i + 1, ns + i + 1] for i in range(ns)], np.int32)
    return start, neigh, o, gen


def straight(orc, ns=3, nmax=4, dt=0.1, ndt=1, min_vel=1e-3, **kw):
    start, neigh, o, gen = tables(ns)
    np_ = ns + 1
    w = orc.PanWake(start, neigh, o, gen, np_, nmax, dt, ndt=ndt, min_vel=min_vel, **kw)
    y = np.linspace(-1.0, 1.0, np_)
    spts = np.stack([np.full(np_, 1.0), y, np.zeros(np_)], 1)
    gdir = np.tile([2.0, 0.0, -0.2], (np_, 1)) + 0.1 * np.stack([np.zeros(np_), y, y], 1)
    scal = np.linspace(0.8, 1.2, np_)
    vte = np.stack([np.zeros(np_), 0.1 * y, np.zeros(np_)], 1)
    return w, spts, gdir, scal, vte


def full_last_row(orc, ns, end, last, chord=1.0, span=0.5, dz=0.01, **kw):
    """A panel wake whose last row is full: stripes of the given chord and span
    along y, end and last intensities as given."""
    start, neigh, o, gen = tables(ns)
    np_ = ns + 1
    nmax = 2
    w = orc.PanWake(start, neigh, o, gen, np_, nmax, 0.1, **kw)
    y = span * np.arange(np_)
    w.pts[nmax] = np.stack([np.full(np_, 3.0), y, np.zeros(np_)], 1)
    w.pend[:] = np.stack([np.full(np_, 3.0 + chord), y, np.full(np_, dz)], 1)
    w.pvel[nmax] = np.tile([10.0, 0.0, 0.0], (np_, 1))
    w.end_idou[:] = end
    w.last_idou[:] = last
    w.idou[nmax - 1] = end
    w.s.len, w.s.full = nmax, 1
    st = start - 1
    return w, w.pts[nmax][st[:, 0]], w.pts[nmax][st[:, 1]], w.pend[st[:, 0]], w.pend[st[:, 1]]


def pool(n):
    return dict(free=np.ones(n, np.uint8), cen=np.zeros((n, 3)), dir=np.zeros((n, 3)),
                mag=np.zeros(n), rv=np.zeros(n), vel=np.zeros((n, 3)), n_prt=0)


def two_edges(orc, gap=0.01, join_fact=0.5, nmax=3):
    """Two one-stripe trailing edges along y, ends facing each other across
    `gap`."""
    start = np.array([[1, 2], [3, 4]], np.int32)
    neigh = np.zeros((2, 2), np.int32)
    o = np.zeros((2, 2), np.int32)
    gen = np.array([[1, 0], [2, 0]], np.int32)
    w = orc.PanWake(start, neigh, o, gen, 4, nmax, 0.1, join_te=True, join_fact=join_fact)
    spts = np.array([[1.0, -1.0, 0.0], [1.0, -gap / 2, 0.0], [1.0, gap / 2, 0.0],
                     [1.0, 1.0, 0.0]])
    gdir = np.array([[1.0, 0.0, 0.0], [1.0, 0.1, 0.0], [1.0, -0.1, 0.0], [1.0, 0.0, 0.0]])
    return w, spts, gdir, np.ones(4), np.zeros((4, 3))
