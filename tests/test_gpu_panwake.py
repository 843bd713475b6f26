"""GPU parity of the panel wake (dust_panwake_* through the C ABI) against
the oracle over several steps on identical inputs: a straight trailing edge
with neighbours, two trailing edges joined by join_te, and refined shedding.
Each step: update (solution, velocities at the points), complete (new
trailing-edge rows), shedding into a particle pool.  Points, velocities,
intensities and the pool agree to 1e-13 (relative max norm)."""
import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import panwake as PW

import panwake_cases as PC

pytestmark = pytest.mark.gpu


class Pool:
    def __init__(self, torch, n):
        self.free = torch.ones(n, dtype=torch.uint8, device="cuda")
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device="cuda")
        self.cen, self.dir, self.vel = z(n, 3), z(n, 3), z(n, 3)
        self.mag, self.rv = z(n), z(n)
        self.n_prt = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.status = torch.zeros(1, dtype=torch.int32, device="cuda")


def relmax(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


@pytest.mark.parametrize("case", ["straight", "joined", "refined"])
def test_panel_wake_vs_oracle(orc, case):
    torch = pytest.importorskip("torch")
    nmax, steps = 3, 8
    if case == "joined":
        _, spts, gdir, scal, vte = PC.two_edges(orc, gap=0.01, join_fact=0.5, nmax=nmax)
        start = np.array([[1, 2], [3, 4]], np.int32)
        neigh, o = np.zeros((2, 2), np.int32), np.zeros((2, 2), np.int32)
        gen = np.array([[1, 0], [2, 0]], np.int32)
        kw = dict(join_te=True, join_fact=0.5)
        nel = 2
    else:
        ns = 5
        start, neigh, o, gen = PC.tables(ns)
        _, spts, gdir, scal, vte = PC.straight(orc, ns=ns, nmax=nmax)
        kw = dict(refine=True, k_refine=2, kvr=0.2) if case == "refined" else {}
        nel = 2 * ns
    npt = len(spts)
    dt = 0.02
    ow = orc.PanWake(start, neigh, o, gen, npt, nmax, dt, ndt=2, **kw)
    gw = PW.PanWake(start, neigh, o, gen, npt, nmax, dt, ndt=2, **kw)
    rng = np.random.default_rng(31)
    u = np.array([10.0, 0.5, -0.3])
    dv = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    ow.start(True, spts, gdir, scal, vte, u)
    gw.start(True, dv(spts), dv(gdir), dv(scal), dv(vte), u)
    nmaxp = 400
    opool = PC.pool(nmaxp)
    gpool = Pool(torch, nmaxp)
    import ctypes as C
    prm = L.WakeParams()
    prm.box_min = (C.c_double * 3)(-1e9, -1e9, -1e9)
    prm.box_max = (C.c_double * 3)(1e9, 1e9, 1e9)
    for k in range(steps):
        print(f"panel wake {case} step {k}", flush=True)
        mu = rng.normal(size=nel)
        n, _ = gw.points()
        assert n == ow.len * npt
        vel = u + rng.normal(scale=0.5, size=(n, 3))
        ow.update(mu, vel)
        gw.update(dv(mu), dv(vel))
        sp = spts + 0.01 * k * np.array([0.0, 0.0, 1.0])
        ow.start(False, sp, gdir, scal, vte, u)
        gw.start(False, dv(sp), dv(gdir), dv(scal), dv(vte), u)
        ow.shed(opool, prm.box_min[:], prm.box_max[:])
        gw.shed(gpool, prm, gpool.n_prt, gpool.status)
        g = gw.get()
        assert gw.state()[:2] == (ow.len, ow.full)
        for key in ("pts", "pvel", "idou", "end_idou", "last_idou", "pend"):
            a, b = g[key], getattr(ow, key)
            if np.abs(b).max() > 0:
                assert relmax(a, b) <= 1e-13, (key, relmax(a, b))
            else:
                assert np.abs(a).max() == 0, key
        # the panels and end vortices handed to the assembly and the elements
        pn = gw.panels()
        assert pn["n_wake"] == ow.len * len(start) and pn["n_impl"] == len(start)
        free = gpool.free.cpu().numpy()
        assert np.array_equal(free, opool["free"])
        assert int(gpool.n_prt.item()) == opool["n_prt"]
        live = free == 0
        if live.any():
            for key, t in (("cen", gpool.cen), ("dir", gpool.dir), ("mag", gpool.mag),
                           ("rv", gpool.rv), ("vel", gpool.vel)):
                assert relmax(t.cpu().numpy()[live], opool[key][live]) <= 1e-13, key
    assert int(gpool.status.item()) == 0
    assert ow.full and opool["n_prt"] > 0
