"""CPU oracle pipeline of a body + wake time step -- TEST INFRASTRUCTURE ONLY
(imported by tests/ and bench.py's CPU leg).  The sequence of
dust_amd.step.BodyWakeStep (DUST's time loop, src/main/dust.f90 537-1108)
with the oracle's restatements: get_vort_vel, initialize_linsys /
assemble_linsys with the implicit wake row and the explicit wake panels,
dgetrf/dgetrs, the octree FMM, the element loops (body panels, wake panels,
end vortices), the panel wake (update_wake / complete_wake panel parts), the
particle update and the shedding from the last panel row
(src/aeroel/mod_wake.f90, src/linsys/mod_linsys.f90, src/fmm/mod_octree.f90).

The trailing-edge tables and the rotation of the geometry are restated here
from their definitions (not imported from the product).
"""
import types

import numpy as np

from . import oracle as orc


def strip_trailing_edge(ver, strips, nc):
    """Trailing edge of a mesh of `strips` strips of 2 nc quads (upper panels
    from the trailing edge forward, then lower panels back): the stripe's
    generating panels are the first upper and the last lower panel of the
    strip, its points at vertex 1 and vertex 4 of the upper panel, each the
    midpoint of that vertex and the lower panel's vertex closest to it; a
    point shared by two stripes makes them neighbours, with orientation +1
    when it is the first point of one and the second of the other."""
    up = np.array([2 * nc * s for s in range(strips)])
    lo = up + (2 * nc - 1)
    coords, where = [], []
    start = np.zeros((strips, 2), np.int32)
    for s in range(strips):
        for side, vtx in enumerate((0, 3)):
            x = ver[up[s], vtx]
            for q, c in enumerate(coords):
                if np.array_equal(c, x):
                    start[s, side] = q + 1
                    break
            else:
                coords.append(x.copy())
                dist = [np.sqrt(((ver[lo[s], q] - x) ** 2).sum()) for q in range(4)]
                where.append((up[s], vtx, lo[s], int(np.argmin(dist))))
                start[s, side] = len(coords)
    neigh = np.zeros((strips, 2), np.int32)
    orient = np.zeros((strips, 2), np.int32)
    for s in range(strips):
        for side in range(2):
            for t in range(strips):
                if t == s:
                    continue
                for tside in range(2):
                    if start[t, tside] == start[s, side]:
                        neigh[s, side] = t + 1
                        orient[s, side] = 1 if tside != side else -1
    gen = np.stack([up + 1, lo + 1], axis=1).astype(np.int32)
    return dict(up=up, lo=lo, start=start, where=np.array(where), neigh=neigh, o=orient,
                gen=gen)


def te_rows_input(ver, te):
    """Trailing-edge points and first-row directions: at every point the sum
    over its stripes of (point - centre of the stripe's two generating
    panels), its component along the summed trailing-edge vectors removed,
    normalised."""
    pts = np.array([(ver[p, v] + ver[q, w]) / 2.0 for p, v, q, w in te["where"]])
    n = len(pts)
    d, e = np.zeros((n, 3)), np.zeros((n, 3))
    for s in range(len(te["up"])):
        p1, p2 = te["start"][s] - 1
        c = (ver[te["up"][s]].sum(axis=0) + ver[te["lo"][s]].sum(axis=0)) / 8.0
        for p in (p1, p2):
            d[p] += pts[p] - c
            e[p] += pts[p2] - pts[p1]
    e = e / np.sqrt((e * e).sum(axis=1))[:, None]
    d = d - (d * e).sum(axis=1)[:, None] * e
    return pts, d / np.sqrt((d * d).sum(axis=1))[:, None]


def rotated_z(x, theta):
    """x (..., 3) turned by theta about the z axis."""
    c, s = np.cos(theta), np.sin(theta)
    y = np.array(x, dtype=np.float64)
    y[..., 0] = c * x[..., 0] - s * x[..., 1]
    y[..., 1] = s * x[..., 0] + c * x[..., 1]
    return y


class OracleStep:
    """The same step with the CPU oracle, on host copies of the pool."""

    def __init__(self, case, strips, nmax, u_inf, dt, vortex_rad=0.1, omega=0.0,
                 fmm_pan=False, n_wake_panels=1, refine_wake=False, k_refine=1,
                 k_vortex_rad=0.0, join_te=False, join_te_factor=0.1, min_vel_at_te=1e-3,
                 solver="fp64", wind=None):
        ver, nv, cen, alpha, rv, tree = case
        # wind = (gust type, the 13 numbers of oracle.wind_params) or None:
        # variable_wind at the step's time where the reference calls it
        self.wind, self.t = wind, 0.0
        v4 = np.zeros((len(nv), 4, 3))
        v4[:, : np.asarray(ver).shape[1]] = np.asarray(ver)[:, :4]
        self.ver0, self.nv = v4, np.asarray(nv, np.int32)
        self.omega, self.theta, self.fmm_pan = float(omega), 0.0, fmm_pan
        self.moving = self.omega != 0.0
        self.u = np.asarray(u_inf, float)
        self.dt, self.tree, self.vr = dt, tree, vortex_rad
        lo = np.asarray(tree["origin"], float)
        self.box = (lo + 1e-9, lo + np.asarray(tree["nbox"]) * tree["box_length"] - 1e-9)
        self.te = strip_trailing_edge(v4, *strips)
        npt = len(self.te["where"])
        self.pw = orc.PanWake(self.te["start"], self.te["neigh"], self.te["o"], self.te["gen"],
                              npt, n_wake_panels, dt, min_vel=min_vel_at_te, join_te=join_te,
                              join_fact=join_te_factor, refine=refine_wake, k_refine=k_refine,
                              kvr=k_vortex_rad, vr=vortex_rad)
        self._set_geometry(v4)
        self._rows(True)
        n = len(nv)
        self.S = types.SimpleNamespace(
            rank=n, nstatic=0 if self.moving else n, kind=np.zeros(n, np.int32), ver=v4,
            n_ver=self.nv, ub=None, uvort=np.zeros((n, 3)), bern=None, u_inf=tuple(self.u),
            n_sp=0, nstatic_sp=0, id_sp=np.zeros(0, np.int32), n_expl=0, nstatic_expl=0,
            ff_ratio=10.0, rankine=0.1, cutoff=0.001)
        self._wake_to_system()
        self.st = orc.initialize_linsys(self.S)
        assert self.st["info"] == 0
        self.pool = dict(free=np.ones(nmax, np.uint8), cen=np.zeros((nmax, 3)),
                         dir=np.zeros((nmax, 3)), mag=np.zeros(nmax), rv=np.zeros(nmax),
                         vel=np.zeros((nmax, 3)), stretch=np.zeros((nmax, 3)),
                         rotu=np.zeros((nmax, 3)), n_prt=0)
        if len(rv):
            assert orc.wake_shed(self.pool, cen, alpha, rv, np.tile(self.u, (len(rv), 1)),
                                 *self.box) == len(rv)
        self.pool["n_prt"] = int((self.pool["free"] == 0).sum())

    def _set_geometry(self, v):
        self.ver = v
        self.G = orc.geo(v, self.nv)
        self.cen_p = self.G[:, orc.G_CEN:orc.G_CEN + 3]
        self.nor = self.G[:, orc.G_NOR:orc.G_NOR + 3]
        c = self.cen_p
        self.ub = np.stack([-self.omega * c[:, 1], self.omega * c[:, 0], np.zeros(len(c))], 1)

    def _wind(self, pts):
        return orc.variable_wind(self.wind[0], self.wind[1], pts, self.t)

    def _rows(self, first):
        pts, gdir = te_rows_input(self.ver, self.te)
        vte = np.stack([-self.omega * pts[:, 1], self.omega * pts[:, 0], np.zeros(len(pts))], 1)
        if self.wind is not None:
            # norm2(wind - vel_te) of the reference with the gust moved into vel_te
            vte = vte - (self._wind(pts) - self.u)
        self.pw.start(first, pts, gdir, np.ones(len(pts)), vte, self.u)

    def _wake_to_system(self):
        wv, wm = self.pw.panels()
        S = self.S
        S.wake_ver, S.wake_nver = np.ascontiguousarray(wv), np.full(len(wv), 4, np.int32)
        S.wake_mag = wm
        S.n_wake, S.n_impl = len(wv), len(self.te["up"])
        S.impl_ind = self.te["gen"]

    def _groups(self, mu, sigma):
        wv, wm = self.pw.panels()
        g = [dict(kind=orc.SURFPAN, ver=self.ver, n_ver=self.nv, mag=mu, sigma=sigma),
             dict(kind=orc.RING, ver=wv, n_ver=np.full(len(wv), 4, np.int32), mag=wm)]
        if self.pw.full:
            lver, lmag, _ = self.pw.end_vortices()
            g.append(dict(kind=orc.LINE, ver=lver, mag=lmag))
        return g

    def step(self):
        P, S = self.pool, self.S
        live = orc.wake_live(P["free"])
        pc, pd, pm, pr = P["cen"][live], P["dir"][live], P["mag"][live], P["rv"][live]
        t = self.tree
        fmm = orc.Octree(t["box_length"], t["nbox"], t["origin"], t["nlevels"], t["min_part"], 4,
                         t["delta"], u_inf=tuple(self.u))
        if self.wind is not None:
            fmm.set_wind(self.wind[0], self.wind[1], self.t)
        wc = self._wind(self.cen_p) if self.wind is not None else self.u[None]
        lver = lmag = None
        if self.pw.full:
            lver, lmag, _ = self.pw.end_vortices()
        # assemble_linsys
        if self.fmm_pan:
            fmm.sort(pc, pd, pm, pr)
            fmm.multipole()
            uvort = fmm.apply_panels(self.cen_p)
            if lver is not None:
                uvort = uvort + orc.vort_vel(self.cen_p, lver=lver, lmag=lmag)
        else:
            uvort = orc.vort_vel(self.cen_p, pc, pd, pm, pr, lver=lver, lmag=lmag)
        S.uvort = uvort if self.wind is None else uvort + (wc - self.u)
        S.ub = self.ub if self.moving else None
        S.ver = self.ver
        self._wake_to_system()
        b, _ = orc.assemble_linsys(S, self.st)
        # solve_linsys
        n = S.rank
        if self.moving:
            LU, ipiv, info = orc.dgetrf(self.st["A"])
            assert info == 0
            mu = orc.dgetrs(LU, ipiv, b)
        else:
            mu = b.copy()
            orc.lib().ao_dgetrs(n, 1, orc._p(self.st["A"]), n, orc._p(self.st["ipiv"]),
                                orc._p(mu), n)
        # update_wake
        g1, g2 = self.te["gen"][:, 0] - 1, self.te["gen"][:, 1] - 1
        self.pw.idou[0] = mu[g1] - np.where(g2 >= 0, mu[np.maximum(g2, 0)], 0.0)
        sigma = np.einsum("ij,ij->i", self.nor, self.ub - wc - uvort)
        groups = self._groups(mu, sigma)
        pts = self.pw.points()
        vpt = orc.vort_vel(pts, pc, pd, pm, pr) + (self.u if self.wind is None else self._wind(pts))
        vpt, _ = orc.elem_apply(groups, pts, vel=vpt, vortex_rad=self.vr)
        if self.fmm_pan:
            vel, st, _ = fmm.apply()
        else:
            vel, st, _ = fmm.run(pc, pd, pm, pr)
        vel, _ = orc.elem_apply(groups, pc, vel=vel, vortex_rad=self.vr)
        P["vel"][live] = vel
        P["stretch"][live] = st
        self.pw.update(mu, vpt)
        # geometry update, complete_wake
        if self.moving:
            self.theta += self.omega * self.dt
            self._set_geometry(rotated_z(self.ver0, self.theta))
        self._rows(False)
        orc.wake_update(P, *self.box, self.dt, use_vs=True)
        P["n_prt"] = int((P["free"] == 0).sum())
        assert self.pw.shed(P, *self.box) == 0
        self.t += self.dt
        return dict(mu=mu, uvort=uvort, vel=vel, stretch=st, live=live)
