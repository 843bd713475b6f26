"""CPU oracle pipeline of a body + wake time step -- TEST INFRASTRUCTURE ONLY
(imported by tests/ and bench.py's CPU leg).  The same sequence as
dust_amd.step.BodyWakeStep with the oracle's restatements: get_vort_vel,
the static right-hand side, dgetrf/dgetrs, the octree FMM, the body panels'
velocity, the particle update and the shedding insertion
(src/aeroel/mod_wake.f90, src/linsys/mod_linsys.f90, src/fmm/mod_octree.f90).
"""
import numpy as np

from . import oracle as orc


class OracleStep:
    """The same step with the CPU oracle, on host copies of the pool."""

    def __init__(self, case, te, nmax, u_inf, dt, sub=1, r_shed=0.1, omega=0.0, strips=None,
                 fmm_pan=False):
        ver, nv, cen, alpha, rv, tree = case
        self.omega, self.theta, self.strips, self.sub = float(omega), 0.0, strips, sub
        self.fmm_pan = fmm_pan
        v4 = np.zeros((len(nv), 4, 3))
        v4[:, : np.asarray(ver).shape[1]] = np.asarray(ver)[:, :4]
        self.ver0 = v4
        self.orc = orc
        self.r_shed = r_shed
        self.G = orc.geo(ver, nv)
        self.cen_p = self.G[:, orc.G_CEN:orc.G_CEN + 3]
        self.nor = self.G[:, orc.G_NOR:orc.G_NOR + 3]
        A, self.B = orc.aic_surfpan(self.G)
        self.LU, self.ipiv, info = orc.dgetrf(A)
        assert info == 0
        self.ver, self.nv = ver, nv
        self.u = np.asarray(u_inf, float)
        self.dt = dt
        self.tree = tree
        lo = np.asarray(tree["origin"], float)
        self.box = (lo + 1e-9, lo + np.asarray(tree["nbox"]) * tree["box_length"] - 1e-9)
        self._set_te(te)
        self.pool = dict(free=np.ones(nmax, np.uint8), cen=np.zeros((nmax, 3)),
                         dir=np.zeros((nmax, 3)), mag=np.zeros(nmax), rv=np.zeros(nmax),
                         vel=np.zeros((nmax, 3)), stretch=np.zeros((nmax, 3)),
                         rotu=np.zeros((nmax, 3)))
        assert orc.wake_shed(self.pool, cen, alpha, rv, np.tile(self.u, (len(rv), 1)),
                             *self.box) == len(rv)

    def _set_te(self, te):
        up, lo_i, te0, te1 = te
        sub = self.sub
        q = (np.arange(sub) + 0.5) / sub
        self.up, self.lo = np.repeat(up, sub), np.repeat(lo_i, sub)
        self.te_vec = np.repeat((te1 - te0) / sub, sub, axis=0)
        self.te_mid = (np.repeat(te0, sub, axis=0) + np.tile(q, len(up))[:, None] *
                       np.repeat(te1 - te0, sub, axis=0) + 0.5 * self.dt * self.u)

    def _move(self):
        from dust_amd.step import trailing_edge
        from dust_amd.workloads import rotate_z
        self.theta += self.omega * self.dt
        v = rotate_z(self.ver0, self.theta)
        self.ver = v
        self.G = orc.geo(v, self.nv)
        self.cen_p = self.G[:, orc.G_CEN:orc.G_CEN + 3]
        self.nor = self.G[:, orc.G_NOR:orc.G_NOR + 3]
        A, self.B = orc.aic_surfpan(self.G)
        self.LU, self.ipiv, info = orc.dgetrf(A)
        assert info == 0
        ns, nc = self.strips
        self._set_te(trailing_edge(v, ns, nc))
        c = self.cen_p
        return np.stack([-self.omega * c[:, 1], self.omega * c[:, 0], np.zeros(len(c))], axis=1)

    def step(self):
        orc, P = self.orc, self.pool
        live = orc.wake_live(P["free"])
        pc, pd, pm, pr = P["cen"][live], P["dir"][live], P["mag"][live], P["rv"][live]
        ub = self._move() if self.omega != 0.0 else np.zeros_like(self.cen_p)
        t = self.tree
        fmm = orc.Octree(t["box_length"], t["nbox"], t["origin"], t["nlevels"], t["min_part"], 4,
                         t["delta"], u_inf=tuple(self.u))
        if self.fmm_pan:
            fmm.sort(pc, pd, pm, pr)
            fmm.multipole()
            uvort = fmm.apply_panels(self.cen_p)
        else:
            uvort = orc.vort_vel(self.cen_p, pc, pd, pm, pr)
        b = orc.rhs_static(self.G, self.B, self.u, uvort - ub)
        mu = orc.dgetrs(self.LU, self.ipiv, b)
        sigma = np.einsum("ij,ij->i", self.nor, ub - self.u[None] - uvort)
        if self.fmm_pan:
            vel, st, _ = fmm.apply()
        else:
            vel, st, _ = fmm.run(pc, pd, pm, pr)
        grp = [dict(kind=orc.SURFPAN, ver=self.ver, n_ver=self.nv, mag=mu, sigma=sigma)]
        vel, _ = orc.elem_apply(grp, pc, vel=vel)
        P["vel"][live] = vel
        P["stretch"][live] = st
        orc.wake_update(P, *self.box, self.dt, use_vs=True)
        pv = (mu[self.up] - mu[self.lo])[:, None] * self.te_vec
        k = len(pv)
        assert orc.wake_shed(P, self.te_mid, pv, np.full(k, self.r_shed), np.tile(self.u, (k, 1)),
                             *self.box) >= 0
        return dict(mu=mu, uvort=uvort, vel=vel, stretch=st, live=live)


