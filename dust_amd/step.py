"""One time step of a body (at rest, or rotating about z) with a vortex-particle
wake, device-resident,
in the order of DUST's time loop (src/main/dust.f90 via src/aeroel/mod_wake.f90
and src/linsys/mod_linsys.f90):

  prepare_wake      live list of the particle pool, particle arrays for the
                    FMM and the element terms (mod_wake.f90:785-849)
  assemble_linsys   uvort at the panel centres from the particles
                    (get_vort_vel, mod_surfpan.f90:962-977) and the static
                    right-hand side b = B nor.(-u_inf - uvort)
                    (mod_surfpan.f90:400-404)
  solve_linsys      b -> panel intensities mu with the LU factors of the
                    static AIC computed once (mod_linsys.f90:236, 473)
  update_wake       particle velocity and stretching: the FMM pass
                    (sort_particles, calculate_multipole, apply_multipole)
                    plus the body panels' velocity (mag = mu, source
                    sigma = nor.(-u_inf - uvort)) (mod_octree.f90:1350-1373)
  complete_wake     particle update (Euler, box culling) and insertion of the
                    particles shed at the trailing edge (mod_wake.f90:1305-1355,
                    1888-1967)

The shedding geometry of the reference (last wake-panel row) is replaced by a
rule of the same shape: `sub` particles per trailing-edge segment, half a
step downstream of the midpoints of its `sub` equal pieces, partvec =
(mu_upper - mu_lower) times the piece's vector, radius `r_shed`, velocity
u_inf.  A rotating body (omega != 0) turns by omega dt at the start of each
step (geometry update of the caller), its AIC is assembled again and solved
with a fresh factorisation (FP64 getrf/getrs, or the mixed solve with
solver="fp32" / "fp16s"), and the panels' velocity omega x r enters the
source strengths.  With fmm_pan (the reference's use_fmm_pan) the tree and
expansions are built first and uvort comes from apply_multipole_panels on
them instead of the direct get_vort_vel sum.  Every stage runs on the GPU
through
the C ABI; the host reads back the live count and the trailing-edge
intensities (2 x segments values) once per step.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import _lib as L
from . import elements as E
from . import octree as O
from . import wake as WK
from . import workloads as WL


def trailing_edge(ver, strips, nc):
    """Trailing-edge segments of a mesh built as `strips` strips of 2*nc quads
    (nc upper panels from the trailing edge forward, then nc lower panels
    back to it, as workloads.wing_case / workloads.rotor lay them out):
    (upper panel index, lower panel index, segment start, segment end)."""
    up = np.arange(strips) * 2 * nc
    lo = up + 2 * nc - 1
    return up, lo, ver[up, 0].copy(), ver[up, 3].copy()


class BodyWakeStep:
    def __init__(self, ver, n_ver, te, u_inf, dt, nmax, tree, degree=4, form="fp64",
                 aic=None, r_shed=0.1, box=None, vs_elems=False, sub=1, omega=0.0,
                 solver="fp64", strips=None, fmm_pan=False, overlap=False):
        import torch
        self.torch = torch
        self.lib = L.lib()
        self.n = len(n_ver)
        self.u_inf = np.asarray(u_inf, dtype=np.float64)
        self.dt = float(dt)
        self.form = form
        self.vs_elems = bool(vs_elems)
        self.aic = dict(ff_ratio_doublet=10.0, rankine_rad=0.1, cutoff_rad=0.001)
        self.aic.update(aic or {})
        # the tuned form also evaluates the element far fields in FP32
        self.eprm = E.params(vortex_rad=r_shed, precision="tuned" if form == "tuned" else "fp64",
                             **self.aic)
        dv = lambda a, dt=np.float64: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).cuda()
        v4 = np.zeros((self.n, 4, 3))
        v4[:, : np.asarray(ver).shape[1]] = np.asarray(ver)[:, :4]
        self.ver0 = v4
        self.nv_h = np.asarray(n_ver, dtype=np.int32)
        self.ver = dv(v4)
        self.nv = dv(n_ver, np.int32)
        self.cen = dv(self._centres(v4))
        self.omega = float(omega)
        self.theta = 0.0
        self.solver = solver
        self.strips = strips  # (strips, nc) of the trailing-edge layout, moving bodies
        self.fmm_pan = bool(fmm_pan)
        # the FMM pass depends only on the particles: with overlap it runs on
        # its own (non-blocking) stream from a second host thread while this
        # thread assembles and solves the panel system.  Off by default: both
        # sides keep the chip busy on their own, so at 1M particles the step
        # gains under 1 %, and at a few thousand particles the thread start
        # costs more than it hides
        self.overlap = bool(overlap) and not self.fmm_pan
        n = self.n
        # static AIC and its LU factors (initialize_linsys, mod_linsys.f90:236)
        self.A = torch.empty((n, n), dtype=torch.float64, device="cuda")
        self.B = torch.empty((n, n), dtype=torch.float64, device="cuda")
        ap = L.AicParams()
        ap.ff_ratio_doublet = self.aic["ff_ratio_doublet"]
        ap.rankine_rad = self.aic["rankine_rad"]
        ap.cutoff_rad = self.aic["cutoff_rad"]
        L.check(self.lib.dust_aic_surfpan_dev(n, self.nv.data_ptr(), self.ver.data_ptr(),
                                              C.byref(ap), self.A.data_ptr(), n,
                                              self.B.data_ptr(), n, None), "aic")
        self.LU = self.A.clone()
        self.ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = self.lib.dust_dgetrf_dev(n, n, self.LU.data_ptr(), n, self.ipiv.data_ptr(), None)
        if info != 0:
            raise RuntimeError(f"static AIC is singular (dgetrf info {info})")
        self.b = torch.zeros(n, dtype=torch.float64, device="cuda")
        self.sigma = torch.zeros(n, dtype=torch.float64, device="cuda")
        self.uvort = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        self.ub = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        self.uvort_rel = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        # wake
        self.sub = sub
        self._set_te(te)
        self.r_shed = float(r_shed)
        self.pool = WK.ParticlePool(nmax)
        if box is None:
            lo = np.asarray(tree["origin"], dtype=np.float64)
            box = (lo + 1e-9, lo + np.asarray(tree["nbox"]) * tree["box_length"] - 1e-9)
        self.wprm = WK.params(box[0], box[1], dt=self.dt, use_vs=True)
        self.fmm = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"],
                                       tree["nlevels"], tree["min_part"], degree, tree["delta"],
                                       capacity=nmax, use_vs=True, u_inf=tuple(self.u_inf),
                                       precision=form)
        d3 = dict(dtype=torch.float64, device="cuda")
        self.p = dict(cen=torch.zeros((nmax, 3), **d3), dir=torch.zeros((nmax, 3), **d3),
                      mag=torch.zeros(nmax, **d3), rv=torch.zeros(nmax, **d3),
                      alpha=torch.zeros((nmax, 3), **d3), vel=torch.zeros((nmax, 3), **d3),
                      stretch=torch.zeros((nmax, 3), **d3))
        self.mu = self.b  # solve_linsys writes res in place
        self.times = {}

    def _centres(self, v4):
        """Vertex means (calc_geo_data): the first 3 or 4 vertices, summed in
        vertex order."""
        s = v4[:, 0] + v4[:, 1] + v4[:, 2]
        quad = self.nv_h == 4
        s[quad] = s[quad] + v4[quad, 3]
        return s / self.nv_h[:, None].astype(np.float64)

    def _set_te(self, te):
        up, lo, te0, te1 = te
        sub = self.sub
        q = (np.arange(sub) + 0.5) / sub
        self.te_up, self.te_lo = np.repeat(up, sub), np.repeat(lo, sub)
        self.te_vec = np.repeat((te1 - te0) / sub, sub, axis=0)
        self.te_mid = (np.repeat(te0, sub, axis=0) + np.tile(q, len(up))[:, None] *
                       np.repeat(te1 - te0, sub, axis=0) + 0.5 * self.dt * self.u_inf)

    def _move(self):
        """Rotating body: the geometry of the new step (rotation about z by
        omega dt; the caller's geometry update, outside the library) and the
        body velocity omega x r at the panel centres."""
        torch = self.torch
        self.theta += self.omega * self.dt
        v = WL.rotate_z(self.ver0, self.theta)
        cen = self._centres(v)
        self.ver.copy_(torch.from_numpy(v))
        self.cen.copy_(torch.from_numpy(cen))
        ub = np.stack([-self.omega * cen[:, 1], self.omega * cen[:, 0], np.zeros(self.n)], axis=1)
        self.ub.copy_(torch.from_numpy(ub))
        ns, nc = self.strips
        self._set_te(trailing_edge(v, ns, nc))

    def seed(self, cen, alpha, rv):
        """Initial particles (inserted like shed ones)."""
        self.pool.shed(cen, alpha, rv, np.tile(self.u_inf, (len(rv), 1)), self.wprm)

    def _shed(self):
        mu = self.b.cpu().numpy()
        pv = (mu[self.te_up] - mu[self.te_lo])[:, None] * self.te_vec
        k = len(pv)
        self.pool.shed(self.te_mid, pv, np.full(k, self.r_shed), np.tile(self.u_inf, (k, 1)),
                       self.wprm)

    def _fmm_pass(self, npart, err):
        """update_wake's FMM pass over the gathered particles into p[vel],
        p[stretch]; returns when its stream is done.  Exceptions go to err."""
        try:
            lib, p, h = self.lib, self.p, self.fmm.handle
            L.check(lib.dust_fmm_set_particles_dev(h, npart, p["cen"].data_ptr(),
                                                   p["dir"].data_ptr(), p["mag"].data_ptr(),
                                                   p["rv"].data_ptr()), "fmm set")
            nl = C.c_int()
            L.check(lib.dust_fmm_pass_dev(h, C.byref(nl)), "fmm pass")
            L.check(lib.dust_fmm_results_dev(h, p["vel"].data_ptr(), p["stretch"].data_ptr(),
                                             None), "fmm results")
            L.check(lib.dust_fmm_synchronize(h), "fmm sync")
        except BaseException as e:  # re-raised by the stepping thread
            err.append(e)

    def step(self):
        torch, lib, p = self.torch, self.lib, self.p
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(7)]
        ev[0].record()
        # prepare_wake: live list and the pointer arrays
        live = self.pool.live()
        npart = live.numel()
        P = self.pool
        L.check(lib.dust_wake_gather_dev(npart, live.data_ptr(), P.cen.data_ptr(),
                                         P.dir.data_ptr(), P.mag.data_ptr(),
                                         P.r_vortex.data_ptr(), p["cen"].data_ptr(),
                                         p["dir"].data_ptr(), p["mag"].data_ptr(),
                                         p["rv"].data_ptr(), p["alpha"].data_ptr(), None),
                "gather")
        h = self.fmm.handle
        if self.fmm_pan and npart:
            # use_fmm_pan: the tree and the expansions are built in prepare_wake
            # (mod_wake.f90:797-801) and serve both uvort and the particles
            torch.cuda.synchronize()
            L.check(lib.dust_fmm_set_particles_dev(h, npart, p["cen"].data_ptr(),
                                                   p["dir"].data_ptr(), p["mag"].data_ptr(),
                                                   p["rv"].data_ptr()), "fmm set")
            nl = C.c_int()
            L.check(lib.dust_fmm_sort_particles(h, C.byref(nl)), "fmm sort")
            L.check(lib.dust_fmm_calculate_multipole(h), "fmm multipole")
            L.check(lib.dust_fmm_synchronize(h), "fmm sync")
        ev[1].record()
        fmm_job, fmm_err = None, []
        if npart and self.overlap:
            torch.cuda.synchronize()  # the gathered particle arrays
            fmm_job = threading.Thread(target=self._fmm_pass, args=(npart, fmm_err))
            fmm_job.start()
        moving = self.omega != 0.0
        if moving:  # moving body: geometry of the step, AIC re-assembled
            self._move()
            ap = L.AicParams()
            ap.ff_ratio_doublet = self.aic["ff_ratio_doublet"]
            ap.rankine_rad = self.aic["rankine_rad"]
            ap.cutoff_rad = self.aic["cutoff_rad"]
            L.check(lib.dust_aic_surfpan_dev(self.n, self.nv.data_ptr(), self.ver.data_ptr(),
                                             C.byref(ap), self.A.data_ptr(), self.n,
                                             self.B.data_ptr(), self.n, None), "aic")
        # assemble_linsys: uvort, right-hand side and source strengths
        if self.fmm_pan:
            # apply_multipole_panels on the tree of this step (no end vortices
            # here, so get_vort_vel contributes nothing)
            self.uvort.zero_()
            if npart:
                torch.cuda.synchronize()
                L.check(lib.dust_fmm_apply_multipole_panels_dev(h, self.n, self.cen.data_ptr(),
                                                                self.uvort.data_ptr()),
                        "fmm panels")
                L.check(lib.dust_fmm_synchronize(h), "fmm sync")
        else:
            L.check(lib.dust_vort_vel_dev(self.n, self.cen.data_ptr(), npart,
                                          p["cen"].data_ptr(), p["dir"].data_ptr(),
                                          p["mag"].data_ptr(), p["rv"].data_ptr(), 0, None,
                                          None, C.byref(self.eprm), self.uvort.data_ptr(),
                                          None), "uvort")
        u = np.ascontiguousarray(self.u_inf)
        # sigma = nor.(ub - u_inf - uvort): the moving panels' velocity enters
        # as uvort - ub
        urel = self.uvort
        if moving:
            torch.sub(self.uvort, self.ub, out=self.uvort_rel)
            urel = self.uvort_rel
        L.check(lib.dust_aic_rhs_dev(self.n, self.nv.data_ptr(), self.ver.data_ptr(),
                                     self.B.data_ptr(), self.n, 1, u.ctypes.data,
                                     urel.data_ptr(), self.b.data_ptr(), self.n,
                                     self.sigma.data_ptr(), None), "rhs")
        ev[2].record()
        # solve_linsys: the static factors, or (moving) a new factorisation
        if not moving:
            L.check(lib.dust_dgetrs_dev(self.n, 1, self.LU.data_ptr(), self.n,
                                        self.ipiv.data_ptr(), self.b.data_ptr(), self.n, None),
                    "getrs")
        elif self.solver == "fp64":
            self.LU.copy_(self.A)
            info = lib.dust_dgetrf_dev(self.n, self.n, self.LU.data_ptr(), self.n,
                                       self.ipiv.data_ptr(), None)
            if info != 0:
                raise RuntimeError(f"AIC is singular (dgetrf info {info})")
            L.check(lib.dust_dgetrs_dev(self.n, 1, self.LU.data_ptr(), self.n,
                                        self.ipiv.data_ptr(), self.b.data_ptr(), self.n, None),
                    "getrs")
        else:
            it, res = C.c_int(), C.c_double()
            prec = {"fp32": L.LU_FP32, "fp16s": L.LU_FP16S}[self.solver]
            L.check(lib.dust_dgesv_mixed_dev(self.n, 1, self.A.data_ptr(), self.n,
                                             self.b.data_ptr(), self.n, prec, 1e-12, 30,
                                             C.byref(it), C.byref(res), None), "mixed solve")
        ev[3].record()
        # update_wake: FMM pass over the live particles
        if fmm_job is not None:
            fmm_job.join()
            if fmm_err:
                raise fmm_err[0]
        torch.cuda.synchronize()
        if npart and fmm_job is None:
            if self.fmm_pan:
                L.check(lib.dust_fmm_apply_multipole_dev(h), "fmm apply")
            else:
                L.check(lib.dust_fmm_set_particles_dev(h, npart, p["cen"].data_ptr(),
                                                       p["dir"].data_ptr(), p["mag"].data_ptr(),
                                                       p["rv"].data_ptr()), "fmm set")
                nl = C.c_int()
                L.check(lib.dust_fmm_pass_dev(h, C.byref(nl)), "fmm pass")
            L.check(lib.dust_fmm_results_dev(h, p["vel"].data_ptr(), p["stretch"].data_ptr(),
                                             None), "fmm results")
            L.check(lib.dust_fmm_synchronize(h), "fmm sync")
        ev[4].record()
        # ... plus the body panels (elem list of apply_multipole)
        grp = (L.ElemGroup * 1)()
        grp[0].kind, grp[0].n = L.ELEM_SURFPAN, self.n
        grp[0].n_ver, grp[0].ver = self.nv.data_ptr(), self.ver.data_ptr()
        grp[0].mag, grp[0].sigma, grp[0].stretch = self.b.data_ptr(), self.sigma.data_ptr(), 1
        L.check(lib.dust_elem_apply_dev(1, grp, C.byref(self.eprm), npart, p["cen"].data_ptr(),
                                        p["alpha"].data_ptr() if self.vs_elems else None,
                                        p["vel"].data_ptr(),
                                        p["stretch"].data_ptr() if self.vs_elems else None,
                                        None), "elements")
        L.check(lib.dust_wake_scatter_dev(npart, live.data_ptr(), p["vel"].data_ptr(),
                                          p["stretch"].data_ptr(), P.vel.data_ptr(),
                                          P.stretch.data_ptr(), None), "scatter")
        ev[5].record()
        # complete_wake: move, cull, shed
        P.update(self.wprm)
        self._shed()
        ev[6].record()
        torch.cuda.synchronize()
        names = ("prepare", "assemble", "solve", "fmm", "elements", "complete")
        # (moving bodies: "assemble" includes the AIC re-assembly; with
        # overlap "fmm" is the part of the FMM pass not hidden behind the
        # assembly and the solve, which it slows down while they share the GPU)
        for i, k in enumerate(names):
            self.times[k] = self.times.get(k, 0.0) + ev[i].elapsed_time(ev[i + 1])
        return npart

    def state(self):
        """Host copies: panel intensities, uvort, live slots and the pool."""
        live = self.pool.live().cpu().numpy()
        P = self.pool
        return dict(mu=self.b.cpu().numpy(), uvort=self.uvort.cpu().numpy(), live=live,
                    cen=P.cen.cpu().numpy(), dir=P.dir.cpu().numpy(), mag=P.mag.cpu().numpy(),
                    vel=P.vel.cpu().numpy(), stretch=P.stretch.cpu().numpy(),
                    free=P.free.cpu().numpy())
