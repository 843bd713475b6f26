"""Host mirror of the device-resident particle pool of DUST's wake
(src/aeroel/mod_wake.f90) on top of libdust_gpu.

    ParticlePool(nmax)        wake_parts(nmax) as device arrays (torch tensors
                              on the GPU): cen, dir, mag, r_vortex, vel,
                              stretch, rotu, free (uint8, 1 = free slot), and
                              n_prt (a device int)
    .update(...)              the particle loop of complete_wake (:1305-1355)
    .live()                   the live slots of prepare_wake (:826-835)
    .shed(...)                insertion of shed particles (:1888-1967)

Slots are 0-based here (the reference's ip - 1).
"""
import ctypes as C

import numpy as np

from . import _lib as L


def params(box_min, box_max, dt, filt_eta=0.0, use_vs=True, use_vd=False, use_divfilt=False):
    """dt = sim_param%dt * ndt_update_wake; filt_eta = sim_param%filt_eta /
    ndt_update_wake."""
    p = L.WakeParams()
    for e in range(3):
        p.box_min[e] = float(box_min[e])
        p.box_max[e] = float(box_max[e])
    p.dt = float(dt)
    p.filt_eta = float(filt_eta)
    p.use_vs, p.use_vd, p.use_divfilt = int(use_vs), int(use_vd), int(use_divfilt)
    return p


class ParticlePool:
    def __init__(self, nmax):
        import torch
        self.torch = torch
        d = dict(dtype=torch.float64, device="cuda")
        self.nmax = int(nmax)
        self.cen = torch.zeros((nmax, 3), **d)
        self.dir = torch.zeros((nmax, 3), **d)
        self.mag = torch.zeros(nmax, **d)
        self.r_vortex = torch.zeros(nmax, **d)
        self.vel = torch.zeros((nmax, 3), **d)
        self.stretch = torch.zeros((nmax, 3), **d)
        self.rotu = torch.zeros((nmax, 3), **d)
        self.free = torch.ones(nmax, dtype=torch.uint8, device="cuda")
        self.n_prt = torch.zeros(1, dtype=torch.int32, device="cuda")
        self._live = torch.zeros(max(1, nmax), dtype=torch.int32, device="cuda")
        self._n_live = torch.zeros(1, dtype=torch.int32, device="cuda")

    def update(self, prm):
        L.check(L.lib().dust_wake_update_dev(
            self.nmax, self.free.data_ptr(), self.cen.data_ptr(), self.dir.data_ptr(),
            self.mag.data_ptr(), self.vel.data_ptr(), self.stretch.data_ptr(),
            self.rotu.data_ptr(), C.byref(prm), self.n_prt.data_ptr(), None), "wake_update")

    def live(self):
        """Device tensor of the live slots, ascending."""
        L.check(L.lib().dust_wake_live_dev(self.nmax, self.free.data_ptr(), self._live.data_ptr(),
                                           self._n_live.data_ptr(), None), "wake_live")
        n = int(self._n_live.item())
        return self._live[:n]

    def shed(self, new_cen, new_partvec, new_rv, new_vel, prm):
        t = self.torch
        k = len(new_rv)
        dv = [t.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
              for a in (new_cen, new_partvec, new_rv, new_vel)]
        L.check(L.lib().dust_wake_shed_dev(
            self.nmax, self.free.data_ptr(), self.cen.data_ptr(), self.dir.data_ptr(),
            self.mag.data_ptr(), self.r_vortex.data_ptr(), self.vel.data_ptr(), k,
            dv[0].data_ptr(), dv[1].data_ptr(), dv[2].data_ptr(), dv[3].data_ptr(), C.byref(prm),
            self.n_prt.data_ptr(), None), "wake_shed")

    def avoid_collision(self, groups, ub, prm):
        """avoid_collision (mod_wake.f90:2350-2534) on the live slots' vel.
        groups: the body's element lists (dust_amd.elements dicts; surface
        panels and vortex-lattice rings); ub: per group a device (n,3) tensor
        of the panels' velocity, or None."""
        from . import elements as E
        arr, keep = E.pack_groups(groups, device=True)
        ptrs = (C.c_void_p * max(1, len(groups)))(
            *[None if u is None else u.data_ptr() for u in ub])
        L.check(L.lib().dust_wake_avoid_collision_dev(
            self.nmax, self.free.data_ptr(), self.cen.data_ptr(), self.r_vortex.data_ptr(),
            self.vel.data_ptr(), len(groups), arr, ptrs, C.byref(prm), None), "avoid_collision")
        del keep


def pa_params(dt, dt_wake, u_ref, rad_mult=1.0, elrad_mult=1.0):
    """sim_param%dt, dt * ndt_update_wake, u_ref, pa_rad_mult, pa_elrad_mult."""
    p = L.PaParams()
    p.dt, p.dt_wake, p.u_ref = float(dt), float(dt_wake), float(u_ref)
    p.rad_mult, p.elrad_mult = float(rad_mult), float(elrad_mult)
    return p
