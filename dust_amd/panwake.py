"""Host mirror of the panel wake of src/aeroel/mod_wake.f90 on libdust_gpu:
initialize_wake / complete_wake trailing-edge rows (start), update_wake's
panel part (update) and complete_wake's shedding from the last panel row
(shed).  Inputs and outputs are device buffers (torch tensors or raw device
addresses); tables are 1-based as in the reference (te%ii, te%neigh, te%o,
pan_gen_elems_id).
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _addr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


class PanWake:
    def __init__(self, start, neigh, neigh_o, gen, n_points, nmax, dt, ndt=1, min_vel=1e-3,
                 join_te=False, join_fact=0.1, refine=False, k_refine=1, kvr=0.0, vr=0.1,
                 box=((-1e30,) * 3, (1e30,) * 3)):
        self.tab = [np.ascontiguousarray(a, dtype=np.int32) for a in (start, neigh, neigh_o, gen)]
        p = L.PanwakeParams()
        p.n_stripes, p.n_points, p.nmax_pan = len(self.tab[0]), int(n_points), int(nmax)
        p.start_points, p.neigh, p.neigh_o, p.gen_elems = [t.ctypes.data for t in self.tab]
        p.dt, p.ndt_update_wake, p.min_vel_at_te = float(dt), int(ndt), float(min_vel)
        p.join_te, p.join_te_factor = int(join_te), float(join_fact)
        p.refine_wake, p.k_refine = int(refine), int(k_refine)
        p.k_vortex_rad, p.vortex_rad = float(kvr), float(vr)
        p.box_min = (C.c_double * 3)(*box[0])
        p.box_max = (C.c_double * 3)(*box[1])
        self.prm = p
        self.ns, self.np, self.nmax = p.n_stripes, p.n_points, p.nmax_pan
        h = C.c_void_p()
        L.check(L.lib().dust_panwake_create(C.byref(p), C.byref(h)), "panwake_create")
        self.h = h

    def destroy(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().dust_panwake_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def start(self, first, spts, gdir, scal, vte, uinf, stream=None):
        u = (C.c_double * 3)(*uinf)
        L.check(L.lib().dust_panwake_start_dev(self.h, int(first), _addr(spts), _addr(gdir),
                                               _addr(scal), _addr(vte), u, stream),
                "panwake_start")

    def panels(self):
        """dict n_wake, n_impl and device addresses n_ver, ver, mag, impl_ind."""
        nw, ni = C.c_int(), C.c_int()
        pv = [C.c_void_p() for _ in range(4)]
        L.check(L.lib().dust_panwake_panels(self.h, C.byref(nw), C.byref(ni),
                                            *[C.byref(x) for x in pv]), "panwake_panels")
        return dict(n_wake=nw.value, n_impl=ni.value, n_ver=pv[0].value, ver=pv[1].value,
                    mag=pv[2].value, impl_ind=pv[3].value)

    def points(self):
        n, p = C.c_int(), C.c_void_p()
        L.check(L.lib().dust_panwake_points(self.h, C.byref(n), C.byref(p)), "panwake_points")
        return n.value, p.value

    def first_row(self, mu, stream=None):
        L.check(L.lib().dust_panwake_first_row_dev(self.h, _addr(mu), stream), "first_row")

    def update(self, mu, vel, stream=None):
        L.check(L.lib().dust_panwake_update_dev(self.h, _addr(mu), _addr(vel), stream),
                "panwake_update")

    def shed(self, pool, prm, n_prt, status, stream=None):
        """pool: object with device tensors free, cen, dir, mag, r_vortex (or
        rv), vel."""
        rv = pool.r_vortex if hasattr(pool, "r_vortex") else pool.rv
        L.check(L.lib().dust_panwake_shed_dev(
            self.h, int(pool.free.numel()), _addr(pool.free), _addr(pool.cen), _addr(pool.dir),
            _addr(pool.mag), _addr(rv), _addr(pool.vel), C.byref(prm), _addr(n_prt),
            _addr(status), stream), "panwake_shed")

    def end_vortices(self):
        n = C.c_int()
        pv = [C.c_void_p() for _ in range(3)]
        L.check(L.lib().dust_panwake_end_vortices(self.h, C.byref(n), *[C.byref(x) for x in pv]),
                "panwake_end_vortices")
        return n.value, pv[0].value, pv[1].value, pv[2].value

    def state(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        L.check(L.lib().dust_panwake_state(self.h, C.byref(a), C.byref(b), C.byref(c)), "state")
        return a.value, bool(b.value), c.value

    def get(self):
        rows = self.nmax + 1
        out = dict(pts=np.zeros((rows, self.np, 3)), pvel=np.zeros((rows, self.np, 3)),
                   idou=np.zeros((self.nmax, self.ns)), end_idou=np.zeros(self.ns),
                   last_idou=np.zeros(self.ns), pend=np.zeros((self.np, 3)))
        L.check(L.lib().dust_panwake_get(self.h, *[out[k].ctypes.data for k in
                                                   ("pts", "pvel", "idou", "end_idou",
                                                    "last_idou", "pend")]), "panwake_get")
        return out

    def checkpoint(self):
        """The device state as bytes (dust_panwake_checkpoint)."""
        n = C.c_size_t()
        L.check(L.lib().dust_panwake_checkpoint(self.h, None, C.byref(n)), "checkpoint")
        buf = np.zeros(n.value, np.uint8)
        L.check(L.lib().dust_panwake_checkpoint(self.h, buf.ctypes.data, C.byref(n)),
                "checkpoint")
        return buf

    def load(self, length, points, pvel, idou, last_idou, stream=None):
        """The panel part of load_wake from a result file's arrays
        (dust_panwake_load): points / pvel [length+1][n_points][3], idou
        [length][n_stripes], last_idou [n_stripes]."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (points, pvel, idou, last_idou)]
        length = int(length)
        if a[0].shape != (length + 1, self.np, 3) or a[1].shape != a[0].shape:
            raise ValueError(f"panel wake points of shape {a[0].shape}, "
                             f"({length + 1}, {self.np}, 3) expected")
        if a[2].shape != (length, self.ns) or a[3].shape != (self.ns,):
            raise ValueError("panel wake intensities of the wrong shape")
        L.check(L.lib().dust_panwake_load(self.h, length, *[x.ctypes.data for x in a], stream),
                "panwake_load")

    def restore(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        L.check(L.lib().dust_panwake_restore(self.h, buf.ctypes.data, buf.size), "restore")
