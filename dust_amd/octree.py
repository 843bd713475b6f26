"""Host mirror of DUST's src/fmm public interface on top of libdust_gpu.

Names, argument meaning and error behaviour follow mod_octree
(src/fmm/mod_octree.f90:78-79):

    initialize_octree(box_length, nbox, origin, nlevels, min_part, degree,
                      delta)                          -> Octree   (251-518)
    sort_particles(parts, octree)                     -> n_leaves (703-996)
    calculate_multipole(parts, octree)                            (1002-1133)
    apply_multipole(parts, octree)                                (1139-1429)
    destroy_octree(octree)                                        (694)

`parts` is a VortParts, the SoA image of t_vortpart (src/aeroel/mod_vortpart.f90:85-97):
cen (n,3), dir (n,3), mag (n), r_vortex (n); apply_multipole fills vel, stretch
and turbvisc in place, as the reference writes through t_vortpart pointers.
Where the reference stops with error() (particle outside the box,
mod_octree.f90:757-762) a DustError is raised.

The sim_param switches the reference FMM reads (use_vs, use_vd, use_tv,
nu_inf, u_inf) are keyword arguments of initialize_octree.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class VortParts:
    """Vortex particles (t_vortpart fields used by the FMM)."""

    def __init__(self, cen, dir, mag, r_vortex):
        # own copies: sort_particles writes the redistributed strengths back
        self.cen = np.array(cen, dtype=np.float64, order="C").reshape(-1, 3)
        n = self.cen.shape[0]
        self.dir = np.array(dir, dtype=np.float64, order="C").reshape(n, 3)
        self.mag = np.array(np.broadcast_to(mag, (n,)), dtype=np.float64, order="C")
        self.r_vortex = np.array(np.broadcast_to(r_vortex, (n,)), dtype=np.float64, order="C")
        self.vel = np.zeros((n, 3))
        self.stretch = np.zeros((n, 3))
        self.turbvisc = np.zeros(n)
        # particles freed by a redistribution in sort_particles, and the
        # count of the others (the reference's n_prt)
        self.free = np.zeros(n, bool)
        self.n_prt = n
        self.rotu = np.zeros((n, 3))

    @classmethod
    def from_alpha(cls, cen, alpha, r_vortex):
        """Build from vorticity vectors alpha = mag * dir."""
        alpha = np.asarray(alpha, dtype=np.float64).reshape(-1, 3)
        mag = np.linalg.norm(alpha, axis=1)
        dir = np.where(mag[:, None] > 0, alpha / np.where(mag > 0, mag, 1.0)[:, None], 0.0)
        return cls(cen, dir, mag, r_vortex)

    @property
    def n(self):
        return self.cen.shape[0]

    @property
    def alpha(self):
        return self.dir * self.mag[:, None]


class Octree:
    """Handle on a device octree (the t_octree of mod_octree.f90:174-223)."""

    def __init__(self, params, capacity):
        self.params = params
        self.capacity = capacity
        h = C.c_void_p()
        L.check(L.lib().dust_fmm_create(C.byref(params), capacity, C.byref(h)), "initialize_octree")
        self._h = h
        self.n_leaves = 0
        # particle redistribution switches (sim_param use_pr, lvl_solid,
        # part_redist_ratio; mod_sim_param.f90:792-800)
        self.use_pr = False
        self.lvl_solid = max(params.nlevels - 2, 1)
        self.part_redist_ratio = 3.0

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("octree destroyed")
        return self._h

    @property
    def degree(self):
        return self.params.degree

    @property
    def nlevels(self):
        return self.params.nlevels

    def destroy(self):
        if self._h is not None:
            L.lib().dust_fmm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -- introspection -----------------------------------------------------
    def leaves(self, n):
        """Leaf membership: (level (1-based), ijk (1-based), npart, members)."""
        nl = self.n_leaves
        lev = np.zeros(nl, np.int32)
        ijk = np.zeros((nl, 3), np.int32)
        npart = np.zeros(nl, np.int32)
        mem = np.zeros(max(n, 1), np.int32)
        L.check(L.lib().dust_fmm_leaves(self.handle, L.ptr(lev), L.ptr(ijk), L.ptr(npart),
                                        L.ptr(mem)), "leaves")
        return lev, ijk, npart, mem[:n]

    def cell_index(self, n):
        ijk = np.zeros((max(n, 1), 3), np.int32)
        L.check(L.lib().dust_fmm_cell_index(self.handle, L.ptr(ijk)), "cell_index")
        return ijk[:n]

    def work(self):
        p = C.c_double()
        m = C.c_double()
        L.check(L.lib().dust_fmm_work(self.handle, C.byref(p), C.byref(m)), "work")
        return p.value, m.value

    def set_tuning(self, name, value):
        """Knob: near_factor, m2l_chunk (tuned form), seg_iters (both forms)."""
        L.check(L.lib().dust_fmm_set_tuning(self.handle, name.encode(), float(value)), "set_tuning")

    def get_tuning(self, name):
        v = C.c_double()
        L.check(L.lib().dust_fmm_get_tuning(self.handle, name.encode(), C.byref(v)), "get_tuning")
        return v.value

    def autotune(self, reps=3):
        """Time full passes over the particles already set for a grid of the
        performance knobs and keep the fastest; returns its ms per pass.  Run
        a pass afterwards before reading results."""
        ms = C.c_double()
        L.check(L.lib().dust_fmm_autotune(self.handle, int(reps), C.byref(ms)), "autotune")
        return ms.value

    def set_timing(self, on):
        L.check(L.lib().dust_fmm_set_timing(self.handle, int(bool(on))), "set_timing")

    def stage_times(self):
        ms = np.zeros(len(L.STAGES))
        cnt = np.zeros(len(L.STAGES), np.int32)
        L.check(L.lib().dust_fmm_stage_times(self.handle, L.ptr(ms), L.ptr(cnt)), "stage_times")
        return {s: (ms[i], int(cnt[i])) for i, s in enumerate(L.STAGES)}

    @property
    def stream(self):
        return L.lib().dust_fmm_stream(self.handle)

    def set_wind(self, wind, time):
        """The free-stream term of the next passes: variable_wind(pos, time)
        at every particle (mod_octree.f90:1376); wind from wind_params, None
        for the uniform u_inf of the parameters."""
        L.check(L.lib().dust_fmm_set_wind(self.handle, None if wind is None else C.byref(wind),
                                          float(time)), "set_wind")


def make_params(box_length, nbox, origin, nlevels, min_part, degree, delta, use_vs=True,
                use_vd=False, use_tv=False, nu_inf=0.0, u_inf=(0.0, 0.0, 0.0),
                precision="fp64", kernel="rosenhead", use_divfilt=False):
    """kernel: the particle core of kernel_coeffs, "rosenhead" (the
    reference's) or "gaussian" (its commented alternative,
    src/aeroel/mod_vortpart.f90:283-291).  use_divfilt (with use_vs):
    apply_multipole also fills parts.rotu, the divergence filter's curl of
    the velocity (mod_octree.f90:1207-1212, 1274-1336)."""
    p = L.FmmParams()
    p.box_length = float(box_length)
    for d in range(3):
        p.nbox[d] = int(nbox[d])
        p.origin[d] = float(origin[d])
        p.u_inf[d] = float(u_inf[d])
    p.nlevels = int(nlevels)
    p.min_part = int(min_part)
    p.degree = int(degree)
    p.delta = float(delta)
    p.use_vs = int(bool(use_vs))
    p.use_vd = int(bool(use_vd))
    p.use_tv = int(bool(use_tv))
    p.nu_inf = float(nu_inf)
    p.precision = {"fp64": L.FMM_FP64, "tuned": L.FMM_TUNED}[precision]
    p.kernel = KERNELS[kernel]
    p.use_divfilt = int(bool(use_divfilt))
    return p


KERNELS = {"rosenhead": L.KERNEL_ROSENHEAD, "gaussian": L.KERNEL_GAUSSIAN}
GUSTS = {None: L.GUST_NONE, "ACM": L.GUST_ACM, "linear": L.GUST_LINEAR}


def wind_params(u_inf, gust_type=None, origin=(0.0, 0.0, 0.0), front_direction=None,
                front_speed=None, u_des=0.0, perturb_direction=(0.0, 0.0, 1.0), gradient=1.0,
                start_time=0.0):
    """variable_wind's parameters (the sim_param gust options,
    mod_sim_param.f90:476-485, 814-835): gust_type "ACM", "linear" or None
    (any other string of the reference adds nothing).  front_direction and
    front_speed default to u_inf and |u_inf|, as the reference fills them."""
    w = L.WindParams()
    u = np.asarray(u_inf, dtype=np.float64)
    fd = u if front_direction is None else np.asarray(front_direction, dtype=np.float64)
    for d in range(3):
        w.u_inf[d] = float(u[d])
        w.origin[d] = float(origin[d])
        w.front_direction[d] = float(fd[d])
        w.perturb_direction[d] = float(perturb_direction[d])
    w.gust_type = GUSTS.get(gust_type, L.GUST_NONE)
    w.front_speed = float(np.linalg.norm(u) if front_speed is None else front_speed)
    w.u_des = float(u_des)
    w.gradient = float(gradient)
    w.start_time = float(start_time)
    return w


def variable_wind(wind, pos, time):
    """variable_wind (src/aeroel/mod_wind.f90:66-121) at the points pos (n,3)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(pos)
    L.check(L.lib().dust_variable_wind(len(pos), L.ptr(pos), C.byref(wind), float(time),
                                       L.ptr(out)), "variable_wind")
    return out


def initialize_octree(box_length, nbox, origin, nlevels, min_part, degree, delta, *,
                      capacity=1 << 20, particles_redistribution=False, octree_level_solid=None,
                      particles_redistribution_ratio=3.0, **flags):
    """initialize_octree (mod_octree.f90:251-518).  The redistribution
    options are the sim_param ones of mod_sim_param.f90:462-467,792-800
    (octree_level_solid defaults to max(nlevels - 2, 1))."""
    o = Octree(make_params(box_length, nbox, origin, nlevels, min_part, degree, delta, **flags),
               capacity)
    o.use_pr = bool(particles_redistribution)
    if octree_level_solid is not None:
        o.lvl_solid = int(octree_level_solid)
    o.part_redist_ratio = float(particles_redistribution_ratio)
    return o


def sort_particles(parts, octree, elem=None):
    """sort_particles (mod_octree.f90:703-996); returns the number of leaves.

    elem: the element centres (m,3) the reference sorts into the tree with the
    particles (777-820); with particles_redistribution on they mark the cells
    near a wall, and the redistribution (827-894) updates parts.mag, parts.dir
    and parts.free (the freed particles) in place and sets parts.n_prt."""
    lib = L.lib()
    L.check(lib.dust_fmm_set_particles(octree.handle, parts.n, L.ptr(parts.cen), L.ptr(parts.dir),
                                       L.ptr(parts.mag), L.ptr(parts.r_vortex)), "sort_particles")
    if octree.use_pr:
        cen = np.ascontiguousarray(np.zeros((0, 3)) if elem is None else elem,
                                   dtype=np.float64).reshape(-1, 3)
        L.check(lib.dust_fmm_set_redistribution(octree.handle, 1, octree.lvl_solid,
                                                octree.part_redist_ratio, len(cen), L.ptr(cen)),
                "sort_particles")
    else:
        L.check(lib.dust_fmm_set_redistribution(octree.handle, 0, 1, 1.0, 0, None),
                "sort_particles")
    nl = C.c_int()
    L.check(lib.dust_fmm_sort_particles(octree.handle, C.byref(nl)), "sort_particles")
    octree.n_leaves = nl.value
    if octree.use_pr and parts.n > 0:
        keep = np.zeros(parts.n, np.uint8)
        nk = C.c_int()
        L.check(lib.dust_fmm_redistributed(octree.handle, C.byref(nk), L.ptr(keep),
                                           L.ptr(parts.dir), L.ptr(parts.mag)), "sort_particles")
        parts.free = keep == 0
    parts.n_prt = int(parts.n - np.count_nonzero(parts.free))
    return nl.value


def calculate_multipole(parts, octree):
    """calculate_multipole (mod_octree.f90:1002-1133); `parts` is unused as in
    the reference."""
    L.check(L.lib().dust_fmm_calculate_multipole(octree.handle), "calculate_multipole")


def apply_multipole(parts, octree):
    """apply_multipole (mod_octree.f90:1139-1429), particle terms + free stream."""
    L.check(L.lib().dust_fmm_apply_multipole(octree.handle, L.ptr(parts.vel), L.ptr(parts.stretch),
                                             L.ptr(parts.turbvisc)), "apply_multipole")
    if octree.params.use_divfilt and octree.params.use_vs:
        L.check(L.lib().dust_fmm_rotu(octree.handle, L.ptr(parts.rotu)), "apply_multipole")


def apply_multipole_panels(octree, cen, uvort=None):
    """apply_multipole_panels (mod_octree.f90:1434-1510): uvort (m,3) +=
    the particle velocity at the panel centres cen (m,3) from the last tree."""
    cen = np.ascontiguousarray(cen, dtype=np.float64).reshape(-1, 3)
    u = (np.zeros((len(cen), 3)) if uvort is None
         else np.array(uvort, dtype=np.float64).reshape(-1, 3))
    L.check(L.lib().dust_fmm_apply_multipole_panels(octree.handle, len(cen), L.ptr(cen),
                                                    L.ptr(u)), "apply_multipole_panels")
    return u


def destroy_octree(octree):
    """destroy_octree (mod_octree.f90:694)."""
    octree.destroy()


def direct(parts, use_vs=True, use_vd=False, nu_inf=0.0, u_inf=(0.0, 0.0, 0.0),
           kernel="rosenhead"):
    """Direct O(N^2) particle terms, the fmm=F branch of update_wake
    (src/aeroel/mod_wake.f90:1050-1113).  Returns (vel, stretch)."""
    vel = np.zeros((parts.n, 3))
    st = np.zeros((parts.n, 3))
    u = np.asarray(u_inf, dtype=np.float64)
    L.check(L.lib().dust_vortpart_direct_k(parts.n, L.ptr(parts.cen), L.ptr(parts.dir),
                                           L.ptr(parts.mag), L.ptr(parts.r_vortex), int(use_vs),
                                           int(use_vd), float(nu_inf), L.ptr(u),
                                           KERNELS[kernel], L.ptr(vel), L.ptr(st)), "direct")
    return vel, st
