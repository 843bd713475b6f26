"""Host mirror of DUST's direct element <-> particle terms on top of libdust_gpu.

    apply_elements(groups, pos, ...)  the element loops of apply_multipole
                                      (src/fmm/mod_octree.f90:1350-1401): the
                                      velocity of body panels, wake panels, wake
                                      rings and end vortices at the particles,
                                      and with alpha the vs_elems stretching
    get_vort_vel(cen, ...)            get_vort_vel_surfpan / _vortlatt
                                      (src/aeroel/mod_surfpan.f90:962-977,
                                      mod_vortlatt.f90:676-693): the particle and
                                      end-vortex velocity at the panel centres

An element group is a dict: kind (SURFPAN, RING or LINE), ver ((n,4,3)
polygons, triangles padded, or (n,2,3) lines), n_ver (polygons), mag
(strength; None = not associated), sigma (surface-panel source strength
nor.(ub - wind - uvort); None = no source term) and stretch (bool: the list
takes part in the stretching term; the reference's wake rings do not).
Arrays are numpy float64; results come back on the host.  Parameters are
sim_param's FarFieldRatioDoublet, RankineRad, CutoffRad and VortexRad.
"""
import ctypes as C

import numpy as np

from . import _lib as L

SURFPAN, RING, LINE = L.ELEM_SURFPAN, L.ELEM_RING, L.ELEM_LINE


def params(ff_ratio_doublet=10.0, rankine_rad=0.1, cutoff_rad=0.001, vortex_rad=0.1,
           precision="fp64"):
    p = L.ElemParams()
    p.precision = {"fp64": L.ELEM_FP64, "tuned": L.ELEM_TUNED}[precision]
    p.ff_ratio_doublet = float(ff_ratio_doublet)
    p.rankine_rad = float(rankine_rad)
    p.cutoff_rad = float(cutoff_rad)
    p.vortex_rad = float(vortex_rad)
    return p


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def pack_groups(groups, device=False):
    """ctypes array of dust_elem_group plus the arrays it points into (kept
    alive by the caller): numpy host arrays, or with device=True torch tensors
    on the GPU."""
    keep = []
    arr = (L.ElemGroup * max(1, len(groups)))()
    for k, g in enumerate(groups):
        kind = int(g["kind"])
        ver = np.asarray(g["ver"], dtype=np.float64)
        n = ver.shape[0]
        if kind == LINE:
            v = _f64(ver[:, :2, :], (n, 2, 3))
            nv = None
        else:
            v = np.zeros((n, 4, 3))
            v[:, : ver.shape[1], :] = ver[:, :4, :]
            nv = np.ascontiguousarray(g["n_ver"], dtype=np.int32)
        mag = _f64(g.get("mag"))
        sig = _f64(g.get("sigma"))
        if device:
            import torch
            v, nv, mag, sig = [None if a is None else torch.from_numpy(a).cuda()
                               for a in (v, nv, mag, sig)]
        keep += [v, nv, mag, sig]
        e = arr[k]
        e.kind, e.n = kind, n
        e.n_ver, e.ver, e.mag, e.sigma = L.ptr(nv), L.ptr(v), L.ptr(mag), L.ptr(sig)
        e.stretch = 1 if g.get("stretch", True) else 0
    return arr, keep


def apply_elements(groups, pos, alpha=None, vel=None, stretch=None, **prm):
    """vel (+= element velocity) and, with alpha, stretch (+= element
    stretching) at the points pos (n,3); returns (vel, stretch)."""
    pos = _f64(pos, (-1, 3))
    n = pos.shape[0]
    vel = np.zeros((n, 3)) if vel is None else np.array(vel, dtype=np.float64).reshape(n, 3)
    al = _f64(alpha, (n, 3)) if alpha is not None else None
    if al is not None:
        stretch = (np.zeros((n, 3)) if stretch is None
                   else np.array(stretch, dtype=np.float64).reshape(n, 3))
    arr, keep = pack_groups(groups)
    L.check(L.lib().dust_elem_apply(len(groups), arr, C.byref(params(**prm)), n, L.ptr(pos),
                                    L.ptr(al), L.ptr(vel), L.ptr(stretch) if al is not None
                                    else None), "dust_elem_apply")
    del keep
    return vel, stretch


def get_vort_vel(cen, pcen=None, pdir=None, pmag=None, prv=None, lver=None, lmag=None, **prm):
    """uvort (m,3) at the centres cen (m,3) from vortex particles (cen, dir,
    mag, r_vortex) and end-vortex lines lver (nl,2,3) of strength lmag."""
    cen = _f64(cen, (-1, 3))
    m = cen.shape[0]
    np_ = 0 if pmag is None else len(pmag)
    nl = 0 if lver is None else np.asarray(lver).shape[0]
    pc, pd = _f64(pcen, (-1, 3)) if np_ else None, _f64(pdir, (-1, 3)) if np_ else None
    pm, pr = _f64(pmag) if np_ else None, _f64(prv) if np_ else None
    lv = _f64(lver, (-1, 2, 3)) if nl else None
    lm = _f64(lmag) if (nl and lmag is not None) else None
    u = np.zeros((m, 3))
    L.check(L.lib().dust_vort_vel(m, L.ptr(cen), np_, L.ptr(pc), L.ptr(pd), L.ptr(pm), L.ptr(pr),
                                  nl, L.ptr(lv), L.ptr(lm), C.byref(params(**prm)), L.ptr(u)),
            "dust_vort_vel")
    return u
