"""Host mirror of DUST's pressure (Bernoulli) integral equation
(src/linsys/mod_pressure_equation.f90) on top of libdust_gpu.

    pres_rhs_rot(cen, b, ...)      the rotational part (c) of
                                   assemble_pressure_sys (:139-236): particles,
                                   end-vortex lines and wake panels acting on
                                   the surface-panel centres, added to b_pres
    solve_pressure_sys(A, P, ns, b) solve_pressure_sys (:251-390): the block
                                   static/moving elimination of the surface-
                                   panel system, as solve_linsys

    pres_rhs_src(B, src)           the source part (b): the source matrix of
                                   the surface panels times their
                                   bernoulli_source (mod_surfpan.f90:405-437)

    press_normvel_der(...)         press_normvel_der (:396-462): the
                                   panels' bernoulli_source from their
                                   velocities and neighbour stencils
"""
import numpy as np

from . import _lib as L


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def pres_rhs_rot(cen, b=None, pcen=None, pdir=None, pmag=None, prv=None, pvel=None, lver=None,
                 lmag=None, lvel=None, wver=None, wnver=None, wmag=None, wvel=None):
    """b_pres (m,) += the rotational terms at the centres cen (m,3).
    Particles: pcen, pdir, pvel (np,3), pmag, prv (np,) -- live ones only;
    end-vortex lines: lver, lvel (nl,2,3), lmag (nl,) or None (not
    associated); wake panels: wver, wvel (nw,4,3) or (nw,3,3), wnver, wmag
    (nw,), the vertex velocities in the vertex order.  Returns b."""
    cen = _f64(cen, (-1, 3))
    m = cen.shape[0]
    b = np.zeros(m) if b is None else np.array(b, dtype=np.float64).reshape(m)
    npart = 0 if pmag is None else len(pmag)
    nl = 0 if lver is None else np.asarray(lver).shape[0]
    nw = 0 if wver is None else np.asarray(wver).shape[0]
    pc, pd, pm, pr, pv = ((_f64(pcen, (-1, 3)), _f64(pdir, (-1, 3)), _f64(pmag), _f64(prv),
                           _f64(pvel, (-1, 3))) if npart else (None,) * 5)
    lv = _f64(lver, (-1, 2, 3)) if nl else None
    lm = _f64(lmag) if (nl and lmag is not None) else None
    lu = _f64(lvel, (-1, 2, 3)) if nl else None
    wv = wu = wn = wm = None
    if nw:
        wv = np.zeros((nw, 4, 3))
        wu = np.zeros((nw, 4, 3))
        wv[:, : np.asarray(wver).shape[1]] = np.asarray(wver)[:, :4]
        wu[:, : np.asarray(wvel).shape[1]] = np.asarray(wvel)[:, :4]
        wn = np.ascontiguousarray(wnver, dtype=np.int32)
        wm = _f64(wmag)
    L.check(L.lib().dust_pres_rhs_rot(m, L.ptr(cen), npart, L.ptr(pc), L.ptr(pd), L.ptr(pm),
                                      L.ptr(pr), L.ptr(pv), nl, L.ptr(lv), L.ptr(lm), L.ptr(lu), nw,
                                      L.ptr(wn), L.ptr(wv), L.ptr(wm), L.ptr(wu), L.ptr(b)),
            "pres_rhs_rot")
    return b


def pres_rhs_src(B, src):
    """The source part (b) of b_pres (mod_surfpan.f90:405-437): B_sp src, with
    B the surface panels' source matrix (Fortran order) and src their
    bernoulli_source.  Returns b_pres."""
    n = B.shape[0]
    if not (B.flags.f_contiguous and B.dtype == np.float64):
        raise ValueError("B must be a Fortran-ordered float64 array")
    src = _f64(src)
    b = np.zeros(n)
    L.check(L.lib().dust_pres_rhs_src(n, L.ptr(B), n, L.ptr(src), L.ptr(b)), "pres_rhs_src")
    return b


def solve_pressure_sys(A, P, nstatic_sp, b):
    """solve_pressure_sys (mod_pressure_equation.f90:251-390).  A (n_sp, n_sp,
    Fortran order) holds the static block factored by dgetrf and P its pivots;
    A and P are updated in place as the reference leaves them.  Returns
    (res_pres, info)."""
    n = A.shape[0]
    if not (A.flags.f_contiguous and A.dtype == np.float64):
        raise ValueError("A must be a Fortran-ordered float64 array")
    if not (P.dtype == np.int32 and P.flags.c_contiguous and P.shape == (n,)):
        raise ValueError("P must be a contiguous int32 array of n_sp pivots")
    res = np.zeros(n)
    b = np.ascontiguousarray(b, dtype=np.float64)
    info = L.check(L.lib().dust_solve_pressure_sys(n, nstatic_sp, L.ptr(A), n, L.ptr(P), L.ptr(b),
                                                   L.ptr(res)), "solve_pressure_sys")
    return res, info


def press_normvel_der(nor, ub, surf_vel, surf_vel_old, stencil, n_ver, neigh, R, dt):
    """press_normvel_der (mod_pressure_equation.f90:396-462) for the n surface
    panels in idSurfPan order: nor, ub, surf_vel, surf_vel_old (n,3);
    stencil (n,4,3) = pot_vel_stencil; n_ver (n); neigh (n,4) the 0-based
    surface-panel index across each edge, -1 for no neighbour (the edge takes
    -2 surf_vel), -2 for a neighbour that is not a surface panel (no term);
    R (n,3,3) the rotation R_g applied to panel i (R[i][row, column]; the
    reference takes that of the component of elems(i)).  Returns the
    bernoulli_source (n)."""
    n = len(n_ver)
    out = np.zeros(n)
    Rc = np.ascontiguousarray(np.transpose(np.asarray(R, np.float64), (0, 2, 1)))
    L.check(L.lib().dust_press_normvel_der(
        n, L.ptr(_f64(nor)), L.ptr(_f64(ub)), L.ptr(_f64(surf_vel)), L.ptr(_f64(surf_vel_old)),
        L.ptr(_f64(stencil)), L.ptr(np.ascontiguousarray(n_ver, np.int32)),
        L.ptr(np.ascontiguousarray(neigh, np.int32)), L.ptr(Rc), float(dt), L.ptr(out)),
        "press_normvel_der")
    return out
