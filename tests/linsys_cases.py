"""Systems for the assembly tests (initialize_linsys / assemble_linsys with a
panel wake): a vortex-lattice flat plate and a surface-panel NACA0012 wing,
each with a first row of implicit wake panels behind the trailing edge."""
import numpy as np

from dust_amd import linsys as LS
from dust_amd import workloads as W


def helmbold_cla(ar):
    """Lift slope of a straight wing of aspect ratio ar (Helmbold's formula)."""
    return 2.0 * np.pi * ar / (2.0 + np.sqrt(ar * ar + 4.0))


def vlm_plate(ar=8.0, nc=6, ns=24, wake_len=200.0):
    """Flat rectangular plate (chord 1, span ar) in z = 0 as vortex rings set
    back a quarter panel (collocation at the panel three-quarter chord), and
    one long implicit wake panel per spanwise strip behind the last ring.
    Returns (ver, n_ver, strip width, wake dict)."""
    xs = np.linspace(0.0, 1.0, nc + 1)
    dx = xs[1] - xs[0]
    ys = np.linspace(-ar / 2, ar / 2, ns + 1)
    ver = []
    for j in range(ns):
        for i in range(nc):
            x0, x1 = xs[i] + dx / 4, xs[i + 1] + dx / 4
            ver.append([[x0, ys[j], 0.0], [x1, ys[j], 0.0], [x1, ys[j + 1], 0.0],
                        [x0, ys[j + 1], 0.0]])
    ver = np.array(ver)
    n_ver = np.full(len(ver), 4, np.int32)
    xt = 1.0 + dx / 4
    wv, ind = [], []
    for j in range(ns):
        wv.append([[xt, ys[j + 1], 0.0], [xt, ys[j], 0.0], [xt + wake_len, ys[j], 0.0],
                   [xt + wake_len, ys[j + 1], 0.0]])
        ind.append([j * nc + nc, 0])  # last ring of the strip, 1-based; no lower element
    wake = dict(ver=np.array(wv), n_ver=np.full(ns, 4, np.int32), mag=np.zeros(ns), n_impl=ns,
                impl_ind=np.array(ind, np.int32))
    return ver, n_ver, ys[1] - ys[0], wake


def morino_wing(nc=12, ns=16, ar=8.0, alpha=5.0, wake_len=200.0):
    """The NACA0012 wing of workloads.wing_case as surface panels with one long
    implicit wake panel per strip (generated by the upper and lower trailing-
    edge panels).  Returns (ver, n_ver, strip width, wake dict, upper, lower)."""
    ver, nv, *_ = W.wing_case(nc, ns, ar, alpha, nrows=1)
    up = np.arange(ns) * 2 * nc
    lo = up + 2 * nc - 1
    p1, p2 = ver[up, 0], ver[up, 3]
    d = np.array([wake_len, 0.0, 0.0])
    wv = np.stack([p1, p2, p2 + d, p1 + d], axis=1)
    wake = dict(ver=wv, n_ver=np.full(ns, 4, np.int32), mag=np.zeros(ns), n_impl=ns,
                impl_ind=np.stack([up + 1, lo + 1], axis=1).astype(np.int32))
    return ver, nv, ar / ns, wake, up, lo


def system(ver, n_ver, kind, nstatic, wake=None, **kw):
    return LS.System(np.full(len(n_ver), kind, np.int32) if np.isscalar(kind) else kind, ver,
                     n_ver, nstatic, wake=wake, **kw)


def mixed_case(nc=6, ns=6):
    """Surface-panel wing (explicit wake panels behind the implicit row, two
    lifting-line panels ahead) with a vortex-lattice plate behind it; the
    surface panels are the pressure system."""
    ver, nv, dy, wake, up, lo = morino_wing(nc=nc, ns=ns, wake_len=0.5)
    pv, pnv, _, pw = vlm_plate(ar=4.0, nc=2, ns=4, wake_len=0.5)
    pv = pv + np.array([3.0, 0.0, 0.2])
    pw["ver"] = pw["ver"] + np.array([3.0, 0.0, 0.2])
    n_w = len(nv)
    rows2 = wake["ver"] + np.array([0.5, 0.0, 0.0])
    wv = np.concatenate([wake["ver"], pw["ver"], rows2])
    impl = np.concatenate([wake["impl_ind"], pw["impl_ind"] + np.array([n_w, 0])])
    wk = dict(ver=wv, n_ver=np.full(len(wv), 4, np.int32),
              mag=np.concatenate([np.zeros(len(impl)), np.linspace(-0.3, 0.4, len(rows2))]),
              n_impl=len(impl), impl_ind=impl.astype(np.int32))
    ev = np.array([[[-1.0, -1.0, 0.3], [-0.8, -1.0, 0.3], [-0.8, 0.0, 0.3], [-1.0, 0.0, 0.3]],
                   [[-1.0, 0.0, 0.3], [-0.8, 0.0, 0.3], [-0.8, 1.0, 0.3], [-1.0, 1.0, 0.3]]])
    expl = dict(ver=ev, n_ver=np.full(2, 4, np.int32), mag=np.array([0.7, -0.2]))
    ver_all = np.concatenate([ver, pv])
    nv_all = np.concatenate([nv, pnv])
    kind = np.concatenate([np.full(n_w, LS.ROW_SURFPAN), np.full(len(pnv), LS.ROW_VORTLATT)])
    rng = np.random.default_rng(5)
    n = len(nv_all)
    return dict(ver=ver_all, n_ver=nv_all, kind=kind.astype(np.int32), wake=wk, expl=expl,
                uvort=rng.normal(scale=0.3, size=(n, 3)), bern=rng.normal(size=n),
                id_sp=np.arange(n_w, dtype=np.int32), strip=2 * nc, n_wing=n_w)


def mixed_system(c, nstatic, nstatic_sp, nstatic_expl, ub=None):
    return LS.System(c["kind"], c["ver"], c["n_ver"], nstatic, ub=ub, uvort=c["uvort"],
                     bern=c["bern"], u_inf=(10.0, 0.0, 1.0), id_sp=c["id_sp"],
                     nstatic_sp=nstatic_sp, wake=c["wake"],
                     expl=dict(c["expl"], nstatic=nstatic_expl))
