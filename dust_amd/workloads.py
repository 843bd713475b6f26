"""Seeded synthetic inputs for the hot path (numpy only, no device work).

helix_wake   -- vortex-particle wake: three helical tubes plus a uniform
                background (the 1,000,000-particle FMM workload, seed 20261203)
ellipsoid    -- perturbed ellipsoid surface mesh for the dense AIC
                (128 x 128 panels = 16,384, seed 20261204)
wing_case    -- rectangular NACA0012 wing (40 chordwise per side x 20 spanwise)
                with a trailing vortex-particle sheet (1,600 panels, 8,000
                particles)
rotor        -- b rectangular NACA0012 blades (aspect ratio 10) around the z
                axis, the surface-panel mesh of one rotor case
element_lists -- element lists of apply_multipole around a panel mesh: the
                body surface panels (doublet + source strengths), a first
                row of wake panels and rings behind given trailing-edge
                points, and the end vortices closing them
batch_cases  -- the many-case workload: 48 rotor cases with 2-5 blades and
                500-4,000 panels, each with a helix wake of 50k-400k
                particles (pitch 0.4-0.6)
Each generator returns plain numpy arrays plus the octree settings that go
with them.
"""
import numpy as np

HELIX_SEED = 20261203
ELLIPSOID_SEED = 20261204


def helix_wake(n=1_000_000, seed=HELIX_SEED, tube_fraction=0.9, radius=1.0, pitch=0.5,
               sigma=0.05, core=0.02):
    """Three helical vortex tubes (radius 1, pitch 0.5, Gaussian cross-section
    sigma) with tangent strengths ~N(1e-3, 2e-4), plus uniform background
    particles in [-2,2]^3 with isotropic N(0, 1e-4^2) strengths.

    Returns (cen (n,3), alpha (n,3), r_vortex (n,), octree dict)."""
    rng = np.random.default_rng(seed)
    nt = int(round(n * tube_fraction))
    nb = n - nt
    k = rng.integers(0, 3, nt)
    phi = 2.0 * np.pi * k / 3.0
    c = pitch / (2.0 * np.pi)
    s = rng.uniform(0.0, 3.5 / c, nt)  # z from -1.75 to 1.75 (14 pi at pitch 0.5)
    a = s + phi
    ca, sa = np.cos(a), np.sin(a)
    centre = np.stack([radius * ca, radius * sa, -1.75 + c * s], axis=1)
    t = np.stack([-radius * sa, radius * ca, np.full(nt, c)], axis=1)
    t /= np.linalg.norm(t, axis=1)[:, None]
    nrm = np.stack([-ca, -sa, np.zeros(nt)], axis=1)
    bn = np.cross(t, nrm)
    g = rng.standard_normal((nt, 2))
    pos_t = centre + sigma * (g[:, :1] * nrm + g[:, 1:] * bn)
    m = rng.normal(1e-3, 2e-4, nt)
    alpha_t = m[:, None] * t
    pos_b = rng.uniform(-2.0, 2.0, (nb, 3))
    alpha_b = rng.normal(0.0, 1e-4, (nb, 3))
    cen = np.concatenate([pos_t, pos_b])
    alpha = np.concatenate([alpha_t, alpha_b])
    perm = rng.permutation(n)
    cen = np.ascontiguousarray(cen[perm])
    alpha = np.ascontiguousarray(alpha[perm])
    rv = np.full(n, core)
    tree = dict(box_length=4.0 + 2e-6, nbox=(1, 1, 1), origin=(-2.0 - 1e-6,) * 3, nlevels=7,
                min_part=64, delta=core)
    return cen, alpha, rv, tree


def ellipsoid(nt=128, npf=128, axes=(4.0, 1.0, 0.5), jitter=0.01, seed=ELLIPSOID_SEED):
    """Closed ellipsoid surface, nt rows in theta x npf columns in phi; the
    two pole rows are triangles (n_ver = 3), the rest quads; outward normals.
    Non-pole vertices are jittered by jitter * (local minimum edge) * U[-1,1]^3.

    Returns (ver (n,4,3), n_ver (n,))."""
    rng = np.random.default_rng(seed)
    th = np.pi * np.arange(nt + 1) / nt
    ph = 2.0 * np.pi * np.arange(npf) / npf
    T, P = np.meshgrid(th[1:-1], ph, indexing="ij")  # interior rings
    ring = np.stack([axes[0] * np.sin(T) * np.cos(P), axes[1] * np.sin(T) * np.sin(P),
                     axes[2] * np.cos(T)], axis=-1)  # (nt-1, npf, 3)
    # local minimum edge of every interior vertex
    d_phi = np.linalg.norm(ring - np.roll(ring, 1, axis=1), axis=-1)
    d_phi = np.minimum(d_phi, np.roll(d_phi, -1, axis=1))
    north = np.array([0.0, 0.0, axes[2]])
    south = np.array([0.0, 0.0, -axes[2]])
    up = np.concatenate([np.linalg.norm(ring[:1] - north, axis=-1),
                         np.linalg.norm(ring[1:] - ring[:-1], axis=-1)])
    dn = np.concatenate([np.linalg.norm(ring[1:] - ring[:-1], axis=-1),
                         np.linalg.norm(ring[-1:] - south, axis=-1)])
    lmin = np.minimum(np.minimum(d_phi, up), dn)
    ring = ring + jitter * lmin[..., None] * rng.uniform(-1.0, 1.0, ring.shape)
    ver = np.zeros((nt * npf, 4, 3))
    nv = np.zeros(nt * npf, np.int32)
    q = 0
    for i in range(nt):
        for j in range(npf):
            j1 = (j + 1) % npf
            if i == 0:
                v = [north, ring[0, j], ring[0, j1]]
            elif i == nt - 1:
                v = [ring[nt - 2, j], south, ring[nt - 2, j1]]
            else:
                v = [ring[i - 1, j], ring[i, j], ring[i, j1], ring[i - 1, j1]]
            ver[q, : len(v)] = v
            nv[q] = len(v)
            q += 1
    # orient outward: normal as in calc_geo_data_surfpan, flipped if inward
    for p in range(len(nv)):
        v = ver[p, : nv[p]]
        if nv[p] == 4:
            nr = np.cross(v[2] - v[0], v[3] - v[1])
        else:
            nr = np.cross(v[2] - v[1], v[0] - v[1])
        if np.dot(nr, v.mean(axis=0)) < 0.0:
            ver[p, : nv[p]] = v[::-1]
    return ver, nv


def freestreams(nrhs=8, seed=ELLIPSOID_SEED):
    """Unit free-stream directions for the right-hand sides."""
    rng = np.random.default_rng(seed + 1)
    u = rng.standard_normal((nrhs, 3))
    return u / np.linalg.norm(u, axis=1)[:, None]


def naca0012(x):
    """Half-thickness of the NACA 0012 section (chord 1)."""
    return 0.6 * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3
                  - 0.1015 * x**4)


def wing_case(nc=40, ns=20, aspect=8.0, alpha_deg=5.0, u=30.0, nrows=200, seed=7):
    """Rectangular NACA0012 wing (chord 1, span = aspect) of 2*nc*ns quads with
    cosine chordwise spacing, pitched by alpha, and a trailing sheet of
    nrows x (2*ns) particles spaced by dt*U = chord/20 behind the trailing
    edge with an elliptic circulation distribution.  Core radius 0.1 chord."""
    xc = 0.5 * (1.0 - np.cos(np.linspace(0.0, np.pi, nc + 1)))
    ys = np.linspace(-aspect / 2, aspect / 2, ns + 1)
    zt = naca0012(xc)
    ca, sa = np.cos(np.radians(alpha_deg)), np.sin(np.radians(alpha_deg))

    def pt(x, y, z):
        return np.array([x * ca + z * sa, y, -x * sa + z * ca])

    ver = []
    for j in range(ns):
        for i in range(nc):  # upper surface, trailing edge -> leading edge order
            a, b = nc - i, nc - i - 1
            ver.append([pt(xc[a], ys[j], zt[a]), pt(xc[b], ys[j], zt[b]),
                        pt(xc[b], ys[j + 1], zt[b]), pt(xc[a], ys[j + 1], zt[a])])
        for i in range(nc):  # lower surface
            a, b = i, i + 1
            ver.append([pt(xc[a], ys[j], -zt[a]), pt(xc[b], ys[j], -zt[b]),
                        pt(xc[b], ys[j + 1], -zt[b]), pt(xc[a], ys[j + 1], -zt[a])])
    # reversed vertex order: outward normals (calc_geo_data_surfpan's
    # (v3 - v1) x (v4 - v2)); the trailing-edge points stay vertices 1 and 4
    # of the upper panels
    ver = np.ascontiguousarray(np.array(ver)[:, ::-1])
    nv = np.full(len(ver), 4, np.int32)
    # particle sheet
    rng = np.random.default_rng(seed)
    dx = 1.0 / 20.0
    nsp = 2 * ns
    yp = (np.arange(nsp) + 0.5) / nsp * aspect - aspect / 2
    gam = np.sqrt(np.maximum(1.0 - (2.0 * yp / aspect) ** 2, 0.0))
    dgam = -np.gradient(gam, yp)
    cen, alpha = [], []
    te = pt(1.0, 0.0, 0.0)
    for r in range(nrows):
        x = te[0] + dx * (r + 0.5)
        for s in range(nsp):
            cen.append([x, yp[s], te[2] + 1e-3 * rng.standard_normal()])
            # trailing vorticity along x: -dGamma/dy * dy * dx (Gamma ~ 0.1 U c)
            alpha.append([-0.1 * u * dgam[s] * (aspect / nsp) * dx, 0.0, 0.0])
    cen = np.array(cen)
    alpha = np.array(alpha)
    alpha[:, 1] = 1e-3 * rng.standard_normal(len(cen))
    rv = np.full(len(cen), 0.1)
    lo = np.minimum(cen.min(axis=0), ver.reshape(-1, 3).min(axis=0)) - 0.5
    hi = np.maximum(cen.max(axis=0), ver.reshape(-1, 3).max(axis=0)) + 0.5
    box = float((hi - lo).max())
    tree = dict(box_length=box, nbox=(1, 1, 1), origin=tuple(lo), nlevels=5, min_part=64,
                delta=0.1)
    return ver, nv, cen, alpha, rv, tree


def rotor_strips(blades, panels, aspect):
    """(nc, ns) of rotor(): chordwise panels per surface and spanwise strips
    per blade.  The mesh is blades x ns strips of 2*nc quads."""
    per = max(8, panels // blades)
    ns = max(2, int(round(np.sqrt(per * aspect / 8.0))))
    nc = max(2, int(round(per / (2 * ns))))
    return nc, ns


def rotor(blades=3, panels=1000, aspect=10.0, root=0.2, pitch_deg=8.0):
    """`blades` rectangular NACA0012 blades (chord 1, span = aspect), root at
    radius `root`, pitched by pitch_deg, evenly spaced around the z axis;
    about `panels` quads in total (upper and lower surfaces, cosine chordwise
    spacing).  Returns (ver (n,4,3), n_ver)."""
    nc, ns = rotor_strips(blades, panels, aspect)
    xc = 0.5 * (1.0 - np.cos(np.linspace(0.0, np.pi, nc + 1))) - 0.25
    rs = np.linspace(root, root + aspect, ns + 1)
    zt = naca0012(xc + 0.25)
    cp, sp = np.cos(np.radians(pitch_deg)), np.sin(np.radians(pitch_deg))
    ver = []
    for b in range(blades):
        th = 2.0 * np.pi * b / blades
        ct, st = np.cos(th), np.sin(th)

        def pt(x, r, z):
            xx, zz = x * cp + z * sp, -x * sp + z * cp  # blade pitch about its span axis
            return np.array([r * ct - xx * st, r * st + xx * ct, zz])
        for j in range(ns):
            for i in range(nc):
                a, c = nc - i, nc - i - 1
                ver.append([pt(xc[a], rs[j], zt[a]), pt(xc[c], rs[j], zt[c]),
                            pt(xc[c], rs[j + 1], zt[c]), pt(xc[a], rs[j + 1], zt[a])])
            for i in range(nc):
                a, c = i, i + 1
                ver.append([pt(xc[a], rs[j], -zt[a]), pt(xc[c], rs[j], -zt[c]),
                            pt(xc[c], rs[j + 1], -zt[c]), pt(xc[a], rs[j + 1], -zt[a])])
    ver = np.array(ver)
    return ver, np.full(len(ver), 4, np.int32)


def batch_cases(ncase=48, seed=20261205):
    """The many-case workload: ncase rotor cases, each with 2-5 blades and
    500-4,000 surface panels and a helix wake of 50k-400k particles (pitch
    0.4-0.6, seed varied per case).  Returns a list of dicts with the rotor
    mesh (ver, n_ver) and the wake (cen, alpha, rv, tree)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(ncase):
        blades = int(rng.integers(2, 6))
        panels = int(rng.integers(500, 4001))
        n = int(rng.integers(50_000, 400_001))
        pitch = float(rng.uniform(0.4, 0.6))
        ver, nv = rotor(blades, panels)
        cen, alpha, rv, tree = helix_wake(n, seed=seed + 1 + k, pitch=pitch)
        # the rare tube particles beyond 5 sigma of the cross-section are kept
        # inside the octree box
        cen = np.clip(cen, -2.0 + 1e-3, 2.0 - 1e-3)
        tree = dict(tree, nlevels=7 - int(round(np.log(1_000_000 / n) / np.log(8))))
        out.append(dict(blades=blades, ver=ver, n_ver=nv, cen=cen, alpha=alpha, rv=rv,
                        tree=tree))
    return out


def element_lists(ver, n_ver, te, u_inf=(30.0, 0.0, 0.0), dx=0.05, wake_dir=(1.0, 0.0, 0.0),
                  seed=11):
    """Element lists of apply_multipole around the surface-panel mesh
    (ver, n_ver) with the trailing edge te (k+1, 3) (a polyline of k
    segments): a list of dicts (kind 0 surface panel, 1 vortex ring,
    2 vortex line; the kinds of dust_amd.elements) in the reference's order
    elem, wpan, wrin, wvort.

    - elem: the surface panels with a smooth doublet strength and the source
      strength nor.(-u_inf) of a body at rest;
    - wpan: one row of wake panels (quads te_i, te_i + d, te_i+1 + d, te_i+1,
      d = dx wake_dir, wake_dir a unit vector across the trailing edge) with
      circulation Gamma_i;
    - wrin: the same row moved one more dx downstream (vortex rings that take
      no part in the stretching term);
    - wvort: the k end vortices (te_i + 2 d -> te_i+1 + 2 d) closing it."""
    rng = np.random.default_rng(seed)
    ver = np.asarray(ver, dtype=np.float64)
    n_ver = np.asarray(n_ver, dtype=np.int32)
    cen = np.array([ver[i, :n_ver[i]].mean(axis=0) for i in range(len(ver))])
    a = np.where(n_ver[:, None] == 4, ver[:, 2] - ver[:, 0], ver[:, 2] - ver[:, 1])
    b = np.where(n_ver[:, None] == 4, ver[:, 3] - ver[:, 1], ver[:, 0] - ver[:, 1])
    nor = np.cross(a, b)
    nor /= np.linalg.norm(nor, axis=1)[:, None]
    mag = 0.5 * np.sin(cen[:, 0] * 1.3) + 0.2 * np.cos(cen[:, 1] * 0.7) \
        + 0.05 * rng.standard_normal(len(cen))
    sigma = nor @ (-np.asarray(u_inf, dtype=np.float64))
    te = np.asarray(te, dtype=np.float64)
    k = len(te) - 1
    s = np.linspace(-1.0, 1.0, k)
    gam = np.sqrt(np.maximum(1.0 - s**2, 0.0)) + 0.1
    d = dx * np.asarray(wake_dir, dtype=np.float64)
    wp = np.array([[te[i], te[i] + d, te[i + 1] + d, te[i + 1]] for i in range(k)])
    wr = wp + d
    wl = np.array([[te[i] + 2 * d, te[i + 1] + 2 * d] for i in range(k)])
    return [dict(kind=0, ver=ver, n_ver=n_ver, mag=mag, sigma=sigma, stretch=True),
            dict(kind=1, ver=wp, n_ver=np.full(k, 4, np.int32), mag=gam, stretch=True),
            dict(kind=1, ver=wr, n_ver=np.full(k, 4, np.int32), mag=0.8 * gam, stretch=False),
            dict(kind=2, ver=wl, n_ver=np.full(k, 2, np.int32), mag=0.8 * gam, stretch=True)]


def wing_elements(nc=40, ns=20, aspect=8.0, alpha_deg=5.0, u=30.0):
    """element_lists of the wing_case wing (trailing edge at chord 1)."""
    ver, nv, *_ = wing_case(nc, ns, aspect, alpha_deg, u, nrows=1)
    ca, sa = np.cos(np.radians(alpha_deg)), np.sin(np.radians(alpha_deg))
    ys = np.linspace(-aspect / 2, aspect / 2, ns + 1)
    te = np.array([[ca, y, -sa] for y in ys])
    return element_lists(ver, nv, te, u_inf=(u, 0.0, 0.0))


def rotate_z(x, theta):
    """Points x (..., 3) rotated by theta about the z axis through the origin."""
    c, s_ = np.cos(theta), np.sin(theta)
    x = np.asarray(x, dtype=np.float64)
    out = x.copy()
    out[..., 0] = c * x[..., 0] - s_ * x[..., 1]
    out[..., 1] = s_ * x[..., 0] + c * x[..., 1]
    return out
