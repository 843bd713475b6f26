"""GPU parity of the FMM path (libdust_gpu through the C ABI) against the CPU
oracle: same inputs, FP64, relative L2 <= 1e-10 on velocity and stretching;
Morton/cell keys and leaf membership bit-exact."""
import os

import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import octree as O
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-10


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run_both(orc, cen, alpha, rv, tree, degree, **flags):
    parts = O.VortParts.from_alpha(cen, alpha, rv)
    print(f"case n={parts.n} levels={tree['nlevels']} p={degree} flags={flags}", flush=True)
    g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                            tree["min_part"], degree, tree["delta"], capacity=max(parts.n, 1),
                            **flags)
    nl = O.sort_particles(parts, g)
    O.calculate_multipole(parts, g)
    O.apply_multipole(parts, g)
    oflags = {k: v for k, v in flags.items() if k != "precision"}
    if "kernel" in oflags:
        oflags["kernel"] = {"rosenhead": orc.ROSENHEAD, "gaussian": orc.GAUSSIAN}[oflags["kernel"]]
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                   tree["min_part"], degree, tree["delta"], **oflags)
    v, s, tv = c.run(parts.cen, parts.dir, parts.mag, parts.r_vortex)
    return parts, g, nl, c, v, s, tv


def check_leaves(parts, g, c):
    lev_g, ijk_g, np_g, mem_g = g.leaves(parts.n)
    lev_c, ijk_c, np_c, mem_c = c.leaves()
    key_g = {(int(a), *map(int, b)): int(n) for a, b, n in zip(lev_g, ijk_g, np_g)}
    key_c = {(int(a), *map(int, b)): int(n) for a, b, n in zip(lev_c, ijk_c, np_c)}
    assert key_g == key_c
    # membership: same particles in every leaf (order inside a leaf is the
    # stable input order on both sides)
    def groups(lev, ijk, npart, mem):
        out, w = {}, 0
        for a, b, n in zip(lev, ijk, npart):
            out[(int(a), *map(int, b))] = tuple(mem[w:w + n])
            w += n
        return out
    assert groups(lev_g, ijk_g, np_g, mem_g) == groups(lev_c, ijk_c, np_c, mem_c)


@pytest.mark.parametrize("nlev,p", [(1, 2), (3, 2), (4, 4), (5, 4), (4, 6), (4, 8)])
def test_helix_vs_oracle(orc, nlev, p):
    cen, alpha, rv, tree = W.helix_wake(12000, seed=11)
    tree = dict(tree, nlevels=nlev)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, p, use_vs=True)
    assert rel(parts.vel, v) <= TOL
    assert rel(parts.stretch, s) <= TOL
    check_leaves(parts, g, c)
    assert np.array_equal(g.work(), c.work())


def test_reg_test_b_particles(orc):
    """Particle field of the reference's FMM regression case with its octree
    settings (tests/reg_test_b/dust_base.in)."""
    d = np.load(os.path.join(HERE, "golden", "reg_test_b_particles.npz"))
    tree = dict(box_length=float(d["box_length"]), nbox=tuple(d["nbox"]), origin=tuple(d["origin"]),
                nlevels=int(d["nlevels"]), min_part=int(d["min_part"]), delta=float(d["rankine_rad"]))
    rv = np.full(len(d["cen"]), float(d["vortex_rad"]))
    for vs in (False, True):
        parts, g, nl, c, v, s, tv = run_both(orc, d["cen"], d["alpha"], rv, tree, int(d["degree"]),
                                             use_vs=vs)
        assert rel(parts.vel, v) <= TOL
        if vs:
            assert rel(parts.stretch, s) <= TOL
        check_leaves(parts, g, c)


def test_diffusion_turbvisc_freestream(orc):
    cen, alpha, rv, tree = W.helix_wake(6000, seed=5)
    rv = rv * (1.0 + 0.5 * np.random.default_rng(1).random(len(rv)))  # per-particle radii
    tree = dict(tree, nlevels=4)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, 4, use_vs=True, use_vd=True,
                                         use_tv=True, nu_inf=1e-3, u_inf=(1.0, -0.5, 0.25))
    assert rel(parts.vel, v) <= TOL
    assert rel(parts.stretch, s) <= TOL
    assert rel(parts.turbvisc, tv) <= TOL


def test_multibox_nbox(orc):
    """nbox > 1: level-1 interaction list over the whole box array."""
    rng = np.random.default_rng(3)
    n = 5000
    cen = rng.uniform([0.0, 0.0, 0.0], [3.0, 2.0, 1.0], (n, 3))
    alpha = rng.normal(0, 1e-3, (n, 3))
    rv = np.full(n, 0.03)
    tree = dict(box_length=1.0, nbox=(3, 2, 1), origin=(-1e-7,) * 3, nlevels=3, min_part=16,
                delta=0.03)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, 4, use_vs=True)
    assert rel(parts.vel, v) <= TOL
    assert rel(parts.stretch, s) <= TOL
    check_leaves(parts, g, c)


def test_cell_index_bitexact():
    """Finest-cell index = ceiling((x - xmin)/csize), including points that sit
    exactly on cell faces."""
    rng = np.random.default_rng(9)
    n = 4096
    box, lev, org = 2.0, 5, -1.0
    cs = box / 2 ** (lev - 1)
    grid = org + cs * rng.integers(1, 2 ** (lev - 1), (n, 3)).astype(np.float64)
    cen = np.where(rng.random((n, 3)) < 0.5, grid, rng.uniform(org + 1e-9, org + box, (n, 3)))
    parts = O.VortParts(cen, np.tile([1.0, 0, 0], (n, 1)), np.ones(n), np.full(n, 0.01))
    g = O.initialize_octree(box, (1, 1, 1), (org,) * 3, lev, 8, 2, 0.01, capacity=n)
    O.sort_particles(parts, g)
    ijk = g.cell_index(n)
    ref = np.ceil((cen - org) / cs).astype(np.int64)
    assert np.array_equal(ijk, ref)


def test_outside_box_raises():
    parts = O.VortParts(np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0]]), np.eye(3)[:2], np.ones(2),
                        np.full(2, 0.1))
    g = O.initialize_octree(1.0, (1, 1, 1), (-0.5,) * 3, 3, 1, 2, 0.1, capacity=4)
    with pytest.raises(O.L.DustError) as e:
        O.sort_particles(parts, g)
    assert e.value.code == O.L.DUST_E_OUTSIDE


def test_empty_and_capacity():
    g = O.initialize_octree(1.0, (1, 1, 1), (-0.5,) * 3, 3, 1, 2, 0.1, capacity=4)
    parts = O.VortParts(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    assert O.sort_particles(parts, g) == 1
    O.calculate_multipole(parts, g)
    O.apply_multipole(parts, g)
    big = O.VortParts(np.zeros((5, 3)), np.tile([1.0, 0, 0], (5, 1)), np.ones(5), np.full(5, 0.1))
    with pytest.raises(O.L.DustError) as e:
        O.sort_particles(big, g)
    assert e.value.code == O.L.DUST_E_CAPACITY


def test_direct_vs_oracle(orc):
    d = np.load(os.path.join(HERE, "golden", "mir_static_p_particles.npz"))
    parts = O.VortParts.from_alpha(d["cen"], d["alpha"], d["r_vortex"])
    vg, sg = O.direct(parts, use_vs=True, use_vd=True, nu_inf=1e-2, u_inf=(10.0, 0.0, 1.0))
    vc, sc = orc.direct(parts.cen, parts.dir, parts.mag, parts.r_vortex, True, True, 1e-2,
                        (10.0, 0.0, 1.0))
    assert rel(vg, vc) <= TOL
    assert rel(sg, sc) <= TOL


@pytest.mark.parametrize("form", ["fp64", "tuned"])
def test_gaussian_core_vs_oracle(orc, form):
    """The Gaussian particle core (kernel_coeffs' commented alternative,
    mod_vortpart.f90:283-291) in P2P: FP64 form <= 1e-10, tuned form (FP32
    erf pairs beyond one core radius, FP64 inside) within the tuned bounds."""
    cen, alpha, rv, tree = W.helix_wake(12000, seed=11)
    tree = dict(tree, nlevels=4)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, 4, use_vs=True,
                                         precision=form, kernel="gaussian")
    ev, es = rel(parts.vel, v), rel(parts.stretch, s)
    print(f"gaussian {form}: vel {ev:.2e} stretch {es:.2e}")
    assert ev <= (TOL if form == "fp64" else TUNED_VEL)
    assert es <= (TOL if form == "fp64" else TUNED_STR)
    check_leaves(parts, g, c)
    # the core matters: the Rosenhead oracle differs
    cr = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], 4, tree["min_part"], 4,
                    tree["delta"], use_vs=True)
    vr, _, _ = cr.run(parts.cen, parts.dir, parts.mag, parts.r_vortex)
    assert rel(vr, v) > 1e-4


def test_gaussian_direct_vs_oracle(orc):
    d = np.load(os.path.join(HERE, "golden", "mir_static_p_particles.npz"))
    parts = O.VortParts.from_alpha(d["cen"], d["alpha"], d["r_vortex"])
    vg, sg = O.direct(parts, use_vs=True, u_inf=(10.0, 0.0, 1.0), kernel="gaussian")
    vc, sc = orc.direct(parts.cen, parts.dir, parts.mag, parts.r_vortex, True, False, 0.0,
                        (10.0, 0.0, 1.0), kernel=orc.GAUSSIAN)
    assert rel(vg, vc) <= TOL
    assert rel(sg, sc) <= TOL


def test_repeat_calls_reuse_handle(orc):
    """sort/multipole/apply twice on one handle with different particle sets."""
    cen, alpha, rv, tree = W.helix_wake(8000, seed=2)
    tree = dict(tree, nlevels=4)
    g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], 4, 64, 4,
                            tree["delta"], capacity=8000)
    for m in (8000, 3000):
        parts = O.VortParts.from_alpha(cen[:m], alpha[:m], rv[:m])
        O.sort_particles(parts, g)
        O.calculate_multipole(parts, g)
        O.apply_multipole(parts, g)
        c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], 4, 64, 4, tree["delta"])
        v, s, _ = c.run(parts.cen, parts.dir, parts.mag, parts.r_vortex)
        assert rel(parts.vel, v) <= TOL
        assert rel(parts.stretch, s) <= TOL


# ---- tuned form: FP32 P2P with FP64 near pairs + FP32 MFMA M2L -------------------
TUNED_VEL = 1e-6   # relative L2 velocity error vs the FP64 oracle
TUNED_STR = 2e-5   # relative L2 stretching error vs the FP64 oracle


@pytest.mark.parametrize("nlev,p", [(4, 4), (5, 4), (4, 6), (4, 8), (3, 2)])
def test_tuned_vs_oracle(orc, nlev, p):
    cen, alpha, rv, tree = W.helix_wake(12000, seed=11)
    tree = dict(tree, nlevels=nlev)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, p, use_vs=True,
                                         precision="tuned")
    ev, es = rel(parts.vel, v), rel(parts.stretch, s)
    print(f"tuned nlev={nlev} p={p}: vel {ev:.2e} stretch {es:.2e}")
    assert ev <= TUNED_VEL
    assert es <= TUNED_STR
    check_leaves(parts, g, c)


def test_tuned_coincident_particles(orc):
    """Distinct particles at the same position: the tuned P2P drops only the
    self pair (by index) and evaluates the coincident pair in FP64, whose
    stretching term is not zero under the Rosenhead core."""
    cen, alpha, rv, tree = W.helix_wake(6000, seed=21)
    rng = np.random.default_rng(22)
    k = 300
    cen = np.concatenate([cen, cen[:k]])
    alpha = np.concatenate([alpha, rng.normal(0.0, 1e-3, (k, 3))])
    rv = np.concatenate([rv, rv[:k]])
    tree = dict(tree, nlevels=4)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, 4, use_vs=True,
                                         precision="tuned")
    dup = np.r_[0:k, len(rv) - k:len(rv)]
    ev, es, ed = rel(parts.vel, v), rel(parts.stretch, s), rel(parts.stretch[dup], s[dup])
    print(f"coincident: vel {ev:.2e} stretch {es:.2e} duplicated particles {ed:.2e}")
    assert ev <= TUNED_VEL
    assert es <= TUNED_STR and ed <= TUNED_STR


def test_tuned_velocity_only_and_nbox(orc):
    rng = np.random.default_rng(3)
    n = 5000
    cen = rng.uniform([0.0, 0.0, 0.0], [3.0, 2.0, 1.0], (n, 3))
    alpha = rng.normal(0, 1e-3, (n, 3))
    rv = np.full(n, 0.03)
    tree = dict(box_length=1.0, nbox=(3, 2, 1), origin=(-1e-7,) * 3, nlevels=4, min_part=16,
                delta=0.03)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, 4, use_vs=False,
                                         precision="tuned")
    assert rel(parts.vel, v) <= TUNED_VEL


def test_tuned_with_diffusion_falls_back_exact(orc):
    """The PSE diffusion and turbulent viscosity run in FP64 in the tuned form."""
    cen, alpha, rv, tree = W.helix_wake(4000, seed=5)
    tree = dict(tree, nlevels=4)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, 4, use_vs=True, use_vd=True,
                                         use_tv=True, nu_inf=1e-3, precision="tuned")
    assert rel(parts.vel, v) <= TUNED_VEL
    assert rel(parts.stretch, s) <= TUNED_STR
    assert rel(parts.turbvisc, tv) <= 1e-5


def test_tuned_m2l_chunk_knob(orc):
    """The M2L job size changes how the offsets are split between jobs.  A job
    sums its offsets in FP32 before adding into the FP64 locals, so results
    differ at FP32 rounding; each stays within the tuned bounds of the FP64
    oracle."""
    cen, alpha, rv, tree = W.helix_wake(12000, seed=11)
    tree = dict(tree, nlevels=5)
    out = {}
    for chunk in (None, 7, 189):
        parts = O.VortParts.from_alpha(cen, alpha, rv)
        print(f"m2l_chunk {chunk}", flush=True)
        g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                                tree["min_part"], 4, tree["delta"], capacity=parts.n, use_vs=True,
                                precision="tuned")
        if chunk:
            g.set_tuning("m2l_chunk", chunk)
        O.sort_particles(parts, g)
        O.calculate_multipole(parts, g)
        O.apply_multipole(parts, g)
        out[chunk] = (parts.vel.copy(), parts.stretch.copy())
        if chunk == 7:
            with pytest.raises(L.DustError):
                g.set_tuning("m2l_chunk", 0)
            with pytest.raises(L.DustError):
                g.set_tuning("m2l_chunk", 190)
        g.destroy()
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                   tree["min_part"], 4, tree["delta"], use_vs=True)
    parts = O.VortParts.from_alpha(cen, alpha, rv)
    v, st, _ = c.run(parts.cen, parts.dir, parts.mag, parts.r_vortex)
    for chunk, (vel, stretch) in out.items():
        print(f"chunk {chunk}: vel {rel(vel, v):.2e} stretch {rel(stretch, st):.2e}")
        assert rel(vel, v) <= TUNED_VEL
        assert rel(stretch, st) <= TUNED_STR
    for chunk in (7, 189):
        assert rel(out[chunk][0], out[None][0]) <= 5e-7
        assert rel(out[chunk][1], out[None][1]) <= 5e-6


@pytest.mark.parametrize("precision", ["fp64", "tuned"])
def test_autotune(precision):
    """The autotuner only picks performance knobs from its grid; a pass with
    the chosen setting gives the default setting's results up to the
    regrouping of partial sums (FP64 items; FP32 M2L jobs in the tuned
    form)."""
    cen, alpha, rv, tree = W.helix_wake(20000, seed=13)
    tree = dict(tree, nlevels=5)

    def handle():
        return O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"],
                                   tree["nlevels"], tree["min_part"], 4, tree["delta"],
                                   capacity=len(cen), use_vs=True, precision=precision)

    def run(g):
        parts = O.VortParts.from_alpha(cen, alpha, rv)
        O.sort_particles(parts, g)
        O.calculate_multipole(parts, g)
        O.apply_multipole(parts, g)
        return parts.vel.copy(), parts.stretch.copy()

    g0 = handle()
    v0, s0 = run(g0)
    g0.destroy()
    g = handle()
    parts = O.VortParts.from_alpha(cen, alpha, rv)
    O.sort_particles(parts, g)
    print(f"autotune {precision}", flush=True)
    ms = g.autotune(reps=1)
    seg = g.get_tuning("seg_iters")
    print(f"{precision}: {ms:.3f} ms/pass, seg_iters {seg}")
    assert 0.0 < ms < 1e4
    assert seg in (768, 1536, 3072)
    if precision == "tuned":
        assert g.get_tuning("m2l_chunk") in (7, 14, 21, 27)
        assert g.get_tuning("near_factor") == 1.0
    else:
        with pytest.raises(L.DustError):
            g.get_tuning("m2l_chunk")
    with pytest.raises(L.DustError):
        g.set_tuning("seg_iters", 10)
    v, s = run(g)
    g.destroy()
    tol_v, tol_s = (1e-13, 1e-13) if precision == "fp64" else (5e-7, 5e-6)
    assert rel(v, v0) <= tol_v
    assert rel(s, s0) <= tol_s


@pytest.mark.parametrize("form", ["fp64", "tuned"])
def test_apply_multipole_panels_vs_oracle(orc, form):
    """apply_multipole_panels (use_fmm_pan) at panel-like points in and around
    the helix wake, after a pass: FP64 form to 1e-10, tuned form to 1e-6 of
    the oracle; uvort is accumulated; host and device entry points agree."""
    import torch
    cen, alpha, rv, tree = W.helix_wake(20000, seed=12)
    tree = dict(tree, nlevels=4)
    parts, g, nl, c, v, s, tv = run_both(orc, cen, alpha, rv, tree, 4, use_vs=True,
                                         precision=form)
    rng = np.random.default_rng(13)
    pts = np.clip(cen[rng.choice(len(cen), 3000, replace=False)] +
                  0.02 * rng.standard_normal((3000, 3)), -1.999, 1.999)
    pts = np.concatenate([pts, rng.uniform(-1.999, 1.999, (500, 3))])
    u0 = rng.standard_normal(pts.shape)
    print(f"apply_multipole_panels form={form} m={len(pts)}", flush=True)
    u = O.apply_multipole_panels(g, pts, u0)
    uo = c.apply_panels(pts, u0)
    tol = TOL if form == "fp64" else 1e-6
    assert rel(u - u0, uo - u0) <= tol
    tp = torch.from_numpy(np.ascontiguousarray(pts)).cuda()
    tu = torch.from_numpy(np.ascontiguousarray(u0)).cuda()
    L.check(L.lib().dust_fmm_apply_multipole_panels_dev(g.handle, len(pts), tp.data_ptr(),
                                                        tu.data_ptr()), "panels_dev")
    L.check(L.lib().dust_fmm_synchronize(g.handle), "sync")
    assert np.array_equal(tu.cpu().numpy(), u)
    with pytest.raises(L.DustError) as ei:
        O.apply_multipole_panels(g, np.array([[5.0, 0.0, 0.0]]))
    assert ei.value.code == L.DUST_E_OUTSIDE
