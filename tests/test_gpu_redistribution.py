"""GPU particle redistribution (sort_particles, mod_octree.f90:827-894)
against the oracle on the same inputs: the freed particles and the new
strengths bit-exact (the same unfused FP64 operations in the same order),
leaves and membership bit-exact, the FMM results within the FMM tolerances;
the tree after a redistribution equals the tree of the survivors; passes
with the redistribution switched on, off and on again, and a batch holding a
redistributing case, repeat the single passes."""
import ctypes as C

import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import octree as O

pytestmark = pytest.mark.gpu

TREE = dict(box_length=1.0, nbox=(2, 1, 1), origin=(-1.0, -0.5, -0.5), nlevels=5, min_part=16,
            delta=0.01)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def cloud(n, seed):
    """Uniform background plus four dense blobs; log-normal magnitudes."""
    rng = np.random.default_rng(seed)
    nb = n // 2
    cen = np.vstack([rng.uniform([-0.999, -0.499, -0.499], [0.999, 0.499, 0.499], size=(nb, 3))]
                    + [np.clip(c + 0.05 * rng.normal(size=((n - nb) // 4, 3)),
                               [-0.999, -0.499, -0.499], [0.999, 0.499, 0.499])
                       for c in ([-0.6, 0.2, 0.1], [0.5, -0.2, 0.0], [0.05, 0.0, -0.3],
                                 [0.7, 0.3, 0.3])])
    n = len(cen)
    a = rng.normal(size=(n, 3))
    dir = a / np.linalg.norm(a, axis=1)[:, None]
    mag = rng.lognormal(0.0, 1.2, size=n)
    return cen, dir, mag, np.full(n, 0.02)


def panels(m=400):
    """Element centres of a flat plate at z = 0 (near-wall cells around it)."""
    u, v = np.meshgrid(np.linspace(-0.45, 0.45, 20), np.linspace(-0.1, 0.1, m // 20))
    return np.column_stack([u.ravel(), v.ravel(), np.zeros(u.size)])


def gpu_pass(cen, dir, mag, rv, precision="fp64", pr=True, elem=None, lvl=3, ratio=3.0):
    parts = O.VortParts(cen, dir, mag, rv)
    g = O.initialize_octree(TREE["box_length"], TREE["nbox"], TREE["origin"], TREE["nlevels"],
                            TREE["min_part"], 4, TREE["delta"], capacity=parts.n,
                            precision=precision, particles_redistribution=pr,
                            octree_level_solid=lvl, particles_redistribution_ratio=ratio)
    nl = O.sort_particles(parts, g, elem=elem)
    O.calculate_multipole(parts, g)
    O.apply_multipole(parts, g)
    return parts, g, nl


def oracle_pass(orc, cen, dir, mag, rv, elem, lvl=3, ratio=3.0):
    c = orc.Octree(TREE["box_length"], TREE["nbox"], TREE["origin"], TREE["nlevels"],
                   TREE["min_part"], 4, TREE["delta"])
    c.set_redistribution(lvl, ratio, elem)
    v, s, _ = c.run(cen, dir, mag, rv)
    return c, v, s


def leaf_groups(lev, ijk, npart, mem):
    out, w = {}, 0
    for a, b, n in zip(lev, ijk, npart):
        out[(int(a), *map(int, b))] = tuple(int(x) for x in mem[w:w + n])
        w += n
    return out


@pytest.mark.parametrize("precision,ratio", [("fp64", 3.0), ("fp64", 10.0), ("tuned", 3.0)])
def test_redistribution_vs_oracle(orc, precision, ratio):
    cen, dir, mag, rv = cloud(40000, seed=21)
    elem = panels()
    print(f"redistribution n={len(cen)} {precision} ratio={ratio}", flush=True)
    parts, g, nl = gpu_pass(cen, dir, mag, rv, precision, elem=elem, ratio=ratio)
    c, v, s = oracle_pass(orc, cen, dir, mag, rv, elem, ratio=ratio)
    keep, d, m, nk = c.redistributed()
    print(f"kept {nk} of {len(cen)}", flush=True)
    assert 0 < nk < len(cen)
    assert np.array_equal(~parts.free, keep) and parts.n_prt == nk
    assert np.array_equal(parts.mag, m) and np.array_equal(parts.dir, d)
    assert nl == orc.lib().orc_octree_nleaves(c._h)
    lev, ijk, npart, mem = g.leaves(nk)
    assert leaf_groups(lev, ijk, npart, mem) == leaf_groups(*c.leaves()[:3], c.leaves()[3][:nk])
    assert not parts.vel[parts.free].any() and not parts.stretch[parts.free].any()
    tol_v, tol_s = (1e-10, 1e-10) if precision == "fp64" else (1e-6, 2e-5)
    assert rel(parts.vel, v) <= tol_v
    assert rel(parts.stretch, s) <= tol_s
    g.destroy()


def test_tree_of_the_survivors_and_switching():
    cen, dir, mag, rv = cloud(30000, seed=5)
    elem = panels()
    print("redistribution: survivors, off/on switching", flush=True)
    p1, g1, nl1 = gpu_pass(cen, dir, mag, rv, elem=elem)
    idx = np.flatnonzero(~p1.free)
    p2, g2, nl2 = gpu_pass(cen[idx], p1.dir[idx], p1.mag[idx], rv[idx], pr=False)
    assert nl1 == nl2
    assert np.array_equal(p1.vel[idx], p2.vel) and np.array_equal(p1.stretch[idx], p2.stretch)
    g2.destroy()
    # the same handle: off (plain pass over all particles), on again
    plain = O.VortParts(cen, dir, mag, rv)
    g1.use_pr = False
    O.sort_particles(plain, g1, elem=elem)
    O.calculate_multipole(plain, g1)
    O.apply_multipole(plain, g1)
    assert not plain.free.any() and np.array_equal(plain.mag, mag)
    p3, g3, _ = gpu_pass(cen, dir, mag, rv, pr=False)
    assert np.array_equal(plain.vel, p3.vel) and np.array_equal(plain.stretch, p3.stretch)
    g3.destroy()
    again = O.VortParts(cen, dir, mag, rv)
    g1.use_pr = True
    O.sort_particles(again, g1, elem=elem)
    O.calculate_multipole(again, g1)
    O.apply_multipole(again, g1)
    assert np.array_equal(again.free, p1.free) and np.array_equal(again.vel, p1.vel)
    g1.destroy()


def test_every_cell_near_the_wall_keeps_everything():
    cen, dir, mag, rv = cloud(8000, seed=9)
    print("redistribution: lvl_solid 1, every cell near the wall", flush=True)
    p, g, _ = gpu_pass(cen, dir, mag, rv, elem=panels(), lvl=1)
    assert not p.free.any() and np.array_equal(p.mag, mag) and np.array_equal(p.dir, dir)
    g.destroy()


def test_batch_with_a_redistributing_case():
    import torch
    lib = L.lib()
    elem = panels()
    cases = [cloud(n, seed=30 + k) for k, n in enumerate((12000, 5000, 9000))]
    print("redistribution inside a batch of 3 FMM passes", flush=True)

    def make(cen, dir, mag, rv, pr):
        g = O.initialize_octree(TREE["box_length"], TREE["nbox"], TREE["origin"],
                                TREE["nlevels"], TREE["min_part"], 4, TREE["delta"],
                                capacity=len(cen), precision="tuned")
        dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (cen, dir, mag, rv)]
        L.check(lib.dust_fmm_set_particles_dev(g.handle, len(cen), *[t.data_ptr() for t in dev]),
                "set")
        ec = np.ascontiguousarray(elem)
        L.check(lib.dust_fmm_set_redistribution(g.handle, int(pr), 3, 3.0, len(ec), L.ptr(ec)),
                "pr")
        return g, dev

    def results(g, n):
        vel = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        keep = torch.empty(n, dtype=torch.uint8, device="cuda")
        L.check(lib.dust_fmm_results_dev(g.handle, vel.data_ptr(), None, None), "res")
        L.check(lib.dust_fmm_redistributed_dev(g.handle, keep.data_ptr(), None, None), "red")
        L.check(lib.dust_fmm_synchronize(g.handle), "sync")
        return vel.cpu().numpy(), keep.cpu().numpy()

    single = []
    for k, cs in enumerate(cases):
        g, dev = make(*cs, pr=(k != 1))
        L.check(lib.dust_fmm_pass_dev(g.handle, None), "pass")
        single.append(results(g, len(cs[0])))
        g.destroy()
    hs = [make(*cs, pr=(k != 1)) for k, cs in enumerate(cases)]
    arr = (C.c_void_p * len(hs))(*[g.handle.value for g, _ in hs])
    L.check(lib.dust_fmm_pass_batch_dev(len(hs), arr, None), "batch")
    for k, (g, _) in enumerate(hs):
        v, keep = results(g, len(cases[k][0]))
        assert np.array_equal(keep, single[k][1])
        assert (keep == 0).any() == (k != 1)
        assert rel(v, single[k][0]) <= 1e-12
        g.destroy()
