"""Direct element <-> particle terms on the GPU (dust_elem_apply,
dust_vort_vel) against the FP64 CPU oracle: the element loops of
apply_multipole (src/fmm/mod_octree.f90:1350-1401) and get_vort_vel
(src/aeroel/mod_surfpan.f90:962-977), to 1e-10 relative L2, with points in
the near field, inside the Rankine cores and on the far-field switch."""
import ctypes as C

import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import elements as E
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_groups(groups):
    return [dict(g, ver=np.asarray(g["ver"])) for g in groups]


@pytest.fixture(scope="module")
def wing():
    groups = W.wing_elements()
    ver, nv, cen, alpha, rv, tree = W.wing_case()
    rng = np.random.default_rng(4)
    body = groups[0]
    pc = np.array([body["ver"][i].mean(axis=0) for i in range(len(body["ver"]))])
    # near-field probes: just off the panel centres and edges, and some on
    # the wake panels and end vortices
    near = pc[rng.choice(len(pc), 300, replace=False)] + 0.01 * rng.standard_normal((300, 3))
    edge = body["ver"][rng.choice(len(pc), 100, replace=False), 0] + 1e-3 * rng.standard_normal(
        (100, 3))
    wake = groups[3]["ver"][:, 0] + [0.0, 0.02, 0.001]
    pts = np.concatenate([cen, near, edge, wake])
    al = np.concatenate([alpha, 1e-3 * rng.standard_normal((len(pts) - len(cen), 3))])
    return groups, pts, al


@pytest.mark.parametrize("vs_elems", [False, True], ids=["vel", "vel+stretch"])
def test_elem_apply_wing_vs_oracle(orc, wing, vs_elems):
    groups, pts, al = wing
    rng = np.random.default_rng(1)
    v0 = rng.standard_normal(pts.shape)
    s0 = rng.standard_normal(pts.shape) if vs_elems else None
    prm = dict(ff_ratio_doublet=10.0, rankine_rad=0.1, cutoff_rad=0.001, vortex_rad=0.1)
    print(f"elem_apply wing n={len(pts)} vs_elems={vs_elems}", flush=True)
    vel, st = E.apply_elements(groups, pts, alpha=al if vs_elems else None, vel=v0, stretch=s0,
                               **prm)
    ov, os_ = orc.elem_apply(oracle_groups(groups), pts, alpha=al if vs_elems else None, vel=v0,
                             stretch=s0, ff_ratio=10.0, rankine=0.1, cutoff=0.001,
                             vortex_rad=0.1)
    print(f"vel rel {rel(vel - v0, ov - v0):.2e}")
    assert rel(vel - v0, ov - v0) <= TOL
    assert np.all(np.isfinite(vel))
    if vs_elems:
        print(f"stretch rel {rel(st - s0, os_ - s0):.2e}")
        assert rel(st - s0, os_ - s0) <= TOL


def test_elem_apply_kinds_and_missing_strengths(orc, wing):
    """Each kind alone; a surface-panel list with sources only (mag None);
    lines with no strength associated; a smaller far-field ratio."""
    groups, pts, al = wing
    pts = pts[-600:]
    al = al[-600:]
    cases = [[groups[0]], [groups[1]], [groups[3]], [dict(groups[0], mag=None)],
             [dict(groups[3], mag=None)], [dict(groups[0], sigma=None), groups[2]]]
    for ff in (10.0, 4.0):
        for gs in cases:
            print(f"elem_apply kinds={[g['kind'] for g in gs]} ff={ff}", flush=True)
            vel, st = E.apply_elements(gs, pts, alpha=al, ff_ratio_doublet=ff, vortex_rad=0.05)
            ov, os_ = orc.elem_apply(oracle_groups(gs), pts, alpha=al, ff_ratio=ff,
                                     vortex_rad=0.05)
            if np.linalg.norm(ov) == 0.0:
                assert np.all(vel == 0.0) and np.all(st == 0.0)
                continue
            assert rel(vel, ov) <= TOL
            if np.linalg.norm(os_) > 0.0:
                assert rel(st, os_) <= TOL
            else:
                assert np.all(st == 0.0)


@pytest.mark.parametrize("n", [900, 9000], ids=["input-order", "morton"])
def test_elem_apply_triangles_vs_oracle(orc, wing, n):
    """Triangular surface panels (each wing quad cut along a diagonal, both
    strengths) and a mixed list with the quads: velocity and stretching with
    probes in the near field and on the vertices' neighbourhood."""
    groups, pts, al = wing
    body = groups[0]
    q = np.asarray(body["ver"])
    tri = np.zeros((2 * len(q), 4, 3))
    tri[0::2, :3] = q[:, [0, 1, 2]]
    tri[1::2, :3] = q[:, [0, 2, 3]]
    rng = np.random.default_rng(6)
    tg = dict(kind=E.SURFPAN, ver=tri, n_ver=np.full(len(tri), 3, np.int32),
              mag=np.repeat(body["mag"], 2), sigma=np.repeat(body["sigma"], 2))
    cen = tri[:, :3].mean(axis=1)
    probes = cen[rng.choice(len(cen), n, replace=True)] + 0.02 * rng.standard_normal((n, 3))
    a = 1e-3 * rng.standard_normal((n, 3))
    for gs in ([tg], [tg, groups[1]]):
        print(f"elem_apply triangles n={n} groups={len(gs)}", flush=True)
        vel, st = E.apply_elements(gs, probes, alpha=a, vortex_rad=0.05)
        ov, os_ = orc.elem_apply(oracle_groups(gs), probes, alpha=a, vortex_rad=0.05)
        print(f"vel rel {rel(vel, ov):.2e} stretch rel {rel(st, os_):.2e}")
        assert rel(vel, ov) <= TOL and rel(st, os_) <= TOL


def test_vort_vel_wing_vs_oracle(orc, wing):
    groups, pts, al = wing
    ver, nv, cen, alpha, rv, tree = W.wing_case()
    body = groups[0]
    pc = np.array([body["ver"][i].mean(axis=0) for i in range(len(body["ver"]))])
    mag = np.linalg.norm(alpha, axis=1)
    dirs = alpha / mag[:, None]
    lv, lm = groups[3]["ver"], groups[3]["mag"]
    print(f"vort_vel m={len(pc)} np={len(cen)} nl={len(lv)}", flush=True)
    u = E.get_vort_vel(pc, cen, dirs, mag, rv, lv, lm)
    ou = orc.vort_vel(pc, cen, dirs, mag, rv, lv, lm)
    print(f"uvort rel {rel(u, ou):.2e}")
    assert rel(u, ou) <= TOL
    # lines only, particles only, no strength on the lines
    assert rel(E.get_vort_vel(pc, lver=lv, lmag=lm), orc.vort_vel(pc, lver=lv, lmag=lm)) <= TOL
    assert rel(E.get_vort_vel(pc, cen, dirs, mag, rv), orc.vort_vel(pc, cen, dirs, mag, rv)) <= TOL
    assert np.all(E.get_vort_vel(pc, lver=lv, lmag=None) == 0.0)


@pytest.mark.parametrize("m,npart", [(1, 1), (300, 70_001), (1601, 200_000)])
def test_vort_vel_split_sources(orc, m, npart):
    """Few targets, many sources: the source splits and the ragged last
    chunk (sizes that are not multiples of 256)."""
    cen, alpha, rv, tree = W.helix_wake(npart, seed=5)
    rng = np.random.default_rng(m)
    tg = rng.uniform(-2.0, 2.0, (m, 3))
    mag = np.linalg.norm(alpha, axis=1)
    dirs = alpha / mag[:, None]
    print(f"vort_vel split m={m} np={npart}", flush=True)
    u = E.get_vort_vel(tg, cen, dirs, mag, rv)
    ou = orc.vort_vel(tg, cen, dirs, mag, rv)
    assert rel(u, ou) <= TOL


def test_device_entry_points_match_host(wing):
    import torch
    groups, pts, al = wing
    n = len(pts)
    dev = torch.device("cuda")
    keep = []
    arr = (L.ElemGroup * len(groups))()
    for k, g in enumerate(groups):
        kind = g["kind"]
        v = np.zeros((len(g["ver"]), 2 if kind == 2 else 4, 3))
        v[:] = np.asarray(g["ver"])[:, : v.shape[1]]
        tv = torch.from_numpy(v).to(dev)
        tn = torch.from_numpy(np.asarray(g["n_ver"], np.int32)).to(dev)
        tm = torch.from_numpy(np.asarray(g["mag"], np.float64)).to(dev)
        ts = (torch.from_numpy(np.asarray(g["sigma"], np.float64)).to(dev)
              if g.get("sigma") is not None else None)
        keep += [tv, tn, tm, ts]
        arr[k].kind, arr[k].n = kind, len(v)
        arr[k].n_ver, arr[k].ver, arr[k].mag = tn.data_ptr(), tv.data_ptr(), tm.data_ptr()
        arr[k].sigma = ts.data_ptr() if ts is not None else None
        arr[k].stretch = 1 if g.get("stretch", True) else 0
    tp = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    ta = torch.from_numpy(np.ascontiguousarray(al)).to(dev)
    tvel = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    tst = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    prm = E.params()
    print(f"elem_apply_dev n={n}", flush=True)
    L.check(L.lib().dust_elem_apply_dev(len(groups), arr, C.byref(prm), n, tp.data_ptr(),
                                        ta.data_ptr(), tvel.data_ptr(), tst.data_ptr(), None),
            "elem_apply_dev")
    torch.cuda.synchronize()
    vel, st = E.apply_elements(groups, pts, alpha=al)
    assert np.array_equal(tvel.cpu().numpy(), vel)
    assert np.array_equal(tst.cpu().numpy(), st)
    # uvort through the device entry point
    body = groups[0]
    pc = np.ascontiguousarray([body["ver"][i].mean(axis=0) for i in range(len(body["ver"]))])
    mag = np.linalg.norm(al, axis=1) + 1e-3
    dirs = np.ascontiguousarray(al / mag[:, None])
    rv = np.full(n, 0.1)
    t = [torch.from_numpy(x).to(dev) for x in (pc, pts, dirs, mag, rv)]
    tu = torch.zeros((len(pc), 3), dtype=torch.float64, device=dev)
    L.check(L.lib().dust_vort_vel_dev(len(pc), t[0].data_ptr(), n, t[1].data_ptr(),
                                      t[2].data_ptr(), t[3].data_ptr(), t[4].data_ptr(), 0, None,
                                      None, C.byref(prm), tu.data_ptr(), None), "vort_vel_dev")
    torch.cuda.synchronize()
    assert np.array_equal(tu.cpu().numpy(), E.get_vort_vel(pc, pts, dirs, mag, rv))


def test_empty_inputs():
    groups = W.wing_elements()
    vel, _ = E.apply_elements(groups, np.zeros((0, 3)))
    assert vel.shape == (0, 3)
    v0 = np.ones((5, 3))
    vel, _ = E.apply_elements([], np.zeros((5, 3)) + 3.0, vel=v0)
    assert np.array_equal(vel, v0)
    assert E.get_vort_vel(np.zeros((0, 3))).shape == (0, 3)
    assert np.all(E.get_vort_vel(np.ones((4, 3))) == 0.0)


def test_morton_order_path_matches_input_order(wing):
    """Targets above the sorting threshold are evaluated in Morton order and
    scattered back; the two paths sum each target's elements in list order
    (split into element ranges when there are few targets, a split that
    depends on the target count), so they agree to rounding."""
    groups, pts, al = wing
    big = np.concatenate([pts, pts[:3000] + 0.37])          # sorted path
    print(f"elem_apply sorted n={len(big)} vs input order n=4000", flush=True)
    vb, sb = E.apply_elements(groups, big, alpha=np.concatenate([al, al[:3000]]))
    vs, ss = E.apply_elements(groups, big[:4000], alpha=al[:4000])   # below the threshold
    assert np.allclose(vb[:4000], vs, rtol=1e-13, atol=1e-13 * np.abs(vs).max())
    assert np.allclose(sb[:4000], ss, rtol=1e-13, atol=1e-13 * np.abs(ss).max())


def test_tuned_far_field_vs_oracle(orc, wing):
    """precision = tuned: far-field element velocities in FP32 (FP64 offsets,
    switches and sums), near field with FP64 geometry and sums and FP32
    logarithms and arctangents: within 1e-6 of the oracle, also when every
    pair is near (the near field alone against the oracle and the FP64 form)."""
    groups, pts, al = wing
    print(f"elem_apply tuned n={len(pts)}", flush=True)
    vel, _ = E.apply_elements(groups, pts, precision="tuned")
    ov, _ = orc.elem_apply(oracle_groups(groups), pts)
    print(f"tuned vel rel {rel(vel, ov):.2e}")
    assert rel(vel, ov) <= 1e-6
    vn, _ = E.apply_elements(groups, pts, precision="tuned", ff_ratio_doublet=1e9)
    v64, _ = E.apply_elements(groups, pts, ff_ratio_doublet=1e9)
    on, _ = orc.elem_apply(oracle_groups(groups), pts, ff_ratio=1e9)
    print(f"near only: tuned vs fp64 {rel(vn, v64):.2e}, fp64 vs oracle {rel(v64, on):.2e}")
    assert rel(v64, on) <= TOL
    assert 0 < rel(vn, v64) <= 1e-6
