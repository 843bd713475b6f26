"""Batched entry points: several independent cases in one call give the same
results as the cases run one by one (FMM passes) and as LAPACK (mixed
solves)."""
import ctypes as C

import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import linsys as LS
from dust_amd import octree as O
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _device_case(torch, cen, alpha, rv, tree, precision):
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                            tree["min_part"], 4, tree["delta"], capacity=len(cen), use_vs=True,
                            precision=precision)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (cen, dr, mag, rv)]
    L.check(L.lib().dust_fmm_set_particles_dev(g.handle, len(cen), *[d.data_ptr() for d in dev]),
            "set")
    return g, dev


def _results(torch, g, n):
    vel = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    st = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    L.check(L.lib().dust_fmm_results_dev(g.handle, vel.data_ptr(), st.data_ptr(), None), "res")
    L.check(L.lib().dust_fmm_synchronize(g.handle), "sync")
    return vel.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp64", "tuned"])
def test_fmm_batch_equals_single(orc, precision):
    """The passes of a batch equal the passes one by one; one case of the
    batch against the oracle on the same particles (the tuned passes run
    their M2L and P2P in the batch's two shared persistent launches)."""
    import torch
    lib = L.lib()
    cases = []
    for k, n in enumerate((9000, 3000, 15000, 500, 7000)):
        cen, alpha, rv, tree = W.helix_wake(n, seed=100 + k, pitch=0.4 + 0.05 * k)
        tree = dict(tree, nlevels=4)
        cases.append((cen, alpha, rv, tree))
    print(f"batch of {len(cases)} FMM cases ({precision})", flush=True)
    single = []
    for cen, alpha, rv, tree in cases:
        g, dev = _device_case(torch, cen, alpha, rv, tree, precision)
        nl = C.c_int()
        L.check(lib.dust_fmm_pass_dev(g.handle, C.byref(nl)), "pass")
        single.append((nl.value, *_results(torch, g, len(cen))))
        g.destroy()
    handles, keep = [], []
    for cen, alpha, rv, tree in cases:
        g, dev = _device_case(torch, cen, alpha, rv, tree, precision)
        handles.append(g)
        keep.append(dev)
    arr = (C.c_void_p * len(handles))(*[g.handle.value for g in handles])
    nls = (C.c_int * len(handles))()
    L.check(lib.dust_fmm_pass_batch_dev(len(handles), arr, nls), "batch")
    for i, g in enumerate(handles):
        v, s = _results(torch, g, len(cases[i][0]))
        assert nls[i] == single[i][0]
        if precision == "fp64":
            # the FP64 form sums in a fixed order: bit-exact
            assert np.array_equal(v, single[i][1])
            assert np.array_equal(s, single[i][2])
        else:
            # the tuned M2L adds into the locals with atomics (order varies)
            assert rel(v, single[i][1]) <= 1e-12
            assert rel(s, single[i][2]) <= 1e-12
        if i == 2:
            cen, alpha, rv, tree = cases[i]
            mag = np.linalg.norm(alpha, axis=1)
            o = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                           tree["min_part"], 4, tree["delta"])
            vo, so, _ = o.run(cen, alpha / mag[:, None], mag, rv)
            tol_v, tol_s = (1e-10, 1e-10) if precision == "fp64" else (1e-6, 2e-5)
            assert rel(v, vo) <= tol_v
            assert rel(s, so) <= tol_s
        g.destroy()


def test_mixed_batch_vs_lapack():
    import torch
    lib = L.lib()
    systems = []
    for k, (b, p) in enumerate(((2, 600), (3, 1300), (5, 900), (4, 2100))):
        ver, nv = W.rotor(b, p)
        A, Bs = LS.aic_surfpan(ver, nv)
        rhs = LS.rhs_static(ver, nv, Bs, W.freestreams(8, seed=k))
        systems.append((A, rhs))
    dA = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A, _ in systems]
    dB = [torch.from_numpy(np.ascontiguousarray(b.T)).cuda() for _, b in systems]
    nb = len(systems)
    n = (C.c_int * nb)(*[A.shape[0] for A, _ in systems])
    pa = (C.c_void_p * nb)(*[d.data_ptr() for d in dA])
    pb = (C.c_void_p * nb)(*[d.data_ptr() for d in dB])
    it = (C.c_int * nb)()
    res = (C.c_double * nb)()
    print(f"batch of {nb} mixed solves, n = {list(n)}", flush=True)
    L.check(lib.dust_dgesv_mixed_batch_dev(nb, n, pa, n, pb, n, 8, L.LU_FP16S, 1e-12, 40, it, res),
            "mixed batch")
    for i, (A, rhs) in enumerate(systems):
        X = dB[i].cpu().numpy().T
        assert it[i] >= 1 and res[i] <= 1e-12
        assert rel(X, np.linalg.solve(A, rhs)) <= 1e-10


@pytest.mark.parametrize("prec", ["fp32", "fp16s"])
def test_mixed_batch_equals_single(prec):
    """Uneven batch: empty, tiny, non-multiple-of-128 and multi-chunk systems,
    a rotor AIC with multipliers above 2 and a system that needs the FP64
    fallback.  The batched factorisation runs the same tile operations in the
    same order as the single solve, so solutions, sweep counts and backward
    errors agree exactly."""
    import torch
    lib = L.lib()
    rng = np.random.default_rng(31)
    nrhs = 11
    systems = []
    for n in (0, 1, 127, 129, 300, 512, 1000):
        A = rng.standard_normal((n, n)) / np.sqrt(max(n, 1)) + 2.0 * np.eye(n)
        if n == 512:
            A = np.roll(A, 200, axis=0)  # pivots outside the diagonal blocks
        systems.append(A)
    ver, nv = W.rotor(5, 704)
    systems.insert(3, LS.aic_surfpan(ver, nv)[0])
    rhs = [rng.standard_normal((A.shape[0], nrhs)) for A in systems]
    code = {"fp32": L.LU_FP32, "fp16s": L.LU_FP16S}[prec]
    single = []
    for A, b in zip(systems, rhs):
        if A.shape[0] == 0:
            single.append((b, 0, 0.0))
            continue
        single.append(LS.dgesv_mixed(np.asfortranarray(A), np.asfortranarray(b.copy()), prec=prec,
                                     tol=1e-12, max_iter=40))
    nb = len(systems)
    dA = [torch.from_numpy(np.ascontiguousarray(A.T)).cuda() for A in systems]
    dB = [torch.from_numpy(np.ascontiguousarray(b.T)).cuda() for b in rhs]
    n = (C.c_int * nb)(*[A.shape[0] for A in systems])
    ld = (C.c_int * nb)(*[max(1, A.shape[0]) for A in systems])
    pa = (C.c_void_p * nb)(*[d.data_ptr() for d in dA])
    pb = (C.c_void_p * nb)(*[d.data_ptr() for d in dB])
    it = (C.c_int * nb)()
    res = (C.c_double * nb)()
    print(f"batch of {nb} mixed solves ({prec}), n = {list(n)}, nrhs {nrhs}", flush=True)
    L.check(lib.dust_dgesv_mixed_batch_dev(nb, n, pa, ld, pb, ld, nrhs, code, 1e-12, 40, it, res),
            "mixed batch")
    for i, A in enumerate(systems):
        X = dB[i].cpu().numpy().T
        Xs, its, bes = single[i]
        print(f"  n={A.shape[0]}: iters {it[i]} / {its}, berr {res[i]:.2e}")
        assert it[i] == its
        assert res[i] == bes
        assert np.array_equal(X, Xs)
        if A.shape[0] == 512:
            assert it[i] == -1
        elif A.shape[0] > 0:
            assert it[i] >= 1
