"""Project autotuner: expansion schedules of the FMM and launch knobs of the
LU solvers, chosen by measurement on the caller's own data.

FMM (tune_fmm): for every candidate schedule -- expansion order, tree depth,
leaf occupancy (min_part) -- the library's own knob search
(dust_fmm_autotune: P2P items per wavefront, M2L offsets per job) runs on a
handle of that schedule, the pass is timed, and its velocities are compared
with a reference pass of the FP64 form at a higher order; the fastest
schedule whose relative L2 velocity error is within `tol` wins.  Accuracy is
part of the choice, so the order is never lowered below what `tol` allows.

LU (tune_lu): the number of compute units kept free of the trailing updates
for the diagonal-block chain (dust_lu_set_tuning), timed on the caller's
matrix for the FP64 and the mixed solve.  Results of the solves do not
depend on it.

Inputs are device tensors (torch, cuda); nothing here runs without the HIP
library.  No counterpart in the reference (its octree parameters are user
input, mod_octree.f90:251-260).
"""
import ctypes as C
import time

import numpy as np

from . import _lib as L
from . import octree as O


def choose(table, tol):
    """The fastest entry of `table` (dicts with 'ms' and 'err') whose err is
    within tol; None when no entry meets it.  Ties go to the earlier entry."""
    best = None
    for row in table:
        if row["err"] <= tol and (best is None or row["ms"] < best["ms"]):
            best = row
    return best


def _pass(torch, tree, degree, nlevels, min_part, dev, n, form, reps, knob_search):
    lib = L.lib()
    g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], nlevels,
                            min_part, degree, tree["delta"], capacity=max(n, 1), use_vs=True,
                            precision=form, kernel=tree.get("kernel", "rosenhead"))
    try:
        h = g.handle
        L.check(lib.dust_fmm_set_particles_dev(h, n, dev["cen"].data_ptr(),
                                               dev["dir"].data_ptr(), dev["mag"].data_ptr(),
                                               dev["rv"].data_ptr()), "set_particles")
        knobs = {}
        if knob_search:
            g.autotune(reps=reps)
            knobs = {"seg_iters": g.get_tuning("seg_iters")}
            if form == "tuned":
                knobs["m2l_chunk"] = g.get_tuning("m2l_chunk")
        nl = C.c_int()
        L.check(lib.dust_fmm_pass_dev(h, C.byref(nl)), "pass")  # warm-up
        L.check(lib.dust_fmm_synchronize(h), "sync")
        t0 = time.perf_counter()
        for _ in range(reps):
            L.check(lib.dust_fmm_pass_dev(h, C.byref(nl)), "pass")
        L.check(lib.dust_fmm_synchronize(h), "sync")
        ms = (time.perf_counter() - t0) * 1e3 / reps
        vel = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        L.check(lib.dust_fmm_results_dev(h, vel.data_ptr(), None, None), "results")
        L.check(lib.dust_fmm_synchronize(h), "sync")
        return ms, vel, knobs, nl.value
    finally:
        g.destroy()


def tune_fmm(torch, dev, tree, tol=1e-6, form="tuned", degrees=(3, 4, 5, 6),
             levels=None, min_parts=(32, 64, 128), ref_degree=8, reps=2):
    """Searches order x depth x leaf occupancy for the particles in `dev`
    (device tensors cen (n,3), dir (n,3), mag (n), rv (n)) with the octree box
    of `tree` (box_length, nbox, origin, delta, nlevels; optional kernel).
    Returns dict(best=..., table=[...], ref_ms=...); best is None when no
    schedule meets tol."""
    n = int(dev["mag"].shape[0])
    L0 = int(tree["nlevels"])
    if levels is None:
        levels = tuple(l for l in (L0 - 1, L0, L0 + 1) if 1 <= l <= 10)
    ref_ms, ref, _, _ = _pass(torch, tree, ref_degree, L0, int(tree.get("min_part", 64)), dev,
                              n, "fp64", 1, False)
    nref = torch.linalg.norm(ref).item()
    table = []
    for p in degrees:
        for nlev in levels:
            for mp in min_parts:
                ms, vel, knobs, nleaves = _pass(torch, tree, p, nlev, mp, dev, n, form, reps, True)
                err = torch.linalg.norm(vel - ref).item() / nref
                table.append(dict(degree=p, nlevels=nlev, min_part=mp, ms=ms, err=err,
                                  leaves=nleaves, **knobs))
    return dict(best=choose(table, tol), table=table, ref_ms=ref_ms, tol=tol, form=form,
                ref=dict(degree=ref_degree, nlevels=L0))


def tune_lu(torch, A, nrhs=8, values=(0, 16, 32, 48, 64), reps=2, mixed_prec=None):
    """Times the FP64 getrf + getrs and the mixed solve (prec mixed_prec,
    default FP16S) on the device matrix A (n x n, column-major as torch
    row-major of A^T) for each number of reserved compute units; sets and
    returns the fastest of each."""
    lib = L.lib()
    n = int(A.shape[0])
    prec = L.LU_FP16S if mixed_prec is None else mixed_prec
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    X = torch.randn(nrhs, n, dtype=torch.float64, device="cuda")
    res = {"getrf_reserved_cus": [], "mixed_reserved_cus": []}
    for v in values:
        L.check(lib.dust_lu_set_tuning(b"getrf_reserved_cus", int(v)), "lu tuning")
        L.check(lib.dust_lu_set_tuning(b"mixed_reserved_cus", int(v)), "lu tuning")
        ts, tm = [], []
        for _ in range(reps + 1):
            Am, Xm = A.clone(), X.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = lib.dust_dgetrf_dev(n, n, Am.data_ptr(), n, ipiv.data_ptr(), None)
            L.check(lib.dust_dgetrs_dev(n, nrhs, Am.data_ptr(), n, ipiv.data_ptr(),
                                        Xm.data_ptr(), n, None), "getrs")
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            if info != 0:
                raise RuntimeError(f"dgetrf info {info}")
            Xm = X.clone()
            it, rs = C.c_int(), C.c_double()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L.check(lib.dust_dgesv_mixed_dev(n, nrhs, A.data_ptr(), n, Xm.data_ptr(), n, prec,
                                             1e-12, 50, C.byref(it), C.byref(rs), None), "mixed")
            torch.cuda.synchronize()
            tm.append(time.perf_counter() - t0)
        res["getrf_reserved_cus"].append(dict(value=int(v), ms=1e3 * float(np.median(ts[1:]))))
        res["mixed_reserved_cus"].append(dict(value=int(v), ms=1e3 * float(np.median(tm[1:]))))
    best = {}
    for k, rows in res.items():
        b = min(rows, key=lambda r: r["ms"])
        L.check(lib.dust_lu_set_tuning(k.encode(), b["value"]), "lu tuning")
        best[k] = b
    return dict(best=best, table=res)
