"""Benchmark of the hot path on one MI355X.

A step is one full FMM pass -- sort_particles + calculate_multipole +
apply_multipole (velocity and stretching) -- over the synthetic 3-helix
vortex wake of 1,000,000 particles, particles resident in HBM.  The metric is
particle updates per second of the tuned form (the plain FP64 form is timed
beside it).  The dense AIC solve (N = 16,384 panels of a perturbed ellipsoid,
8 right-hand sides) is timed after the FMM, plain FP64 LU vs the mixed
precision LU with FP64 refinement.

Prints ONE JSON line (see README of the driver contract).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3, "hbm": 8000.0}  # TFLOP/s, GB/s (MI355X)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fmm_handle(O, tree, degree, n, form):
    return O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                               tree["min_part"], degree, tree["delta"], capacity=n, use_vs=True,
                               precision=form)


def time_fmm(torch, L, O, tree, degree, dev, n, form, steps, warmup, knobs=None):
    """Returns dict with ms per pass, stage times and the work counts."""
    g = fmm_handle(O, tree, degree, n, form)
    for k, v in (knobs or {}).items():
        g.set_tuning(k, v)
    lib = L.lib()
    h = g.handle
    L.check(lib.dust_fmm_set_particles_dev(h, n, dev["cen"].data_ptr(), dev["dir"].data_ptr(),
                                           dev["mag"].data_ptr(), dev["rv"].data_ptr()), "set")
    import ctypes as C
    nl = C.c_int()
    for _ in range(warmup):
        L.check(lib.dust_fmm_pass_dev(h, C.byref(nl)), "pass")
    L.check(lib.dust_fmm_synchronize(h), "sync")
    stream = torch.cuda.ExternalStream(lib.dust_fmm_stream(h))
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        L.check(lib.dust_fmm_pass_dev(h, C.byref(nl)), "pass")
    e1.record(stream)
    L.check(lib.dust_fmm_synchronize(h), "sync")
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ev = e0.elapsed_time(e1) / steps
    # stage timing pass (HIP events around every stage on the handle's stream)
    g.set_timing(True)
    for _ in range(steps):
        L.check(lib.dust_fmm_pass_dev(h, C.byref(nl)), "pass")
    L.check(lib.dust_fmm_synchronize(h), "sync")
    stages = g.stage_times()
    g.set_timing(False)
    p2p_pairs, m2l_pairs = g.work()
    vel = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    st = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    L.check(lib.dust_fmm_results_dev(h, vel.data_ptr(), st.data_ptr(), None), "results")
    L.check(lib.dust_fmm_synchronize(h), "sync")
    res = dict(ms=wall, ms_events=ev, n_leaves=nl.value, stages={k: v[0] / steps for k, v in stages.items()},
               stage_launches={k: v[1] for k, v in stages.items()}, p2p_pairs=p2p_pairs,
               m2l_pairs=m2l_pairs, vel=vel.cpu().numpy(), stretch=st.cpu().numpy())
    g.destroy()
    return res


def time_lu(torch, L, LS, W, n, nrhs, steps):
    """AIC assembly + plain FP64 getrf/getrs and the mixed solve, device resident."""
    import ctypes as C
    lib = L.lib()
    ver, nv = W.ellipsoid(128, 128) if n == 16384 else W.ellipsoid(int(np.sqrt(n)), int(np.sqrt(n)))
    n = len(nv)
    dver = torch.from_numpy(np.ascontiguousarray(ver.reshape(n, 12))).cuda()
    dnv = torch.from_numpy(nv.astype(np.int32)).cuda()
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")  # column-major view: A.T rows
    B = torch.empty((n, n), dtype=torch.float64, device="cuda")
    prm = L.AicParams()
    prm.ff_ratio_doublet = 10.0
    prm.rankine_rad = 0.1
    prm.cutoff_rad = 0.001
    torch.cuda.synchronize()
    out = {}
    t0 = time.perf_counter()
    L.check(lib.dust_aic_surfpan_dev(n, dnv.data_ptr(), dver.data_ptr(), C.byref(prm),
                                     A.data_ptr(), n, B.data_ptr(), n, None), "aic")
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    out["assembly_ms"] = t_asm * 1e3
    out["assembly_GBps"] = 16.0 * n * n / t_asm / 1e9
    # right-hand sides b = B s (column-major B stored as B.T in torch)
    U = torch.from_numpy(W.freestreams(nrhs)).cuda()
    G = torch.from_numpy(ver.reshape(n, 4, 3)).cuda()
    # normals as in calc_geo_data
    v = G
    q = torch.where(dnv[:, None] == 4, torch.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 1], dim=1),
                    torch.cross(v[:, 2] - v[:, 1], v[:, 0] - v[:, 1], dim=1))
    nor = q / q.norm(dim=1, keepdim=True)
    S = -(nor @ U.T)  # (n, nrhs)
    Bm = B.T  # logical column-major matrix
    rhs = (Bm @ S).contiguous()  # (n, nrhs) row-major -> need column-major
    rhs_cm = rhs.T.contiguous()  # (nrhs, n) = column-major (n, nrhs)
    Am = A.clone()
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    # plain FP64
    ts = []
    for it in range(steps):
        Am.copy_(A)
        X = rhs_cm.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = lib.dust_dgetrf_dev(n, n, Am.data_ptr(), n, ipiv.data_ptr(), None)
        L.check(lib.dust_dgetrs_dev(n, nrhs, Am.data_ptr(), n, ipiv.data_ptr(), X.data_ptr(), n,
                                    None), "getrs")
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        assert info == 0, info
    out["fp64_ms"] = 1e3 * float(np.median(ts))
    Xp = X.T  # (n, nrhs)
    r = (A.T @ Xp - rhs)
    out["fp64_residual"] = float((r.abs().max() / (A.abs().sum(dim=0).max() * Xp.abs().max())).item())
    out["getrf_TFLOPs"] = (2.0 / 3.0) * n ** 3 / (out["fp64_ms"] / 1e3) / 1e12
    # mixed
    for prec, code in (("fp16s", L.LU_FP16S), ("fp32", L.LU_FP32)):
        ts = []
        ok = True
        for it in range(steps):
            X2 = rhs_cm.clone()
            iters = C.c_int()
            resid = C.c_double()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc = lib.dust_dgesv_mixed_dev(n, nrhs, A.data_ptr(), n, X2.data_ptr(), n, code, 1e-12,
                                          50, C.byref(iters), C.byref(resid), None)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            if rc != 0:
                ok = False
                break
        if ok:
            Xm = X2.T
            out[f"{prec}_ms"] = 1e3 * float(np.median(ts))
            out[f"{prec}_iters"] = iters.value
            out[f"{prec}_backward_error"] = resid.value
            out[f"{prec}_rel_diff_vs_fp64"] = float(((Xm - Xp).norm() / Xp.norm()).item())
            out[f"{prec}_speedup"] = out["fp64_ms"] / out[f"{prec}_ms"]
        else:
            out[f"{prec}_error"] = L.lib().dust_last_error().decode()
    out["n"] = n
    out["nrhs"] = nrhs
    return out


def cpu_baseline(W, n_sample, degree):
    """The CPU oracle (FP64 restatement, OpenMP) on a bounded sample of the same
    generator; particle updates per second."""
    from oracle import oracle as orc
    cen, alpha, rv, tree = W.helix_wake(n_sample)
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    lev = tree["nlevels"] - max(0, int(round(np.log(1_000_000 / n_sample) / np.log(8))))
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], lev, tree["min_part"], degree,
                   tree["delta"])
    t0 = time.perf_counter()
    c.run(cen, dr, mag, rv)
    dt = time.perf_counter() - t0
    return dict(value=n_sample / dt, unit="particle-updates/s", cores=orc.num_threads(),
                kind="port",
                sample=f"oracle FMM pass (sort+multipole+apply, FP64, OpenMP) on a "
                       f"{n_sample}-particle draw of the same helix generator, {lev} levels, "
                       f"order {degree}, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--degree", type=int, default=4)
    ap.add_argument("--lu-n", type=int, default=16384)
    ap.add_argument("--no-lu", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=125_000)
    ap.add_argument("--near", type=float, default=1.0,
                    help="tuned form: FP64 for pairs closer than this many core radii")
    ap.add_argument("--no-fp64", action="store_true", help="skip timing the plain FP64 form")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this path has no "
              "multi-GPU form", file=sys.stderr)
        sys.exit(2)
    import torch

    from dust_amd import _lib as L
    from dust_amd import linsys as LS
    from dust_amd import octree as O
    from dust_amd import workloads as W

    L.lib()
    n = args.n
    log(f"generating helix wake n={n}")
    cen, alpha, rv, tree = W.helix_wake(n)
    mag = np.linalg.norm(alpha, axis=1)
    dr = alpha / mag[:, None]
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda()
           for k, v in dict(cen=cen, dir=dr, mag=mag, rv=rv).items()}
    torch.cuda.synchronize()
    log("timing plain FP64 FMM")
    plain = time_fmm(torch, L, O, tree, args.degree, dev, n, "fp64", args.steps, args.warmup)
    log(f"plain: {plain['ms']:.2f} ms/pass stages {plain['stages']}")
    tuned = None
    try:
        log("timing tuned FMM")
        tuned = time_fmm(torch, L, O, tree, args.degree, dev, n, "tuned", args.steps, args.warmup,
                         knobs={"near_factor": args.near})
        log(f"tuned: {tuned['ms']:.2f} ms/pass stages {tuned['stages']}")
    except Exception as e:  # tuned form not built in this library
        log(f"tuned form unavailable: {e}")
    head = tuned if tuned is not None else plain
    form = "tuned" if tuned is not None else "fp64"
    value = n / (head["ms"] / 1e3)
    # roofline of the dominant kernel (largest stage)
    st = head["stages"]
    dom = max(st, key=lambda k: st[k])
    if dom == "p2p":
        flops = 60.0 * head["p2p_pairs"]
        peak = PEAK["fp32"] if form == "tuned" else PEAK["fp64"]
    elif dom == "m2l":
        nm = (args.degree + 1) * (args.degree + 2) * (args.degree + 3) // 6
        flops = 72.0 * nm * nm * head["m2l_pairs"]
        peak = PEAK["fp32"] if form == "tuned" else PEAK["fp64"]
    else:
        flops = None
    roof = None
    if flops is not None:
        ach = flops / (st[dom] / 1e3) / 1e12
        roof = dict(bound="valu" if dom == "p2p" else "mfma", kernel=dom, achieved=ach, peak=peak,
                    unit="TFLOP/s", frac=ach / peak, traffic=None,
                    algorithmic=f"{'60 flop/pair x %.4g pairs' % head['p2p_pairs'] if dom == 'p2p' else '72 n_mon^2 flop x %.4g M2L pairs' % head['m2l_pairs']}")
    acc = None
    if tuned is not None:
        dv = np.linalg.norm(tuned["vel"] - plain["vel"]) / np.linalg.norm(plain["vel"])
        ds = np.linalg.norm(tuned["stretch"] - plain["stretch"]) / np.linalg.norm(plain["stretch"])
        acc = dict(vel_rel_l2_vs_fp64=dv, stretch_rel_l2_vs_fp64=ds)
    lu = None
    if not args.no_lu:
        log(f"timing AIC + LU n={args.lu_n}")
        try:
            lu = time_lu(torch, L, LS, W, args.lu_n, 8, 3)
            log(f"lu: {lu}")
        except Exception as e:
            log(f"LU timing failed: {e}")
            lu = dict(error=str(e))
    cpu = None
    if not args.no_cpu:
        log("timing CPU oracle sample")
        cpu = cpu_baseline(W, args.cpu_sample, args.degree)
        log(f"cpu: {cpu}")
    out = {
        "metric": "wake particle updates/s (FMM velocity + stretching)",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if form == "tuned" else "fp64",
        "data": "synthetic",
        "config": {"workload": f"3-helix vortex wake, {n} particles, Rosenhead core 0.02, "
                               f"octree {tree['nlevels']} levels, min_part {tree['min_part']}, "
                               f"expansion order {args.degree}, vel+stretch",
                   "form": form, "parallelism": "single GPU"},
        "fp64_path": {"value": n / (plain["ms"] / 1e3), "ms_per_step": plain["ms"],
                      "stages_ms": plain["stages"], "n_leaves": plain["n_leaves"]},
        "tuned_path": None if tuned is None else {"ms_per_step": tuned["ms"],
                                                  "stages_ms": tuned["stages"],
                                                  "speedup_vs_fp64": plain["ms"] / tuned["ms"]},
        "accuracy": acc,
        "work": {"p2p_pairs": head["p2p_pairs"], "m2l_pairs": head["m2l_pairs"]},
        "roofline": roof,
        "lu": lu,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out, default=float))


if __name__ == "__main__":
    main()
