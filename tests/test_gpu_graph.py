"""Graph form of the FMM pass: every count on the device, captured once into a
HIP graph and replayed for wakes of other sizes and trees.  Replays against
the host-ordered pass (dust_fmm_pass_dev) on the same particles: bit for bit
in the FP64 form, to the tuned M2L's atomic order (1e-12) in the tuned form;
capacity overflow and out-of-box particles reported after the replay."""
import ctypes as C

import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import octree as O
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu

E_OUTSIDE, E_CAPACITY = -3, -4


def _tree_handle(tree, precision, cap):
    return O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                               tree["min_part"], 4, tree["delta"], capacity=cap, use_vs=True,
                               precision=precision)


class Bufs:
    """Fixed device buffers the captured graph reads and writes."""

    def __init__(self, torch, cap):
        f64 = dict(dtype=torch.float64, device="cuda")
        self.n = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.cen = torch.zeros((cap, 3), **f64)
        self.dir = torch.zeros((cap, 3), **f64)
        self.mag = torch.zeros(cap, **f64)
        self.rv = torch.zeros(cap, **f64)
        self.vel = torch.zeros((cap, 3), **f64)
        self.st = torch.zeros((cap, 3), **f64)

    def load(self, torch, cen, alpha, rv):
        n = len(cen)
        mag = np.linalg.norm(alpha, axis=1)
        self.cen[:n] = torch.from_numpy(cen)
        self.dir[:n] = torch.from_numpy(alpha / mag[:, None])
        self.mag[:n] = torch.from_numpy(mag)
        self.rv[:n] = torch.from_numpy(rv)
        self.n.fill_(n)
        torch.cuda.synchronize()

    def args(self):
        return [t.data_ptr() for t in (self.n, self.cen, self.dir, self.mag, self.rv, self.vel,
                                       self.st)]


def _eager(torch, g, b, n):
    """Host-ordered pass over the first n particles of the buffers."""
    lib = L.lib()
    L.check(lib.dust_fmm_set_particles_dev(g.handle, n, b.cen.data_ptr(), b.dir.data_ptr(),
                                           b.mag.data_ptr(), b.rv.data_ptr()), "set")
    nl = C.c_int()
    L.check(lib.dust_fmm_pass_dev(g.handle, C.byref(nl)), "pass")
    vel = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    st = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    L.check(lib.dust_fmm_results_dev(g.handle, vel.data_ptr(), st.data_ptr(), None), "res")
    L.check(lib.dust_fmm_synchronize(g.handle), "sync")
    return nl.value, vel.cpu().numpy(), st.cpu().numpy()


def _draws(cap):
    out = []
    for k, n in enumerate((20000, 16500, 23000)):
        cen, alpha, rv, tree = W.helix_wake(n, seed=300 + k, pitch=0.45 + 0.05 * k)
        out.append((cen, alpha, rv, dict(tree, nlevels=5)))
    assert all(len(d[0]) <= cap for d in out)
    return out


@pytest.mark.parametrize("precision", ["fp64", "tuned"])
def test_graph_replays_match_host_ordered_pass(precision):
    import torch
    lib = L.lib()
    cap = 24000
    draws = _draws(cap)
    tree = draws[0][3]
    ref = _tree_handle(tree, precision, cap)
    g = _tree_handle(tree, precision, cap)
    b = Bufs(torch, cap)
    # capacities from one host-ordered pass on the first draw
    b.load(torch, *draws[0][:3])
    _eager(torch, g, b, len(draws[0][0]))
    L.check(lib.dust_fmm_graph_reserve(g.handle, 2.5), "reserve")
    print(f"graph form ({precision}): capture", flush=True)
    L.check(lib.dust_fmm_graph_capture(g.handle, *b.args()), "capture")
    counts = (C.c_int * 4)()
    for rep in range(2):
        for k, (cen, alpha, rv, _) in enumerate(draws):
            n = len(cen)
            print(f"graph form ({precision}): replay {rep}, draw {k}, n={n}", flush=True)
            b.load(torch, cen, alpha, rv)
            b.vel.zero_()
            b.st.zero_()
            torch.cuda.synchronize()
            L.check(lib.dust_fmm_graph_launch(g.handle), "launch")
            L.check(lib.dust_fmm_graph_status(g.handle, counts), "status")
            vel = b.vel[:n].cpu().numpy()
            st = b.st[:n].cpu().numpy()
            nl, rvel, rst = _eager(torch, ref, b, n)
            assert counts[0] == n and counts[2] == nl
            if precision == "fp64":
                assert np.array_equal(vel, rvel) and np.array_equal(st, rst)
            else:
                assert np.linalg.norm(vel - rvel) / np.linalg.norm(rvel) < 1e-12
                assert np.linalg.norm(st - rst) / np.linalg.norm(rst) < 1e-12
            assert not b.vel[n:].any()  # nothing written past the count
    # the same sequence enqueued without a capture
    cen, alpha, rv, _ = draws[1]
    n = len(cen)
    b.load(torch, cen, alpha, rv)
    L.check(lib.dust_fmm_pass_graph_dev(g.handle, *b.args()), "graph pass")
    L.check(lib.dust_fmm_graph_status(g.handle, counts), "status")
    nl, rvel, rst = _eager(torch, ref, b, n)
    vel = b.vel[:n].cpu().numpy()
    if precision == "fp64":
        assert np.array_equal(vel, rvel)
    else:
        assert np.linalg.norm(vel - rvel) / np.linalg.norm(rvel) < 1e-12
    ref.destroy()
    g.destroy()


def test_graph_capacity_and_outside_reported():
    import torch
    lib = L.lib()
    cap = 24000
    small = W.helix_wake(3000, seed=401)
    tree = dict(small[3], nlevels=5)
    g = _tree_handle(tree, "tuned", cap)
    b = Bufs(torch, cap)
    b.load(torch, *small[:3])
    _eager(torch, g, b, 3000)
    L.check(lib.dust_fmm_graph_reserve(g.handle, 1.0), "reserve")
    L.check(lib.dust_fmm_graph_capture(g.handle, *b.args()), "capture")
    L.check(lib.dust_fmm_graph_launch(g.handle), "launch")
    L.check(lib.dust_fmm_graph_status(g.handle, None), "status")
    # a wake with many more leaves than reserved
    big = W.helix_wake(22000, seed=402)
    print("graph form: replay over capacity", flush=True)
    b.load(torch, *big[:3])
    L.check(lib.dust_fmm_graph_launch(g.handle), "launch")
    assert lib.dust_fmm_graph_status(g.handle, None) == E_CAPACITY
    # a particle count beyond the handle
    b.n.fill_(cap + 1)
    torch.cuda.synchronize()
    L.check(lib.dust_fmm_graph_launch(g.handle), "launch")
    assert lib.dust_fmm_graph_status(g.handle, None) == E_CAPACITY
    # a particle outside the box
    cen, alpha, rv = (a.copy() for a in small[:3])
    cen[17] = [50.0, 0.0, 0.0]
    b.load(torch, cen, alpha, rv)
    L.check(lib.dust_fmm_graph_launch(g.handle), "launch")
    assert lib.dust_fmm_graph_status(g.handle, None) == E_OUTSIDE
    # the handle still runs host-ordered passes, and the graph again
    b.load(torch, *small[:3])
    _eager(torch, g, b, 3000)
    L.check(lib.dust_fmm_graph_launch(g.handle), "launch")
    L.check(lib.dust_fmm_graph_status(g.handle, None), "status")
    g.destroy()


@pytest.mark.parametrize("precision", ["fp64", "tuned"])
def test_graph_replay_against_oracle(orc, precision):
    """A replay captured on one draw, run on another, against the FP64 oracle
    on that draw: FP64 form within 1e-10, tuned within 1e-6 (relative L2 of
    velocity; stretching 1e-10 / 1e-5)."""
    import torch
    lib = L.lib()
    cap = 12000
    first = W.helix_wake(9000, seed=501)
    tree = dict(first[3], nlevels=5)
    g = _tree_handle(tree, precision, cap)
    b = Bufs(torch, cap)
    b.load(torch, *first[:3])
    _eager(torch, g, b, 9000)
    L.check(lib.dust_fmm_graph_reserve(g.handle, 2.5), "reserve")
    L.check(lib.dust_fmm_graph_capture(g.handle, *b.args()), "capture")
    cen, alpha, rv, _ = W.helix_wake(11000, seed=502, pitch=0.55)
    n = len(cen)
    print(f"graph form ({precision}) vs oracle, n={n}", flush=True)
    b.load(torch, cen, alpha, rv)
    L.check(lib.dust_fmm_graph_launch(g.handle), "launch")
    L.check(lib.dust_fmm_graph_status(g.handle, None), "status")
    vel = b.vel[:n].cpu().numpy()
    st = b.st[:n].cpu().numpy()
    g.destroy()
    mag = np.linalg.norm(alpha, axis=1)
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], tree["nlevels"],
                   tree["min_part"], 4, tree["delta"])
    v, s, _ = c.run(cen, alpha / mag[:, None], mag, rv)
    tv, ts = (1e-10, 1e-10) if precision == "fp64" else (1e-6, 1e-5)
    assert np.linalg.norm(vel - v) / np.linalg.norm(v) < tv
    assert np.linalg.norm(st - s) / np.linalg.norm(s) < ts
