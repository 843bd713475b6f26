"""Device particle pool (dust_wake_update_dev, dust_wake_live_dev,
dust_wake_shed_dev) against the CPU oracle of src/aeroel/mod_wake.f90
(complete_wake :1305-1355 and :1888-1967, prepare_wake :826-835): slot flags,
live lists and inserted slots bit-exact, positions, directions and stretching
to 1e-14 relative (the device contracts a*b + c into FMAs)."""
import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import wake as WK
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu
BOX = (np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, 1.0]))


def host_pool(n, seed, free_frac=0.3):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return dict(free=(rng.random(n) < free_frac).astype(np.uint8),
                cen=rng.uniform(-0.95, 0.95, (n, 3)), dir=d, mag=rng.uniform(0.1, 1.0, n),
                rv=rng.uniform(0.01, 0.05, n), vel=rng.standard_normal((n, 3)),
                stretch=rng.standard_normal((n, 3)), rotu=rng.standard_normal((n, 3)))


def to_device(hp):
    import torch
    P = WK.ParticlePool(len(hp["free"]))
    P.free.copy_(torch.from_numpy(hp["free"]))
    for k, attr in (("cen", "cen"), ("dir", "dir"), ("mag", "mag"), ("rv", "r_vortex"),
                    ("vel", "vel"), ("stretch", "stretch"), ("rotu", "rotu")):
        getattr(P, attr).copy_(torch.from_numpy(hp[k]))
    P.n_prt.fill_(int(np.count_nonzero(hp["free"] == 0)))
    return P


def close(a, b, tol=1e-14):
    return np.allclose(a, b, rtol=tol, atol=tol * np.abs(b).max())


@pytest.mark.parametrize("n,divfilt", [(1, False), (1000, True), (300_001, True),
                                       (300_001, False)])
def test_update_vs_oracle(orc, n, divfilt):
    hp = host_pool(n, n)
    P = to_device(hp)
    n0 = int(P.n_prt.item())
    prm = WK.params(*BOX, dt=0.05, filt_eta=0.3, use_vs=True, use_divfilt=divfilt)
    print(f"wake_update n={n} divfilt={divfilt}", flush=True)
    P.update(prm)
    freed = orc.wake_update(hp, *BOX, 0.05, 0.3, use_vs=True, use_divfilt=divfilt)
    assert np.array_equal(P.free.cpu().numpy(), hp["free"])
    assert int(P.n_prt.item()) == n0 - freed
    assert close(P.cen.cpu().numpy(), hp["cen"])
    assert close(P.dir.cpu().numpy(), hp["dir"])
    assert close(P.stretch.cpu().numpy(), hp["stretch"])
    assert np.array_equal(P.live().cpu().numpy(), orc.wake_live(hp["free"]))


def test_many_steps_and_shedding(orc):
    """Twenty steps of update + shedding of 64 particles per step in a pool
    that fills up: the device pool follows the oracle slot by slot."""
    rng = np.random.default_rng(11)
    hp = host_pool(5000, 3, free_frac=0.9)
    P = to_device(hp)
    prm = WK.params(*BOX, dt=0.02, use_vs=True)
    for step in range(20):
        print(f"wake step {step}", flush=True)
        P.vel.copy_(P.vel.new_tensor(rng.standard_normal((5000, 3))))
        hp["vel"] = P.vel.cpu().numpy()
        P.update(prm)
        orc.wake_update(hp, *BOX, 0.02, use_vs=True)
        k = 64
        ncen = rng.uniform(-1.05, 1.05, (k, 3))
        npv = rng.standard_normal((k, 3))
        nrv = np.full(k, 0.02)
        nvel = rng.standard_normal((k, 3))
        P.shed(ncen, npv, nrv, nvel, prm)
        assert orc.wake_shed(hp, ncen, npv, nrv, nvel, *BOX) >= 0
        assert np.array_equal(P.free.cpu().numpy(), hp["free"])
    assert int(P.n_prt.item()) == np.count_nonzero(hp["free"] == 0)
    assert close(P.cen.cpu().numpy(), hp["cen"], 1e-13)
    assert close(P.dir.cpu().numpy(), hp["dir"], 1e-13)
    assert close(P.mag.cpu().numpy(), hp["mag"], 1e-15)
    assert np.array_equal(P.r_vortex.cpu().numpy(), hp["rv"])


def test_shed_overflow_is_reported(orc):
    hp = host_pool(100, 5, free_frac=0.0)
    hp["free"][[7, 50]] = 1
    P = to_device(hp)
    prm = WK.params(*BOX, dt=0.01)
    k = 3
    print("wake_shed overflow", flush=True)
    with pytest.raises(L.DustError) as ei:
        P.shed(np.zeros((k, 3)), np.ones((k, 3)), np.full(k, 0.1), np.zeros((k, 3)), prm)
    assert ei.value.code == L.DUST_E_CAPACITY
    # the slots that were free are taken, as the reference fills them before it stops
    assert np.count_nonzero(P.free.cpu().numpy()) == 0


def test_empty_pool_and_no_new_particles():
    P = WK.ParticlePool(0)
    prm = WK.params(*BOX, dt=0.01)
    P.update(prm)
    assert P.live().numel() == 0
    Q = WK.ParticlePool(10)
    Q.shed(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros((0, 3)), prm)
    assert Q.live().numel() == 0


def test_avoid_collision_vs_oracle(orc):
    """avoid_collision on a pool around a two-blade rotor mesh given once as
    surface panels and once as vortex-lattice panels (both branches, the
    thickness doubling), panels moving with omega x r, particles placed just
    off the panels on both sides with velocities towards them, plus far
    ones."""
    import torch
    ver, nv = W.rotor(2, 300, aspect=3.0)
    G = orc.geo(ver, nv)
    cen = G[:, orc.G_CEN:orc.G_CEN + 3]
    nor = G[:, orc.G_NOR:orc.G_NOR + 3]
    me = np.array([G[j, orc.G["elen"]:orc.G["elen"] + 4].max() for j in range(len(nv))])
    ub = np.stack([-5.0 * cen[:, 1], 5.0 * cen[:, 0], np.zeros(len(nv))], axis=1)
    rng = np.random.default_rng(21)
    m = 6000
    pick = rng.integers(0, len(nv), m)
    h = rng.uniform(-0.03, 0.03, m)
    pos = cen[pick] + h[:, None] * nor[pick] + 0.02 * rng.standard_normal((m, 3))
    vel = -np.sign(h)[:, None] * nor[pick] * rng.uniform(0.5, 5.0, (m, 1)) + \
        rng.standard_normal((m, 3))
    far = rng.uniform(-6.0, 6.0, (2000, 3))
    hp = host_pool(m + 2000, 4, free_frac=0.1)
    hp["cen"] = np.concatenate([pos, far])
    hp["vel"] = np.concatenate([vel, rng.standard_normal((2000, 3))])
    hp["rv"] = rng.uniform(0.005, 0.03, m + 2000)
    P = to_device(hp)
    kinds = np.concatenate([np.zeros(len(nv), np.int32), np.ones(len(nv), np.int32)])
    groups = [dict(kind=0, ver=ver, n_ver=nv), dict(kind=1, ver=ver, n_ver=nv)]
    tub = torch.from_numpy(np.ascontiguousarray(ub)).cuda()
    prm = WK.pa_params(0.01, 0.02, 10.0, 1.0, 1.0)
    print(f"avoid_collision n={m + 2000} elements={2 * len(nv)}", flush=True)
    P.avoid_collision(groups, [tub, tub], prm)
    orc.avoid_collision_pool(kinds, np.concatenate([cen, cen]), np.concatenate([nor, nor]),
                             np.concatenate([me, me]), np.concatenate([ub, ub]), hp, 0.01, 0.02,
                             10.0, 1.0, 1.0)
    gv = P.vel.cpu().numpy()
    changed = np.any(hp["vel"] != np.concatenate([vel, hp["vel"][m:]]), axis=1)
    print(f"changed {np.count_nonzero(changed)}")
    assert np.count_nonzero(changed[:m]) > 100
    assert close(gv, hp["vel"], 1e-13)
