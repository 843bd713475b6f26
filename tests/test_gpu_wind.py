"""variable_wind on the device (src/aeroel/mod_wind.f90:66-121) against the
oracle's restatement, and as the FMM's free-stream term (mod_octree.f90:1376)
in both forms.  The device cos may differ from the host's in the last bit:
1e-15 relative on the wind."""
import ctypes as C

import numpy as np
import pytest

from dust_amd import _lib as L
from dust_amd import octree as O
from dust_amd import workloads as W

pytestmark = pytest.mark.gpu
U = (30.0, 1.0, -2.0)


def oracle_gust(orc, w):
    return orc.wind_params(list(w.u_inf), origin=list(w.origin),
                           front_direction=list(w.front_direction), front_speed=w.front_speed,
                           u_des=w.u_des, gradient=w.gradient, start_time=w.start_time)


@pytest.mark.parametrize("gust", [None, "ACM", "linear"])
def test_variable_wind_vs_oracle(orc, gust):
    import torch
    rng = np.random.default_rng(3)
    pos = rng.uniform(-2, 2, size=(5000, 3))
    w = O.wind_params(U, gust, origin=(0.5, 0.0, 0.0), front_direction=(2.0, 1.0, 0.0),
                      front_speed=4.0, u_des=3.0, gradient=0.6, start_time=0.1)
    print(f"variable_wind {gust}", flush=True)
    got = O.variable_wind(w, pos, 0.4)
    ref = orc.variable_wind(w.gust_type, oracle_gust(orc, w), pos, 0.4)
    assert np.allclose(got, ref, rtol=1e-15, atol=1e-14)
    if gust == "ACM":
        assert np.count_nonzero(got[:, 0] != U[0]) > 500
    dp = torch.from_numpy(pos).cuda()
    dw = torch.empty_like(dp)
    L.check(L.lib().dust_variable_wind_dev(len(pos), dp.data_ptr(), C.byref(w), 0.4,
                                           dw.data_ptr(), None), "wind dev")
    torch.cuda.synchronize()
    assert np.array_equal(dw.cpu().numpy(), got)


@pytest.mark.parametrize("precision", ["fp64", "tuned"])
def test_fmm_free_stream_with_a_gust(orc, precision):
    cen, alpha, rv, tree = W.helix_wake(12000, seed=9)
    x0 = float(np.median(cen[:, 0]))
    w = O.wind_params(U, "ACM", origin=(x0, 0, 0), front_direction=(1, 0, 0), front_speed=2.0,
                      u_des=1.5, gradient=0.3)
    print(f"FMM with an ACM gust ({precision})", flush=True)

    def run(gusty):
        p = O.VortParts.from_alpha(cen, alpha, rv)
        g = O.initialize_octree(tree["box_length"], tree["nbox"], tree["origin"], 4,
                                tree["min_part"], 4, tree["delta"], capacity=p.n, u_inf=U,
                                precision=precision)
        if gusty:
            g.set_wind(w, 0.05)
        O.sort_particles(p, g)
        O.calculate_multipole(p, g)
        O.apply_multipole(p, g)
        g.destroy()
        return p

    pg, p0 = run(True), run(False)
    wind = O.variable_wind(w, pg.cen, 0.05) - np.asarray(U)
    assert np.count_nonzero(wind[:, 0]) > 100
    tol = 1e-13 if precision == "fp64" else 1e-11
    assert np.allclose(pg.vel - p0.vel, wind, rtol=0, atol=tol)
    c = orc.Octree(tree["box_length"], tree["nbox"], tree["origin"], 4, tree["min_part"], 4,
                   tree["delta"], u_inf=U)
    c.set_wind(1, oracle_gust(orc, w), 0.05)
    v, _, _ = c.run(pg.cen, pg.dir, pg.mag, pg.r_vortex)
    rel = np.linalg.norm(pg.vel - v) / np.linalg.norm(v)
    assert rel <= (1e-10 if precision == "fp64" else 1e-6)
