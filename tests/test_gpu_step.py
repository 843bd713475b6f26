"""Whole time steps with DUST's wake model on the device
(dust_amd.step.BodyWakeStep: uvort with the end vortices, assembly with the
implicit wake row and the explicit wake panels, solve, first-row
intensities, velocities at the panel points and at the particles, panel
convection, trailing-edge rows, particle update, shedding from the last panel
row) against the CPU oracle pipeline doing the same (oracle/pipeline.py):
the 1,600-panel wing with its 8,000-particle sheet and a two-blade rotor.
FP64: panel intensities, uvort, velocities, stretching, panel points and
particle states to 1e-10 relative L2, slots bit-exact; the tuned forms to
1e-6 of the FP64 form."""
import numpy as np
import pytest

from dust_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-10
STEPS = 4          # the panel wake (2 rows) is full from the second step on: shedding
NWP = 2
WING = dict(u=(30.0, 0.0, 0.0), dt=1.0 / 600.0, strips=(20, 40))


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def case():
    ver, nv, cen, alpha, rv, tree = W.wing_case()
    return ver, nv, cen, alpha, rv, tree


def make_step(case, form, **kw):
    from dust_amd.step import BodyWakeStep
    ver, nv, cen, alpha, rv, tree = case
    # refined shedding: 2 x 16 particles per stripe and step
    extra = 4000 if kw.get("refine_wake") else 400
    s = BodyWakeStep(ver, nv, WING["strips"], WING["u"], WING["dt"], len(rv) + extra, tree,
                     form=form, n_wake_panels=NWP, **kw)
    s.seed(cen, alpha, rv)
    return s


def compare(g, o, r, tol=TOL):
    assert rel(g["mu"], r["mu"]) <= tol
    assert rel(g["uvort"], r["uvort"]) <= tol
    assert rel(g["vel"][r["live"]], r["vel"]) <= tol
    assert rel(g["stretch"][r["live"]], r["stretch"]) <= tol
    assert np.array_equal(g["free"], o.pool["free"])
    lv = np.flatnonzero(o.pool["free"] == 0)
    assert rel(g["cen"][lv], o.pool["cen"][lv]) <= tol
    assert rel(g["dir"][lv], o.pool["dir"][lv]) <= tol
    assert rel(g["mag"][lv], o.pool["mag"][lv]) <= tol
    for key in ("pts", "idou", "last_idou"):
        assert rel(g["pw_" + key], getattr(o.pw, key)) <= tol, key


def test_wing_steps_with_a_gust_match_oracle(orc, case):
    """An 'ACM' gust front sweeping the wing along x at the wind speed
    (variable_wind, mod_wind.f90:66-121) in the panel sources, the assembled
    right-hand side, the panel-wake points and rows, and the particles."""
    from dust_amd import octree as O
    from oracle.pipeline import OracleStep
    w = O.wind_params(WING["u"], "ACM", origin=(0.5, 0.0, 0.0), front_direction=(1.0, 0.0, 0.0),
                      front_speed=30.0, u_des=3.0, gradient=0.5)
    g13 = orc.wind_params(WING["u"], origin=(0.5, 0.0, 0.0), front_direction=(1.0, 0.0, 0.0),
                          front_speed=30.0, u_des=3.0, gradient=0.5)
    s = make_step(case, "fp64", wind=w)
    o = OracleStep(case, WING["strips"], s.pool.nmax, WING["u"], WING["dt"], n_wake_panels=NWP,
                   wind=(1, g13))
    plain = make_step(case, "fp64")
    for k in range(STEPS):
        print(f"wing step {k} with a gust", flush=True)
        live_g = s.pool.live().cpu().numpy()
        s.step()
        plain.step()
        g = s.state()
        r = o.step()
        assert np.array_equal(live_g, r["live"])
        compare(g, o, r)
    # the gust changed the solution
    assert rel(plain.state()["mu"], g["mu"]) > 1e-4


@pytest.mark.parametrize("kw", [{}, dict(refine_wake=True, k_refine=2, k_vortex_rad=0.5)])
def test_wing_steps_match_oracle(orc, case, kw):
    from oracle.pipeline import OracleStep
    s = make_step(case, "fp64", **kw)
    o = OracleStep(case, WING["strips"], s.pool.nmax, WING["u"], WING["dt"], n_wake_panels=NWP,
                   **kw)
    for k in range(STEPS):
        print(f"wing step {k} {kw}", flush=True)
        live_g = s.pool.live().cpu().numpy()
        s.step()
        g = s.state()
        r = o.step()
        assert np.array_equal(live_g, r["live"])
        print(f"mu {rel(g['mu'], r['mu']):.2e} uvort {rel(g['uvort'], r['uvort']):.2e} "
              f"pts {rel(g['pw_pts'], o.pw.pts):.2e} live {len(r['live'])}")
        compare(g, o, r)
    assert o.pw.full
    print("stage ms", {k: round(v / STEPS, 3) for k, v in s.times.items()})


def test_tuned_fmm_steps_match_fp64(case):
    a = make_step(case, "fp64")
    b = make_step(case, "tuned")
    for k in range(STEPS):
        print(f"tuned step {k}", flush=True)
        a.step()
        b.step()
    ga, gb = a.state(), b.state()
    assert np.array_equal(ga["free"], gb["free"])
    lv = np.flatnonzero(ga["free"] == 0)
    print(f"mu {rel(gb['mu'], ga['mu']):.2e} cen {rel(gb['cen'][lv], ga['cen'][lv]):.2e}")
    assert rel(gb["mu"], ga["mu"]) <= 1e-6
    assert rel(gb["vel"][lv], ga["vel"][lv]) <= 1e-6
    assert rel(gb["cen"][lv], ga["cen"][lv]) <= 1e-9


@pytest.fixture(scope="module")
def rotor_case():
    """Two blades of aspect ratio 4 (512 panels) turning at 20 rad/s, 2 degrees
    per step, in an axial flow of 5 m/s, with 5,000 particles of the helix
    generator (x1.5) around and through the rotor disk."""
    blades, panels, aspect = 2, 512, 4.0
    ver, nv = W.rotor(blades, panels, aspect=aspect)
    nc, ns = W.rotor_strips(blades, panels, aspect)
    cen, alpha, rv, _ = W.helix_wake(5000, seed=9)
    tree = dict(box_length=12.0, nbox=(1, 1, 1), origin=(-6.0, -6.0, -6.0), nlevels=4,
                min_part=32, delta=0.05)
    return (ver, nv, 1.5 * cen, alpha, rv, tree), (blades * ns, nc)


ROTOR = dict(u=(0.0, 0.0, -5.0), dt=np.radians(2.0) / 20.0, omega=20.0)


def make_rotor(rotor_case, form, solver, **kw):
    from dust_amd.step import BodyWakeStep
    (ver, nv, cen, alpha, rv, tree), strips = rotor_case
    s = BodyWakeStep(ver, nv, strips, ROTOR["u"], ROTOR["dt"], len(rv) + 400, tree, form=form,
                     vortex_rad=0.05, omega=ROTOR["omega"], solver=solver, n_wake_panels=NWP,
                     **kw)
    s.seed(cen, alpha, rv)
    return s


def test_rotor_steps_match_oracle(orc, rotor_case):
    from oracle.pipeline import OracleStep
    case, strips = rotor_case
    s = make_rotor(rotor_case, "fp64", "fp64")
    o = OracleStep(case, strips, s.pool.nmax, ROTOR["u"], ROTOR["dt"], vortex_rad=0.05,
                   omega=ROTOR["omega"], n_wake_panels=NWP)
    for k in range(3):
        print(f"rotor step {k}", flush=True)
        s.step()
        g = s.state()
        r = o.step()
        print(f"mu {rel(g['mu'], r['mu']):.2e} uvort {rel(g['uvort'], r['uvort']):.2e} "
              f"vel {rel(g['vel'][r['live']], r['vel']):.2e}")
        compare(g, o, r)


def test_rotor_tuned_matches_fp64(rotor_case):
    a = make_rotor(rotor_case, "fp64", "fp64")
    b = make_rotor(rotor_case, "tuned", "fp16s")
    for k in range(3):
        print(f"rotor tuned step {k}", flush=True)
        a.step()
        b.step()
    ga, gb = a.state(), b.state()
    lv = np.flatnonzero(ga["free"] == 0)
    print(f"mu {rel(gb['mu'], ga['mu']):.2e} vel {rel(gb['vel'][lv], ga['vel'][lv]):.2e}")
    assert np.array_equal(ga["free"], gb["free"])
    assert rel(gb["mu"], ga["mu"]) <= 1e-8
    assert rel(gb["vel"][lv], ga["vel"][lv]) <= 1e-6


def test_fmm_pan_steps_match_oracle(orc, case):
    """use_fmm_pan: the tree and expansions of prepare_wake give uvort through
    apply_multipole_panels (plus the end vortices) and then the particle
    velocities."""
    from oracle.pipeline import OracleStep
    s = make_step(case, "fp64", fmm_pan=True)
    o = OracleStep(case, WING["strips"], s.pool.nmax, WING["u"], WING["dt"], n_wake_panels=NWP,
                   fmm_pan=True)
    for k in range(3):
        print(f"fmm_pan step {k}", flush=True)
        s.step()
        g = s.state()
        r = o.step()
        print(f"mu {rel(g['mu'], r['mu']):.2e} uvort {rel(g['uvort'], r['uvort']):.2e}")
        compare(g, o, r)


@pytest.mark.parametrize("form,solver", [("fp64", "fp64"), ("tuned", "fp16s")])
def test_rotor_overlap_matches_sequential(rotor_case, form, solver):
    """The FMM pass run from a second host thread on its own stream while the
    panel system is assembled and solved gives the same state as the
    sequential order: bit for bit in the FP64 form; within 1e-12 in the tuned
    form, whose M2L adds into the FP64 locals with atomics from several jobs
    per target cell (the order of those additions varies from run to run)."""
    a = make_rotor(rotor_case, form, solver, overlap=False)
    b = make_rotor(rotor_case, form, solver, overlap=True)
    assert not a.overlap and b.overlap
    for k in range(3):
        print(f"rotor overlap {form}/{solver} step {k}", flush=True)
        a.step()
        b.step()
        ga, gb = a.state(), b.state()
        assert np.array_equal(ga["free"], gb["free"])
        for key in ("mu", "uvort", "vel", "cen"):
            if form == "fp64":
                assert np.array_equal(ga[key], gb[key]), key
            else:
                assert rel(gb[key], ga[key]) <= 1e-12, key


@pytest.mark.parametrize("which", ["wing", "rotor"])
def test_restart_continues_bit_for_bit(case, rotor_case, tmp_path, which):
    """Two steps, save_restart, a fresh step object loads it and runs two more:
    the state equals four uninterrupted steps bit for bit (FP64 form)."""
    mk = (lambda: make_step(case, "fp64")) if which == "wing" else \
        (lambda: make_rotor(rotor_case, "fp64", "fp64"))
    a = mk()
    for k in range(4):
        print(f"restart reference {which} step {k}", flush=True)
        a.step()
    b = mk()
    for k in range(2):
        b.step()
    f = tmp_path / "restart.npz"
    b.save_restart(f)
    c = mk()
    c.load_restart(f)
    for k in range(2):
        print(f"restarted {which} step {k}", flush=True)
        c.step()
    ga, gc = a.state(), c.state()
    for key in ga:
        assert np.array_equal(ga[key], gc[key]), key
    with np.load(f) as d:
        assert d["WakePoints"].shape == d["WakeVort"].shape
        assert len(d["VortexRad"]) == len(d["WakePoints"])


def _cloud(g):
    """Live particles of a state, in an order independent of their slots."""
    lv = g["live"]
    c = np.hstack([g["cen"][lv], g["dir"][lv] * g["mag"][lv][:, None]])
    return c[np.lexsort(np.round(c[:, :3], 8).T)]


@pytest.mark.parametrize("which", ["wing", "rotor"])
def test_result_file_restart(case, rotor_case, tmp_path, which):
    """Two steps, save_status to an HDF5 result file in the reference's
    layout, a fresh step object restarts from it through load_wake /
    load_solution and runs two more: the state matches four uninterrupted
    steps (FP64 form; not bit for bit, since load_wake recomputes mag and dir
    from their product and the particles move to the lowest slots)."""
    from dust_amd import h5io as H
    mk = (lambda: make_step(case, "fp64")) if which == "wing" else \
        (lambda: make_rotor(rotor_case, "fp64", "fp64"))
    a = mk()
    for k in range(4):
        print(f"h5 restart reference {which} step {k}", flush=True)
        a.step()
    b = mk()
    for k in range(2):
        b.step()
    f = tmp_path / H.result_name("run", 2)
    b.save_status(f, comp_name="body")
    gb = b.state()
    w = H.read_wake(f)
    assert np.array_equal(w["points"], gb["cen"][gb["live"]])
    assert w["time"] == pytest.approx(2 * b.dt, rel=1e-15)
    ln, _, _ = b.pw.state()
    assert w["pan_points"].shape == (ln + 1, b.npt, 3)
    assert np.array_equal(w["pan_vort"], gb["pw_idou"][:ln])
    c = mk()
    c.load_wake(f)
    assert c.t == b.t
    gc0 = c.state()
    assert np.array_equal(gc0["mu"], gb["mu"])
    assert np.array_equal(gc0["pw_pts"][:ln + 1], gb["pw_pts"][:ln + 1])
    assert np.array_equal(gc0["pw_idou"][:ln], gb["pw_idou"][:ln])
    assert np.array_equal(gc0["pw_last_idou"], gb["pw_last_idou"])
    assert rel(_cloud(gc0), _cloud(gb)) <= 1e-15
    for k in range(2):
        print(f"h5 restarted {which} step {k}", flush=True)
        c.step()
    ga, gc = a.state(), c.state()
    assert rel(gc["mu"], ga["mu"]) <= 1e-10
    assert rel(gc["uvort"], ga["uvort"]) <= 1e-10
    assert len(gc["live"]) == len(ga["live"])
    assert rel(_cloud(gc), _cloud(ga)) <= 1e-10
    for key in ("pts", "idou", "last_idou"):
        assert rel(gc["pw_" + key], ga["pw_" + key]) <= 1e-10, key
    other = mk()
    other.te = dict(other.te, start=other.te["start"][::-1].copy())
    with pytest.raises(ValueError, match="trailing edge connectivity"):
        other.load_wake(f)
