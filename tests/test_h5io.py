"""HDF5 result files (libdust_h5.so, dust_amd.h5io) without a GPU.

- typed datasets, strings and root attributes round-trip through the file;
  missing datasets and size mismatches are errors;
- a file written by save_status has the layout of the reference's own
  result files (tests/golden/reference_h5_layout.json, read from the
  regression fixtures of the reference by tests/golden/make_golden.py):
  the same dataset paths, shapes as h5py reports them and element types;
  read by h5py (an independent reader, when the image has one) with the
  reference's regression-driver comparison (tests/run_b.py: relative L2 of
  /ParticleWake/WakePoints, /ParticleWake/WakeVort, .../Solution/Vort);
- load_wake's particle reconstruction (mag = |WakeVort|, dir = WakeVort /
  mag, WakeVort itself below 1e-13) on the reference's own particle field.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from dust_amd import h5io as H

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CONDA_PY = "/opt/conda/bin/python3.9"


def have_h5py():
    if not os.path.exists(CONDA_PY):
        return False
    return subprocess.run([CONDA_PY, "-c", "import h5py"], capture_output=True).returncode == 0


def test_round_trip(tmp_path):
    f = tmp_path / "a.h5"
    a = np.arange(24, dtype=np.float64).reshape(2, 4, 3) * 0.5
    i = np.array([[1, 2], [3, 4], [5, 6]], dtype=np.int32)
    with H.File(f, "w") as w:
        w.write("G/H/a", a)
        w.write("G/i", i)
        w.write("s", np.float64(2.5))
        w.write("n", np.int32(7))
        w.write("empty", np.zeros((3, 0)))
        w.write_string("G/name", "Wing")
        w.attr("dt", 0.01)
        w.attr("Nbox", np.array([1, 2, 3], np.int32))
        w.attr("use_fmm", True)
        w.attr("basename", "run", H.MAX_CHAR_LEN)
        with pytest.raises(H.H5Error):
            w.write("G/i", i)  # exists already
    with H.File(f) as r:
        assert np.array_equal(r.read("G/H/a"), a)
        assert r.read("G/i").dtype == np.int32 and np.array_equal(r.read("G/i"), i)
        assert r.read("s").shape == () and float(r.read("s")) == 2.5
        assert int(r.read("n")) == 7
        assert r.read("empty").shape == (3, 0)
        assert r.read("G/name") == "Wing"
        assert r.exists("G/H/a") and not r.exists("G/H/b") and not r.exists("X/y")
        assert r.info("G/H/a") == (H.F64, (2, 4, 3))
        with pytest.raises(H.H5Error):
            r.read("missing")
    with pytest.raises(H.H5Error):
        H.File(tmp_path / "nope.h5")


def test_result_name():
    assert H.result_name("out/wing", 11) == "out/wing_res_0011.h5"


def synthetic_state(rng, ne=200, ns=40, npt=42, ln=2, n=360):
    """A state with the dimensions of the reference's mir_test static/p
    result (two surface-panel components of 200 panels, 40 stripes)."""
    comps = [dict(comp_id=c, name=f"Comp{c}", ref_tag="Wing", ref_id=c, vort=rng.normal(size=ne),
                  pres=rng.normal(size=ne), dforce=rng.normal(size=(ne, 3)),
                  dmom=rng.normal(size=(ne, 3)), surf_vel=rng.normal(size=(ne, 3)),
                  didou_dt=rng.normal(size=ne)) for c in (1, 2)]
    pw = dict(points=rng.normal(size=(ln + 1, npt, 3)), vels=rng.normal(size=(ln + 1, npt, 3)),
              start_points=np.stack([np.arange(1, ns + 1), np.arange(2, ns + 2)], 1),
              vort=rng.normal(size=(ln, ns)))
    parts = dict(points=rng.normal(size=(n, 3)), vels=rng.normal(size=(n, 3)),
                 vort=rng.normal(size=(n, 3)), turbvisc=np.zeros(n), vortex_rad=np.full(n, 0.1),
                 last_pan_idou=rng.normal(size=ns))
    params = dict(u_inf=np.array([30.0, 0.0, 0.0]), P_inf=0.0, rho_inf=1.225, mu_inf=1.8e-5,
                  a_inf=340.0, dt=0.01, n_wake_panels=ln, use_fmm=False, basename="./Output/ref")
    return comps, pw, parts, params


def test_layout_matches_the_reference(tmp_path):
    comps, pw, parts, params = synthetic_state(np.random.default_rng(3))
    f = tmp_path / "ours_res_0011.h5"
    H.save_status(f, 0.5, params=params, components=comps, panel_wake=pw, particles=parts,
                  ring_wake_len=2)
    ref = json.load(open(os.path.join(GOLDEN, "reference_h5_layout.json")))
    lay = ref["layout"]["mir_test_static_p"]["datasets"]
    with H.File(f) as r:
        for name, (shape, dt) in lay.items():
            if name.startswith("References/"):
                continue  # reference frames: not produced on this path
            assert r.exists(name), name
            t, sh = r.info(name)
            if dt.startswith("|S"):
                assert t == H.STR, name
                continue
            assert t == (H.F64 if dt == "<f8" else H.I32), name
            assert list(sh) == shape, (name, sh, shape)
        # the reading side of load_wake and load_solution
        w = H.read_wake(f)
        assert np.array_equal(w["points"], parts["points"])
        assert np.array_equal(w["pan_vort"], pw["vort"])
        assert np.array_equal(w["start_points"], pw["start_points"])
        assert np.array_equal(H.read_solution(f, 2), comps[1]["vort"])


@pytest.mark.skipif(not have_h5py(), reason="no h5py interpreter in this image")
def test_h5py_reads_our_files_like_the_regression_driver(tmp_path):
    """h5py reads the datasets tests/run_b.py compares, with the shapes and
    types of the reference's files; the driver's relative-L2 comparison of
    two of our files written from the same state with one perturbed array."""
    rng = np.random.default_rng(5)
    comps, pw, parts, params = synthetic_state(rng)
    a, b = tmp_path / "ref_res_0011.h5", tmp_path / "test_res_0011.h5"
    H.save_status(a, 0.5, params=params, components=comps, panel_wake=pw, particles=parts,
                  ring_wake_len=2)
    parts_b = dict(parts, points=parts["points"] * (1 + 1e-9))
    H.save_status(b, 0.5, params=params, components=comps, panel_wake=pw, particles=parts_b,
                  ring_wake_len=2)
    script = r"""
import json, sys
import h5py, numpy as np
a, b = sys.argv[1], sys.argv[2]
sol = ['/ParticleWake/WakePoints', '/ParticleWake/WakeVort',
       '/Components/Comp001/Solution/Vort', '/Components/Comp001/Solution/Pres']
ra, rb = h5py.File(a, 'r'), h5py.File(b, 'r')
err = [float(np.linalg.norm(np.array(rb[d]) - np.array(ra[d])) / np.linalg.norm(np.array(ra[d])))
       for d in sol]
lay = {}
ra.visititems(lambda n, o: lay.__setitem__(n, [list(o.shape), o.dtype.str])
              if isinstance(o, h5py.Dataset) else None)
print(json.dumps({"err": err, "layout": lay, "attrs": sorted(ra.attrs.keys()),
                  "use_fmm": ra.attrs["use_fmm"].decode(), "NComponents": int(ra['Components/NComponents'][()])}))
"""
    p = subprocess.run([CONDA_PY, "-c", script, str(a), str(b)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    out = json.loads(p.stdout)
    assert out["err"][0] == pytest.approx(1e-9, rel=1e-3)
    assert out["err"][1:] == [0.0, 0.0, 0.0]
    assert out["use_fmm"] == "F" and out["NComponents"] == 2
    ref = json.load(open(os.path.join(GOLDEN, "reference_h5_layout.json")))
    lay = ref["layout"]["mir_test_static_p"]["datasets"]
    for name, (shape, dt) in out["layout"].items():
        assert name in lay, name
        if not dt.startswith("|S"):
            assert [shape, dt] == lay[name], name
    assert set(params) <= set(ref["layout"]["reg_test_b"]["attrs"]) | {"n_wake_panels"}


def test_particles_from_the_reference_field():
    """load_wake's mag/dir on the ParticleWake of the reference's
    reg_test_b result, plus a zero and a tiny vector."""
    g = np.load(os.path.join(GOLDEN, "reg_test_b_particles.npz"))
    vort = np.vstack([g["alpha"], np.zeros(3), np.full(3, 1e-15)])
    mag, d = H.particles_from_file(vort)
    big = mag > 1e-13
    assert np.allclose(mag[big], np.linalg.norm(vort[big], axis=1), rtol=1e-15, atol=0)
    assert np.allclose(np.linalg.norm(d[big], axis=1), 1.0, rtol=0, atol=4e-16)
    assert np.allclose(d[big] * mag[big, None], vort[big], rtol=4e-16, atol=1e-30)
    assert np.array_equal(d[~big], vort[~big]) and (~big).sum() == 2


def test_library_exports_every_declared_symbol():
    import re
    src = open(os.path.join(HERE, "..", "include", "dust_h5.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(dust_h5_[a-z0-9_]+)\s*\(", src)))
    assert len(names) == len(H.SIGNATURES)
    assert {s[0] for s in H.SIGNATURES} == set(names)
    lib = H.lib()
    assert all(hasattr(lib, n) for n in names)
