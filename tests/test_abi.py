"""The C ABI: libdust_gpu.so loads without a GPU and exports every function
declared in include/dust_gpu.h; the Python bindings cover the same set.  No
compute calls (those need a GPU)."""
import ctypes
import os
import re

from dust_amd import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "include", "dust_gpu.h")
TEST_HEADER = os.path.join(HERE, "..", "include", "dust_gpu_test.h")


def declared(header=HEADER):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dust_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("dust_fmm_create", "dust_fmm_sort_particles", "dust_fmm_calculate_multipole",
                 "dust_fmm_apply_multipole", "dust_aic_surfpan", "dust_dgetrf", "dust_dgetrs",
                 "dust_dgesv_mixed", "dust_solve_linsys"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in declared() + declared(TEST_HEADER) if not hasattr(lib, n)]
    assert not missing, missing


def test_test_hooks_stay_out_of_the_product_header():
    hooks = declared(TEST_HEADER)
    assert hooks and all(n.startswith("dust_test_") for n in hooks)
    assert not [n for n in declared() if n.startswith("dust_test_")]


def test_python_signatures_match_header():
    sig = {s[0] for s in L.SIGNATURES}
    names = set(declared()) | set(declared(TEST_HEADER))
    assert names <= sig | {"dust_last_error", "dust_version"}, names - sig


def test_version_and_error_text():
    lib = L.lib()
    assert lib.dust_version().decode()
    # an argument error is reported without touching a GPU
    rc = lib.dust_dgetrf(-1, 1, None, 1, None)
    assert rc == L.DUST_E_ARG
    assert lib.dust_last_error().decode()


def test_element_argument_errors_without_gpu():
    """dust_elem_apply / dust_vort_vel reject bad lists before any device work."""
    import numpy as np
    lib = L.lib()
    prm = L.ElemParams(10.0, 0.1, 0.001, 0.1)
    ver = np.zeros((1, 4, 3))
    nv = np.array([5], np.int32)
    grp = (L.ElemGroup * 1)()
    grp[0].kind, grp[0].n, grp[0].n_ver, grp[0].ver = L.ELEM_RING, 1, nv.ctypes.data, \
        ver.ctypes.data
    pos = np.zeros(3)
    vel = np.zeros(3)
    assert lib.dust_elem_apply(1, grp, ctypes.byref(prm), 1, pos.ctypes.data, None,
                               vel.ctypes.data, None) == L.DUST_E_ARG
    assert "vertices" in lib.dust_last_error().decode()
    grp[0].kind = 7
    assert lib.dust_elem_apply(1, grp, ctypes.byref(prm), 1, pos.ctypes.data, None,
                               vel.ctypes.data, None) == L.DUST_E_ARG
    assert lib.dust_vort_vel(-1, None, 0, None, None, None, None, 0, None, None,
                             ctypes.byref(prm), None) == L.DUST_E_ARG
    assert lib.dust_vort_vel(2, pos.ctypes.data, 3, None, None, None, None, 0, None, None,
                             ctypes.byref(prm), vel.ctypes.data) == L.DUST_E_ARG


def test_pressure_argument_errors_without_gpu():
    """The pressure entry points reject bad sizes and missing arrays before any
    device work."""
    import numpy as np
    lib = L.lib()
    cen = np.zeros(3)
    b = np.zeros(1)
    null = [None] * 5
    assert lib.dust_pres_rhs_rot(-1, None, 0, *null, 0, None, None, None, 0, None, None, None,
                                 None, None) == L.DUST_E_ARG
    # particles announced without their arrays
    assert lib.dust_pres_rhs_rot(1, cen.ctypes.data, 4, *null, 0, None, None, None, 0, None,
                                 None, None, None, b.ctypes.data) == L.DUST_E_ARG
    assert "particle" in lib.dust_last_error().decode()
    # wake panels announced without their arrays
    assert lib.dust_pres_rhs_rot(1, cen.ctypes.data, 0, *null, 0, None, None, None, 2, None,
                                 None, None, None, b.ctypes.data) == L.DUST_E_ARG
    # a wake panel with a vertex count other than 3 or 4
    wn = np.array([4, 5], np.int32)
    wv = np.zeros((2, 4, 3))
    wm = np.zeros(2)
    assert lib.dust_pres_rhs_rot(1, cen.ctypes.data, 0, *null, 0, None, None, None, 2,
                                 wn.ctypes.data, wv.ctypes.data, wm.ctypes.data, wv.ctypes.data,
                                 b.ctypes.data) == L.DUST_E_ARG
    assert "vertices" in lib.dust_last_error().decode()
    A = np.zeros((2, 2), order="F")
    piv = np.zeros(2, np.int32)
    assert lib.dust_solve_pressure_sys(2, 3, A.ctypes.data, 2, piv.ctypes.data, b.ctypes.data,
                                       b.ctypes.data) == L.DUST_E_ARG
    assert lib.dust_pres_rhs_src(2, A.ctypes.data, 1, b.ctypes.data, b.ctypes.data) == L.DUST_E_ARG


def test_assembly_argument_errors_without_gpu():
    """dust_initialize_linsys / dust_assemble_linsys validate the description
    (vertex counts, index ranges) before any device work."""
    import numpy as np
    import sys
    sys.path.insert(0, HERE)
    import linsys_cases as LC
    from dust_amd import linsys as LS
    c = LC.mixed_case(nc=4, ns=2)
    S = LC.mixed_system(c, 0, 0, 0)
    n = S.rank
    A = np.zeros((n, n), order="F")
    b = np.zeros(n)
    lib = L.lib()
    for field, bad in (("impl_ind", n + 5), ("id_sp", -1), ("n_ver", 5)):
        T = LC.mixed_system(c, 0, 0, 0)
        arr = getattr(T, field).copy()
        arr.flat[0] = bad
        setattr(T, field, arr)
        d = T.desc()
        rc = lib.dust_assemble_linsys(ctypes.byref(d), A.ctypes.data, n, A.ctypes.data, n, None,
                                      1, b.ctypes.data, None, 1, None)
        assert rc == L.DUST_E_ARG, field
        assert lib.dust_last_error().decode()
    d = S.desc()
    d.nstatic = n + 1
    assert lib.dust_initialize_linsys(ctypes.byref(d), A.ctypes.data, n, A.ctypes.data, n, None, 1,
                                      None, 1, None, None) == L.DUST_E_ARG


def test_particle_core_argument_errors_without_gpu():
    """An unknown particle core is refused before any device work."""
    import numpy as np

    from dust_amd import octree as O
    lib = L.lib()
    p = O.make_params(4.0, (1, 1, 1), (-2.0, -2.0, -2.0), 3, 32, 4, 0.005)
    p.kernel = 7
    h = ctypes.c_void_p()
    assert lib.dust_fmm_create(ctypes.byref(p), 100, ctypes.byref(h)) == L.DUST_E_ARG
    assert b"kernel" in lib.dust_last_error()
    u = np.zeros(3)
    assert lib.dust_vortpart_direct_k(0, None, None, None, None, 1, 0, 0.0, L.ptr(u), 3, None,
                                      None) == L.DUST_E_ARG


def test_normvel_der_argument_errors_without_gpu():
    import numpy as np
    lib = L.lib()
    n = 2
    z3 = np.zeros((n, 3))
    st = np.zeros((n, 4, 3))
    R = np.tile(np.eye(3), (n, 1, 1))
    out = np.zeros(n)
    args = lambda nv, ng, dt: (n, L.ptr(z3), L.ptr(z3), L.ptr(z3), L.ptr(z3), L.ptr(st),
                               L.ptr(np.asarray(nv, np.int32)), L.ptr(np.asarray(ng, np.int32)),
                               L.ptr(R), dt, L.ptr(out))
    ok_ng = [[1, -1, -1, -2], [0, -1, -1, -1]]
    assert lib.dust_press_normvel_der(*args([4, 5], ok_ng, 0.1)) == L.DUST_E_ARG
    assert lib.dust_press_normvel_der(*args([4, 4], [[2, -1, -1, -1], [0, -1, -1, -1]],
                                            0.1)) == L.DUST_E_ARG
    assert lib.dust_press_normvel_der(*args([4, 4], ok_ng, 0.0)) == L.DUST_E_ARG
